"""The MLP oracle's backward pass against central finite differences (FP64,
no bfloat16 rounding), and its bf16 rounding against hand values."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 65504.0, 1e-3])
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == 1.0078125                       # above the tie -> up
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert r[4] == 65536.0
    assert abs(r[5] - 1e-3) / 1e-3 < 2 ** -8


def test_backward_against_finite_differences():
    nin = 24
    rng = np.random.default_rng(4)
    p = M.init_params(nin, seed=1)
    x = rng.normal(size=(8, nin))
    y = rng.uniform(size=(8, M.OUT))
    loss, g = M.loss_and_grad(p, x, y, bf=False)
    idx = np.concatenate([rng.choice(nin * 256, 20, replace=False),            # W1
                          nin * 256 + rng.choice(256, 5, replace=False),       # b1
                          nin * 256 + 256 + rng.choice(256 * 256, 20, replace=False),  # W2
                          np.arange(len(p) - 10 - 2560, len(p) - 2550),         # W4 first rows
                          np.arange(len(p) - 10, len(p))])                      # b4
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, y, bf=False)[0] - M.loss_and_grad(pm, x, y, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), i


def test_gelu_tanh_form_against_exact_gelu_and_its_derivative():
    z = np.linspace(-12.0, 12.0, 24001)
    # the tanh form stays within 5e-4 of z Phi(z) (largest gap near |z| ~ 2)
    assert np.max(np.abs(M.gelu(z) - M.gelu_erf(z))) < 5e-4
    # known values of the tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
    for v in (-3.0, -1.0, -0.25, 0.5, 1.0, 2.5):
        ref = 0.5 * v * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (v + 0.044715 * v ** 3)))
        assert abs(M.gelu(v) - ref) < 1e-14
    assert abs(M.gelu(1.0) - 0.8411919906082768) < 1e-12
    # saturation: z for large positive z, 0 for large negative z
    assert M.gelu(60.0) == 60.0 and abs(M.gelu(-60.0)) == 0.0
    # derivative against central differences
    h = 1e-5
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(M.gelu_grad(z) - fd)) < 1e-8
