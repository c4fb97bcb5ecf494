"""Fixed parameters (`*_fixed = YES`, npar < 9): the oracle packs the free
parameters' rows and columns (Src/LM_FIT.c:316-347), solves the packed
system and steps only the free parameters (Src/LM_FIT.c:226-233)."""
import numpy as np
import pytest

import oracle as O
from meinv.workloads import get
from test_oracle_math import _numpy_solve

FIXED = {"damp_beta": (5, 8), "angles": (1, 2), "four": (0, 4, 6, 7)}


def _fixed_cfg(name, which, **over):
    w = get(name)
    fx = [0] * 9
    for k in FIXED[which]:
        fx[k] = 1
    return w, w.config(fixed=tuple(fx), **over)


@pytest.mark.parametrize("which", list(FIXED))
def test_packed_normal_equations_are_the_free_submatrix(which):
    w, cfg = _fixed_cfg("niris64", which)
    full = O.Oracle(w.config(), w.wavelengths())
    packed = O.Oracle(cfg, w.wavelengths())
    free = [k for k in range(9) if k not in FIXED[which]]
    assert packed.npar == len(free)
    prof, truth = full.gen_map(w.seed, w.ranges(), 6)
    for i in range(6):
        par = truth[i] * 1.02
        step7 = 0.1 * float(prof[i, 0].max())
        norm0 = (0.01 * float(prof[i, 0].astype(np.float64).mean())) ** 2
        Hf, gf, cf = full.normal_eq(par, prof[i].astype(np.float64), step7, norm0)
        Hp, gp, cp = packed.normal_eq(par, prof[i].astype(np.float64), step7, norm0)
        np.testing.assert_array_equal(Hp, Hf[np.ix_(free, free)])
        np.testing.assert_array_equal(gp, gf[free])
        assert cp == cf


@pytest.mark.parametrize("which", list(FIXED))
def test_packed_solve_against_numpy(which):
    w, cfg = _fixed_cfg("niris64", which)
    orc = O.Oracle(cfg, w.wavelengths())
    prof, truth = orc.gen_map(w.seed, w.ranges(), 8)
    for i in range(8):
        par = truth[i] * 1.05
        step7 = 0.1 * float(prof[i, 0].max())
        norm0 = (0.01 * float(prof[i, 0].astype(np.float64).mean())) ** 2
        H, g, _ = orc.normal_eq(par, prof[i].astype(np.float64), step7, norm0)
        x, U = orc.solve(H, g)
        np.testing.assert_allclose(x, _numpy_solve(H, g), rtol=1e-7, atol=1e-9)
        assert np.abs(U.T @ U - np.eye(len(x))).max() < 1e-12


def test_fixed_parameters_keep_their_starting_value():
    """num_run = 1 draws no random jump (runs 1-3 restart from the best with a
    rotated phi), so fixed parameters end where the explicit start put them."""
    w, cfg = _fixed_cfg("niris64", "damp_beta")
    orc = O.Oracle(cfg, w.wavelengths())
    prof, truth = orc.gen_map(w.seed, w.ranges(), 64)
    init = truth.copy()
    init[:, 0] *= 1.3
    init[:, 3] += 0.4
    init[:, 4] *= 0.9
    out = orc.invert(prof.astype(np.float64), init=init, nthreads=4)
    valid = out["res"][:, 4] > 0
    assert valid.all()
    np.testing.assert_array_equal(out["res"][:, 5], init[:, 5])
    np.testing.assert_array_equal(out["res"][:, 8], init[:, 8])


def test_free_parameters_recover_the_truth_when_the_rest_is_fixed_at_it():
    """Fix damping and beta at their true values and start the others off the
    truth: the packed LM must bring B and v_los back (it cannot if the packed
    rows, the packed step or the packed regularisation are misplaced)."""
    w, cfg = _fixed_cfg("niris64", "damp_beta", regul_theta=False)
    orc = O.Oracle(cfg, w.wavelengths())
    prof, truth = orc.gen_map(w.seed, w.ranges(), 64, sigma=1e-4)
    init = truth.copy()
    init[:, 0] = init[:, 0] * 1.2 + 50
    init[:, 3] += 0.3
    out = orc.invert(prof.astype(np.float64), init=init, nthreads=4)
    res = out["res"]
    strong = truth[:, 0] > 500
    conv = res[:, 9] <= 1.0
    assert conv.mean() > 0.8
    ok = conv & strong
    assert np.median(np.abs(res[ok, 0] - truth[ok, 0]) / truth[ok, 0]) < 1e-2
    assert np.median(np.abs(res[conv, 3] - truth[conv, 3])) < 1e-2
