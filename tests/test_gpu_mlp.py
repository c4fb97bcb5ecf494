"""Learned initialiser on the device: BF16-MFMA gradients against the CPU
oracle, and what a trained network gains over plain random-restart LM."""
import dataclasses

import numpy as np
import pytest

import mlp_oracle as M
from gpu_util import device_tensor, gpu_inversion, oracle_map
from meinv import ME_FP32, ME_INIT_LEARNED, Inversion
from meinv.workloads import get

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_oracle():
    w = get("hmi4k")
    nin = 4 * w.nl
    rng = np.random.default_rng(7)
    p = M.init_params(nin, seed=3).astype(np.float32)
    x = rng.normal(scale=0.5, size=(64, nin)).astype(np.float32)
    y = rng.uniform(size=(64, M.OUT)).astype(np.float32)
    with gpu_inversion("hmi4k") as inv:
        inv.mlp_set_params(p)
        back = inv.mlp_params()
        assert np.array_equal(back, p)
        g, loss = inv.mlp_grad(x, y)
    ref_loss, ref = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64), bf=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    # per layer (W, b): relative error of the gradient block
    dims = [nin, M.H, M.H, M.H, M.OUT]
    o = 0
    for l in range(4):
        for size in (dims[l] * dims[l + 1], dims[l + 1]):
            a, b = g[o:o + size], ref[o:o + size]
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(f"layer {l} block size {size}: relative gradient error {err:.2e}")
            assert err <= 2e-2
            o += size


def test_learned_initialiser_gains_over_plain():
    """Train on device-synthesised profiles, then invert a fresh map both ways."""
    import torch
    w = get("hmi4k")
    n = 20000
    ranges = w.ranges()
    with gpu_inversion("hmi4k", precision=ME_FP32) as inv:
        loss = inv.mlp_train(w.seed + 100, ranges, 262144, batch=2048, epochs=4, lr=2e-3, ff_frac=w.ff_frac,
                             ff_bmax=w.ff_bmax, sigma=w.sigma)
        print(f"final training loss {loss:.4g}")
        assert np.isfinite(loss) and loss < 0.12
        d_prof = device_tensor((n, 4, w.nl), torch.float32)
        d_truth = device_tensor((n, 9), torch.float64)
        inv.gen_map_device(w.seed, ranges, n, d_prof.data_ptr(), d_truth.data_ptr(), ff_frac=w.ff_frac,
                           ff_bmax=w.ff_bmax, sigma=w.sigma)
        d_init = device_tensor((n, 9), torch.float64)
        inv.mlp_predict_device(d_prof.data_ptr(), n, d_init.data_ptr())
        d_res = device_tensor((n, 10), torch.float64)
        plain = inv.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        res_plain = d_res.cpu().numpy()
        pred = d_init.cpu().numpy()
        truth = d_truth.cpu().numpy()
        prof = d_prof.cpu().numpy()
    # the learned mode needs the trained network in the same context: train again there
    with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED) as inv2:
        inv2.mlp_train(w.seed + 100, ranges, 262144, batch=2048, epochs=4, lr=2e-3, ff_frac=w.ff_frac,
                       ff_bmax=w.ff_bmax, sigma=w.sigma)
        d_prof2 = device_tensor((n, 4, w.nl), torch.float32)
        d_prof2.copy_(torch.from_numpy(prof))
        d_res2 = device_tensor((n, 10), torch.float64)
        learned = inv2.invert_device(d_prof2.data_ptr(), n, d_res2.data_ptr())
        res_learned = d_res2.cpu().numpy()
    strong = truth[:, 0] > 300
    err_b = np.median(np.abs(pred[strong, 0] - truth[strong, 0]))
    err_v = np.median(np.abs(pred[:, 3] - truth[:, 3]))
    conv_p = np.mean(res_plain[:, 9] <= w.chisq_criteria)
    conv_l = np.mean(res_learned[:, 9] <= w.chisq_criteria)
    runs_p = plain.n_runs / n
    runs_l = learned.n_runs / n
    print(f"prediction median |dB| {err_b:.1f} G (B > 300 G), |dv| {err_v:.3f} km/s; converged plain {conv_p:.3f} "
          f"in {runs_p:.2f} runs/px, learned {conv_l:.3f} in {runs_l:.2f} runs/px; kernel ms plain "
          f"{plain.kernel_ms:.1f} learned {learned.kernel_ms:.1f}")
    assert err_b < 400 and err_v < 1.0
    assert conv_l >= conv_p + 0.1
    assert runs_l < 0.6 * runs_p


def _numpy_predict(params, prof, nl, ranges):
    """Forward pass of the exported parameters in FP64 with the device's bf16
    rounding points, then the device's decoding into starting points."""
    prof = prof.astype(np.float32)
    imax = np.maximum(prof[:, 0, :].max(axis=1), 0).astype(np.float32)
    inv = np.where(imax > 0, np.float32(1) / np.where(imax > 0, imax, 1), 0).astype(np.float32)
    x = np.concatenate([prof[:, 0, :] * inv[:, None] - np.float32(1),
                        np.float32(10) * prof[:, 1:, :].reshape(len(prof), -1) * inv[:, None]], axis=1)
    x = np.where(np.isfinite(x), x, 0)
    a = M.bf16(x)
    layers = M.unpack(params.astype(np.float64), 4 * nl)
    for l, (W, b) in enumerate(layers):
        z = a @ M.bf16(W) + b
        if l < 3:
            a = M.bf16(M.gelu(z))
    lo, hi = ranges[:, 0], np.where(ranges[:, 1] > ranges[:, 0], ranges[:, 1], ranges[:, 0] + 1)
    out = np.empty((len(prof), 9))
    j = 0
    for k in range(9):
        if k == 2:
            ph = 0.5 * np.arctan2(z[:, j], z[:, j + 1])
            out[:, 2] = np.where(ph < 0, ph + np.pi, ph)
            j += 2
        elif k == 7:
            out[:, 7] = z[:, j] * imax
            j += 1
        else:
            out[:, k] = lo[k] + np.clip(z[:, j], 0, 1) * (hi[k] - lo[k])
            j += 1
    return out, z


def _workload(name):
    """A named workload, or the hmi4k line on an nl-sample grid ("hmi4k-nl12")."""
    if name.startswith("hmi4k-nl"):
        return dataclasses.replace(get("hmi4k"), nl=int(name[len("hmi4k-nl"):]))
    return get(name)


@pytest.mark.parametrize("name", ["hmi4k", "niris64", "hinode", "hmi4k-nl12", "hmi4k-nl20"])
def test_mlp_prediction_matches_numpy_forward(name):
    """Device inference (the fused all-layer kernel for 4*nl <= 256, the
    layer-by-layer GEMMs above) against the numpy forward pass of the same
    exported parameters, on a pixel count that leaves a partial last tile.
    The first layer's k blocks of 32 inputs are 1 (nl 6), 2 (nl 12), 3
    padded with zero weights to 4 (nl 20) and 8 (nl 60)."""
    import torch
    w = _workload(name)
    n = 3001
    ranges = w.ranges()
    with Inversion(w.config(precision=ME_FP32), w.wavelengths()) as inv:
        inv.mlp_train(w.seed + 7, ranges, 16384, batch=1024, epochs=1, lr=2e-3, ff_frac=w.ff_frac,
                      ff_bmax=w.ff_bmax, sigma=w.sigma)
        params = inv.mlp_params()
        d_prof = device_tensor((n, 4, w.nl), torch.float32)
        inv.gen_map_device(w.seed, ranges, n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                           sigma=w.sigma)
        d_init = device_tensor((n, 9), torch.float64)
        d_init.fill_(-1.0)
        inv.mlp_predict_device(d_prof.data_ptr(), n, d_init.data_ptr())
        pred = d_init.cpu().numpy()
        prof = d_prof.cpu().numpy()
    ref, _ = _numpy_predict(params, prof, w.nl, ranges)
    span = ranges[:, 1] - ranges[:, 0]
    for k in (0, 1, 3, 4, 5, 6, 8):
        d = np.abs(pred[:, k] - ref[:, k]) / span[k]
        print(f"{name} param {k}: max normalised difference {d.max():.2e}, 99.9% {np.quantile(d, 0.999):.2e}")
        assert np.quantile(d, 0.999) <= 5e-3 and d.max() <= 3e-2
    dphi = np.abs(pred[:, 2] - ref[:, 2])
    dphi = np.minimum(dphi, np.pi - dphi)
    assert np.quantile(dphi, 0.999) <= 2e-2
    ds = np.abs(pred[:, 7] - ref[:, 7]) / np.maximum(np.abs(ref[:, 7]), 1e-3)
    assert np.quantile(ds, 0.999) <= 5e-3


def test_learned_fp32_error_against_fp64():
    """The learned form's per-parameter error: the same trained network seeds
    the FP32 engine and the FP64 engine (pinned to the oracle by
    test_gpu_parity) on the same map; reported as the fraction of pixels
    within 0.1 G / 0.01 deg / 1 m/s and the median differences."""
    import torch
    from gpu_util import DEG, within_physical
    w = get("hmi4k")
    n = 20000
    ranges = w.ranges()
    with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED) as inv32:
        inv32.mlp_train(w.seed + 100, ranges, 262144, batch=2048, epochs=4, lr=2e-3, ff_frac=w.ff_frac,
                        ff_bmax=w.ff_bmax, sigma=w.sigma)
        params = inv32.mlp_params()
        d_prof = device_tensor((n, 4, w.nl), torch.float32)
        inv32.gen_map_device(w.seed, ranges, n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                             sigma=w.sigma)
        d_res = device_tensor((n, 10), torch.float64)
        st32 = inv32.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        a = d_res.cpu().numpy()
    with gpu_inversion("hmi4k", init_mode=ME_INIT_LEARNED) as inv64:
        inv64.mlp_import(params, ranges)
        st64 = inv64.invert_device(d_prof.data_ptr(), n, d_res.data_ptr())
        b = d_res.cpu().numpy()
    conv_a = np.mean(a[:, 9] <= w.chisq_criteria)
    conv_b = np.mean(b[:, 9] <= w.chisq_criteria)
    both = (a[:, 9] <= w.chisq_criteria) & (b[:, 9] <= w.chisq_criteria)
    phys = within_physical(a[both], b[both]).mean()
    d = np.abs(a[both] - b[both])
    print(f"learned FP32 vs FP64: converged {conv_a:.4f} / {conv_b:.4f}; of pixels converged in both, {phys:.4f} "
          f"within 0.1 G / 0.01 deg / 1 m/s; median |dB| {np.median(d[:, 0]):.3g} G, |dtheta| "
          f"{np.median(d[:, 1]) * DEG:.3g} deg, |dv| {np.median(d[:, 3]) * 1e3:.3g} m/s; fallbacks "
          f"{st32.n_fallback} / {st64.n_fallback}")
    assert abs(conv_a - conv_b) <= 0.02
    assert phys >= 0.6


def test_imported_network_predicts_like_the_trained_one():
    """me_mlp_import: weights exported from one context plus the training
    ranges make another context's predictions (and learned mode) identical."""
    import torch
    w = get("hmi4k")
    n = 4096
    ranges = w.ranges()
    with gpu_inversion("hmi4k", precision=ME_FP32) as a:
        a.mlp_train(w.seed + 3, ranges, 16384, batch=1024, epochs=1, lr=2e-3, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                    sigma=w.sigma)
        params = a.mlp_params()
        d_prof = device_tensor((n, 4, w.nl), torch.float32)
        a.gen_map_device(w.seed, ranges, n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                         sigma=w.sigma)
        d_init = device_tensor((n, 9), torch.float64)
        a.mlp_predict_device(d_prof.data_ptr(), n, d_init.data_ptr())
        pa = d_init.cpu().numpy()
    with gpu_inversion("hmi4k", precision=ME_FP32) as b:
        b.mlp_import(params, ranges)
        d_init.fill_(-1.0)
        b.mlp_predict_device(d_prof.data_ptr(), n, d_init.data_ptr())
        np.testing.assert_array_equal(d_init.cpu().numpy(), pa)


def test_training_is_bit_reproducible():
    """Two trainings with the same seed and settings give the same network
    and report the same final loss: the weight and bias gradients and the
    batch loss are summed in a fixed order (split-K partials added in chunk
    order, column sums and loss partials without atomics)."""
    w = get("hmi4k")
    ranges = w.ranges()
    nets, losses = [], []
    for _ in range(2):
        with gpu_inversion("hmi4k", precision=ME_FP32) as inv:
            losses.append(inv.mlp_train(w.seed + 11, ranges, 65536, batch=4096, epochs=2, lr=2e-3,
                                        ff_frac=w.ff_frac, ff_bmax=w.ff_bmax, sigma=w.sigma))
            nets.append(inv.mlp_params())
    assert np.all(np.isfinite(nets[0]))
    np.testing.assert_array_equal(nets[1], nets[0])
    assert np.isfinite(losses[0]) and losses[0] > 0
    assert losses[1] == losses[0]
