"""Learned initialiser on the device: BF16-MFMA gradients against the CPU
oracle, and what a trained network gains over plain random-restart LM."""
import numpy as np
import pytest

import mlp_oracle as M
from gpu_util import device_tensor, gpu_inversion, oracle_map
from meinv import ME_FP32, ME_INIT_LEARNED
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
