"""me_invert_batch from host buffers: the profiles go to the device and the
records come back through pinned staging blocks moved by several host threads
(8 MiB blocks, here hundreds of them); the records, per-pixel statistics,
best-fit profiles and counters equal me_invert_device on the same device
profiles bit for bit."""
import numpy as np
import pytest

from gpu_util import device_tensor, gpu_inversion
from meinv import ME_FP32, ME_INIT_LEARNED
from meinv.workloads import get

pytestmark = pytest.mark.gpu

N = (1 << 21) * 2 + 12345     # ~400 MB of profiles, the last staging block short


def _map(inv, w, n):
    import torch
    d_prof = device_tensor((n, 4, w.nl), torch.float32)
    inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                       sigma=w.sigma)
    torch.cuda.synchronize()
    return d_prof


def _same_counters(a, b):
    for k in ("n_valid", "n_converged", "n_runs", "n_iters", "n_jac", "n_trial", "n_hess", "n_svd", "n_bk",
              "n_svd_fail", "n_fallback"):
        assert getattr(a, k) == getattr(b, k), k


def test_host_batch_equals_the_device_launch():
    import torch
    w = get("hmi4k")
    with gpu_inversion("hmi4k", precision=ME_FP32, num_run=2) as inv:
        d_prof = _map(inv, w, N)
        d_res = device_tensor((N, 10), torch.float64)
        ref = inv.invert_device(d_prof.data_ptr(), N, d_res.data_ptr())
        want = d_res.cpu().numpy()
        prof = d_prof.cpu().numpy()
        del d_res
        out = inv.invert(prof, per_pixel=True, fit=True)
    st = out["stats"]
    print(f"{N} px from host buffers: total {st.total_ms:.1f} ms, kernel {st.kernel_ms:.1f} ms "
          f"(device-resident launch {ref.kernel_ms:.1f} ms); converged {st.n_converged}/{st.n_valid}")
    np.testing.assert_array_equal(out["res"], want)
    _same_counters(st, ref)
    # restarts (num_run = 2: at least 5 runs per pixel) keyed by the batch index
    assert st.n_runs > 2 * st.n_valid
    pp = st.per_pixel
    assert pp["runs"].sum() == st.n_runs and pp["trials"].sum() == st.n_trial
    assert np.all((pp["status"] & 1) == (out["res"][:, 9] == 0))
    # best-fit profiles of a sample of pixels across staging-block boundaries
    idx = np.r_[0:50, (1 << 21) - 50:(1 << 21) + 50, N - 50:N]
    syn = inv_synth(w, out["res"][idx, :9])
    ok = out["res"][idx, 9] > 0
    np.testing.assert_allclose(out["fit"][idx][ok], syn[ok], rtol=0, atol=1e-5)


def inv_synth(w, par):
    with gpu_inversion("hmi4k", precision=ME_FP32) as inv:
        return inv.synth(par)


def test_host_batch_learned_mode_equals_the_device_launch():
    import torch
    w = get("hmi4k")
    with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED) as inv:
        inv.mlp_train(w.seed + 7, w.ranges(), 16384, batch=1024, epochs=1, lr=2e-3, ff_frac=w.ff_frac,
                      ff_bmax=w.ff_bmax, sigma=w.sigma)
        d_prof = _map(inv, w, N)
        d_res = device_tensor((N, 10), torch.float64)
        ref = inv.invert_device(d_prof.data_ptr(), N, d_res.data_ptr())
        want = d_res.cpu().numpy()
        prof = d_prof.cpu().numpy()
        del d_res
        out = inv.invert(prof)
    np.testing.assert_array_equal(out["res"], want)
    _same_counters(out["stats"], ref)
    assert ref.n_fallback > 0
