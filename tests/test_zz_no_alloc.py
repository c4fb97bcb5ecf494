"""me_invert_device_noalloc: the allocation-free variant of me_invert_device
(static scratch array, no hipMalloc), through the C ABI, against
me_invert_device on the same device inputs: equal records and counters on two
calls in a row, the second with a smaller input, and no change in the
device's free memory across a call."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import device_tensor, gpu_inversion
from meinv import ME_FP32, ME_INIT_LEARNED, MeError
from meinv.abi import ME_EINVAL, MeStats
from meinv.workloads import get

pytestmark = pytest.mark.gpu


def _map(inv, w, n):
    import torch
    d_prof = device_tensor((n, 4, w.nl), torch.float32)
    inv.gen_map_device(w.seed, w.ranges(), n, d_prof.data_ptr(), None, ff_frac=w.ff_frac, ff_bmax=w.ff_bmax,
                       sigma=w.sigma)
    return d_prof


def _same_counters(a, b):
    for k in ("n_valid", "n_converged", "n_runs", "n_iters", "n_jac", "n_trial", "n_hess", "n_svd", "n_bk",
              "n_svd_fail", "n_fallback"):
        assert getattr(a, k) == getattr(b, k), k


# svd_fp32 = 1: the two-wave kernel build, whose default launch (two 256-lane
# blocks per CU) exceeds the static scratch's lanes; the allocation-free call
# caps it at one block per CU, with the same records
@pytest.mark.parametrize("name,prec,svd32", [("hmi4k", ME_FP32, 0), ("hmi4k", ME_FP32, 1), ("niris64", ME_FP32, 0),
                                             ("niris64", 0, 0)])
def test_noalloc_equals_the_allocating_entry_point(name, prec, svd32):
    import torch
    w = get(name)
    with gpu_inversion(name, precision=prec, svd_fp32=svd32) as inv:
        d_prof = _map(inv, w, 20000)
        d_a = device_tensor((20000, 10), torch.float64)
        d_b = device_tensor((20000, 10), torch.float64)
        for n in (20000, 5000):  # two calls in a row, the second smaller
            d_a.fill_(-1.0)
            d_b.fill_(-2.0)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            sb = inv.invert_device_noalloc(d_prof.data_ptr(), n, d_b.data_ptr())
            free1 = torch.cuda.mem_get_info()[0]
            sa = inv.invert_device(d_prof.data_ptr(), n, d_a.data_ptr())
            a, b = d_a.cpu().numpy(), d_b.cpu().numpy()
            print(f"{name} prec={prec} svd_fp32={svd32} n={n}: converged {sb.n_converged}/{sb.n_valid}; kernel {sb.kernel_ms:.2f} ms "
                  f"(allocating path {sa.kernel_ms:.2f}); free memory change {free1 - free0} B")
            np.testing.assert_array_equal(b[:n], a[:n])
            assert np.all(b[n:] == -2.0)
            _same_counters(sa, sb)
            assert sb.n_valid > 0.9 * n
            assert free1 == free0


def test_noalloc_learned_mode_with_given_predictions():
    import torch
    w = get("hmi4k")
    n = 8192
    with gpu_inversion("hmi4k", precision=ME_FP32, init_mode=ME_INIT_LEARNED) as inv:
        inv.mlp_train(w.seed + 5, w.ranges(), 16384, batch=1024, epochs=1, lr=2e-3, ff_frac=w.ff_frac,
                      ff_bmax=w.ff_bmax, sigma=w.sigma)
        d_prof = _map(inv, w, n)
        d_init = device_tensor((n, 9), torch.float64)
        inv.mlp_predict_device(d_prof.data_ptr(), n, d_init.data_ptr())
        d_a = device_tensor((n, 10), torch.float64)
        d_b = device_tensor((n, 10), torch.float64)
        sa = inv.invert_device(d_prof.data_ptr(), n, d_a.data_ptr(), d_init=d_init.data_ptr())
        sb = inv.invert_device_noalloc(d_prof.data_ptr(), n, d_b.data_ptr(), d_init=d_init.data_ptr())
        np.testing.assert_array_equal(d_b.cpu().numpy(), d_a.cpu().numpy())
        _same_counters(sa, sb)
        assert sb.n_fallback > 0
        # learned mode cannot predict without a buffer: the caller's d_init is required
        with pytest.raises(MeError) as e:
            inv.invert_device_noalloc(d_prof.data_ptr(), n, d_b.data_ptr())
        assert e.value.code == ME_EINVAL


def test_noalloc_rejects_per_pixel_statistics():
    import torch
    w = get("niris64")
    with gpu_inversion("niris64") as inv:
        d_prof = _map(inv, w, 64)
        d_res = device_tensor((64, 10), torch.float64)
        runs = np.zeros(64, dtype=np.int32)
        st = MeStats()
        st.runs = runs.ctypes.data_as(C.POINTER(C.c_int32))
        code = inv.lib.me_invert_device_noalloc(inv.handle, C.c_void_p(d_prof.data_ptr()), 0, 64,
                                                C.c_void_p(d_res.data_ptr()), None, C.byref(st))
        assert code == ME_EINVAL
