"""Shared helpers of the GPU tests: device buffers come from torch (plumbing
only); every computation under test goes through libmeinv.so's C ABI."""
import numpy as np

try:  # one HIP runtime per process: torch's must load before libmeinv's
    import torch  # noqa: F401
except ImportError:
    pass

import oracle as O
from meinv import Inversion
from meinv.workloads import get

DEG = 180.0 / np.pi


def device_tensor(shape, dtype):
    import torch
    return torch.empty(shape, dtype=dtype, device="cuda:0")


def oracle_map(name, n, seed=None):
    w = get(name)
    orc = O.Oracle(w.config(), w.wavelengths())
    prof, truth = orc.gen_map(w.seed if seed is None else seed, w.ranges(), n, ff_frac=w.ff_frac,
                              ff_bmax=w.ff_bmax, sigma=w.sigma)
    return w, orc, prof, truth


def within_parity(a, b):
    """Per-pixel agreement at the stated bar: 0.1 G, 0.01 deg, 1 m/s, and
    1e-6 relative on the other five parameters and chi2."""
    ok = np.abs(a[:, 0] - b[:, 0]) <= 0.1
    ok &= np.abs(a[:, 1] - b[:, 1]) * DEG <= 0.01
    dphi = np.abs(a[:, 2] - b[:, 2])
    dphi = np.minimum(dphi, np.pi - dphi)
    ok &= dphi * DEG <= 0.01
    ok &= np.abs(a[:, 3] - b[:, 3]) <= 1e-3
    for k in (4, 5, 6, 7, 8, 9):
        ok &= np.abs(a[:, k] - b[:, k]) <= 1e-6 * np.maximum(1.0, np.abs(b[:, k]))
    return ok


def gpu_inversion(name, **over):
    w = get(name)
    return Inversion(w.config(**over), w.wavelengths())


def within_physical(a, b, dB=0.1, ddeg=0.01, dv=1e-3):
    """Agreement on the four quantities of the stated bar only."""
    ok = np.abs(a[:, 0] - b[:, 0]) <= dB
    ok &= np.abs(a[:, 1] - b[:, 1]) * DEG <= ddeg
    dphi = np.abs(a[:, 2] - b[:, 2])
    dphi = np.minimum(dphi, np.pi - dphi)
    ok &= dphi * DEG <= ddeg
    ok &= np.abs(a[:, 3] - b[:, 3]) <= dv
    return ok


def chaos_floor(orc, prof64, init=None, rel=1e-15, seed=0):
    """Agreement of the oracle with itself when the input moves by ~1 ulp:
    later LM trials solve with the left factor of an orthogonal matrix's SVD
    (H1), which rounding alone decides, so no second implementation can agree
    more often than this."""
    a = orc.invert(prof64, init=init, nthreads=8)
    rng = np.random.default_rng(seed)
    b = orc.invert(prof64 * (1 + rel * rng.standard_normal(prof64.shape)), init=init, nthreads=8)
    return within_parity(b["res"], a["res"]).mean(), a
