"""Shared helpers of the GPU tests: device buffers come from torch (plumbing
only); every computation under test goes through libmeinv.so's C ABI."""
import numpy as np

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
