"""Host-side mirror of the reference's inversion entry point.

``Inversion.invert`` plays the role of INVERSION_MULTI (Src/LM_FIT.c:746-811):
profiles in FITS order [n][4][nl] in, records [n][10] out (9 parameters +
chi2/normp, zero records for pixels that are not inverted), optional best-fit
profiles.  Everything runs through libmeinv.so's C ABI; there is no CPU path
here, and a missing library raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import MeConfig, MeStats

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class MeError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} failed: {abi.STATUS_NAMES.get(code, code)}")
        self.code = code


def _check(code: int, where: str):
    if code != abi.ME_OK:
        raise MeError(code, where)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Stats:
    """Aggregates of one call (me_stats) plus optional per-pixel arrays."""

    def __init__(self, st: MeStats, per_pixel: dict | None):
        for name, _ in MeStats._fields_:
            if name in ("iters", "runs", "trials", "status"):
                continue
            val = getattr(st, name)
            if name == "n_fadd":
                val = [val[i] for i in range(8)]
            setattr(self, name, val)
        self.per_pixel = per_pixel or {}

    def as_dict(self):
        d = {k: v for k, v in self.__dict__.items() if k != "per_pixel"}
        return d


class Inversion:
    """One wavelength grid + configuration on the current HIP device."""

    def __init__(self, cfg: MeConfig, wavelengths, lib_path=None):
        self.lib = abi.load_library(lib_path)
        self.cfg = cfg
        self.lam = np.ascontiguousarray(wavelengths, dtype=np.float64)
        self.nl = int(self.lam.size)
        h = C.c_void_p()
        _check(self.lib.me_ctx_create(C.byref(cfg), _ptr(self.lam, _dp), self.nl, C.byref(h)),
               "me_ctx_create")
        self._h = h

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.me_ctx_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_region_counts(self, on: bool):
        """FP32 path: classify Faddeeva calls by FP64 region for the op model
        (n_fadd); off saves a few per cent of kernel time."""
        _check(self.lib.me_ctx_set_region_counts(self._h, int(bool(on))), "me_ctx_set_region_counts")

    def set_pixel_base(self, base: int):
        """Global index of the first pixel of the following batches: the
        Philox counter of the restart jumps is base + index in the batch."""
        _check(self.lib.me_ctx_set_pixel_base(self._h, int(base)), "me_ctx_set_pixel_base")

    def autotune(self, d_prof: int, n: int, prof_is_f64: bool = False):
        """Pick the inversion kernel's launch shape on the first n device
        profiles (me_ctx_autotune); the choice stays in this context."""
        b, wv, ms = C.c_int(), C.c_int(), C.c_double()
        _check(self.lib.me_ctx_autotune(self._h, C.c_void_p(d_prof), int(bool(prof_is_f64)), int(n), C.byref(b),
                                        C.byref(wv), C.byref(ms)), "me_ctx_autotune")
        return dict(block_size=b.value, waves_per_cu=wv.value, kernel_ms=ms.value)

    def info(self):
        ic, i0, i1, nc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        v = np.zeros(2)
        _check(self.lib.me_ctx_info(self._h, C.byref(ic), C.byref(i0), C.byref(i1), C.byref(nc),
                                    _ptr(v, _dp)), "me_ctx_info")
        return dict(ic=ic.value, i0=i0.value, i1=i1.value, ncut=nc.value, vlos=(v[0], v[1]))

    # -- forward model
    def synth(self, par, jac: bool = False):
        par = np.ascontiguousarray(par, dtype=np.float64).reshape(-1, 9)
        n = par.shape[0]
        out = np.empty((n, 4, self.nl))
        J = np.empty((n, 4, 9, self.nl)) if jac else None
        _check(self.lib.me_synth_batch(self._h, _ptr(par, _dp), n, _ptr(out, _dp), _ptr(J, _dp)),
               "me_synth_batch")
        return (out, J) if jac else out

    # -- inversion
    def invert(self, prof, init=None, fit: bool = False, per_pixel: bool = False):
        """prof: [n][4][nl] float32 or float64 host array."""
        prof = np.asarray(prof)
        if prof.dtype not in (np.float32, np.float64):
            prof = prof.astype(np.float64)
        prof = np.ascontiguousarray(prof).reshape(-1, 4, self.nl)
        n = prof.shape[0]
        res = np.empty((n, 10))
        F = np.empty((n, 4, self.nl)) if fit else None
        ini = None if init is None else np.ascontiguousarray(init, dtype=np.float64).reshape(n, 9)
        st = MeStats()
        pp = None
        if per_pixel:
            pp = {k: np.zeros(n, dtype=np.int32) for k in ("iters", "runs", "trials", "status")}
            for k, a in pp.items():
                setattr(st, k, a.ctypes.data_as(C.POINTER(C.c_int32)))
        if prof.dtype == np.float32:
            code = self.lib.me_invert_batch(self._h, _ptr(prof, _fp), n, _ptr(res, _dp), _ptr(F, _dp),
                                            _ptr(ini, _dp), C.byref(st))
        else:
            code = self.lib.me_invert_batch_f64(self._h, _ptr(prof, _dp), n, _ptr(res, _dp),
                                                _ptr(F, _dp), _ptr(ini, _dp), C.byref(st))
        _check(code, "me_invert_batch")
        out = dict(res=res, stats=Stats(st, pp))
        if fit:
            out["fit"] = F
        return out

    def invert_device(self, d_prof: int, n: int, d_res: int, prof_is_f64: bool = False,
                      d_fit: int | None = None, d_init: int | None = None):
        """All buffers already in device memory (integer addresses)."""
        st = MeStats()
        _check(self.lib.me_invert_device(self._h, C.c_void_p(d_prof), int(prof_is_f64), int(n),
                                         C.c_void_p(d_res), C.c_void_p(d_fit) if d_fit else None,
                                         C.c_void_p(d_init) if d_init else None, C.byref(st)),
               "me_invert_device")
        return Stats(st, None)

    def invert_device_noalloc(self, d_prof: int, n: int, d_res: int, prof_is_f64: bool = False,
                              d_init: int | None = None):
        """me_invert_device_noalloc: the same inversion with the library's
        static scratch, allocating nothing (no fit output, no per-pixel
        statistics; learned mode needs d_init)."""
        st = MeStats()
        _check(self.lib.me_invert_device_noalloc(self._h, C.c_void_p(d_prof), int(prof_is_f64), int(n),
                                                 C.c_void_p(d_res), C.c_void_p(d_init) if d_init else None,
                                                 C.byref(st)), "me_invert_device_noalloc")
        return Stats(st, None)

    # -- synthetic maps and the learned initialiser
    def gen_map_device(self, seed: int, ranges, n: int, d_prof: int, d_truth: int | None = None,
                       ff_frac: float = 0.0, ff_bmax: float = 50.0, sigma: float = 1e-3):
        rg = np.ascontiguousarray(ranges, dtype=np.float64).reshape(9, 2)
        _check(self.lib.me_gen_map_device(self._h, C.c_uint64(seed), _ptr(rg, _dp), ff_frac, ff_bmax,
                                          sigma, int(n), C.c_void_p(d_prof),
                                          C.c_void_p(d_truth) if d_truth else None),
               "me_gen_map_device")

    def mlp_train(self, seed, ranges, n_train, batch=4096, epochs=1, lr=1e-3, ff_frac=0.0,
                  ff_bmax=50.0, sigma=1e-3):
        rg = np.ascontiguousarray(ranges, dtype=np.float64).reshape(9, 2)
        loss = C.c_double()
        _check(self.lib.me_mlp_train(self._h, C.c_uint64(seed), _ptr(rg, _dp), ff_frac, ff_bmax, sigma,
                                     int(n_train), int(batch), int(epochs), lr, C.byref(loss)),
               "me_mlp_train")
        return loss.value

    def mlp_predict_device(self, d_prof: int, n: int, d_init: int):
        _check(self.lib.me_mlp_predict_device(self._h, C.c_void_p(d_prof), int(n), C.c_void_p(d_init)),
               "me_mlp_predict_device")

    def mlp_params(self):
        n = self.lib.me_mlp_num_params(self._h)
        p = np.empty(n, dtype=np.float32)
        _check(self.lib.me_mlp_get_params(self._h, _ptr(p, _fp)), "me_mlp_get_params")
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.lib.me_mlp_set_params(self._h, _ptr(p, _fp)), "me_mlp_set_params")

    def mlp_import(self, p, ranges):
        """A network trained elsewhere: exported weights plus the truth ranges
        it was trained on; learned mode is ready afterwards."""
        p = np.ascontiguousarray(p, dtype=np.float32)
        rg = np.ascontiguousarray(ranges, dtype=np.float64).reshape(9, 2)
        _check(self.lib.me_mlp_import(self._h, _ptr(p, _fp), _ptr(rg, _dp)), "me_mlp_import")

    def mlp_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = self.lib.me_mlp_num_params(self._h)
        g = np.empty(n, dtype=np.float32)
        loss = C.c_double()
        _check(self.lib.me_mlp_grad(self._h, _ptr(x, _fp), _ptr(y, _fp), x.shape[0], _ptr(g, _fp),
                                    C.byref(loss)), "me_mlp_grad")
        return g, loss.value
