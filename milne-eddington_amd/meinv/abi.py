"""ctypes mirror of include/meinv.h and the loader of libmeinv.so.

The structures here have the field order of the C declarations; the C-ABI is
the product boundary, this module only binds it for Python callers.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

ME_OK, ME_EINVAL, ME_EDEVICE, ME_ENOMEM, ME_ENOTREADY = 0, 1, 2, 3, 4
ME_FP64, ME_FP32 = 0, 1
ME_INIT_HEURISTIC, ME_INIT_LEARNED = 0, 1
ME_PIX_INVALID, ME_PIX_SVD_FAIL, ME_PIX_FALLBACK = 1, 2, 4

STATUS_NAMES = {0: "ME_OK", 1: "ME_EINVAL", 2: "ME_EDEVICE", 3: "ME_ENOMEM", 4: "ME_ENOTREADY"}

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_PATH = PACKAGE_DIR.parent / "csrc" / "libmeinv.so"


class MeConfig(C.Structure):
    _fields_ = [
        ("lambda0", C.c_double),
        ("geff", C.c_double),
        ("num_iter", C.c_int),
        ("num_run", C.c_int),
        ("convg_criteria", C.c_double),
        ("chisq_criteria", C.c_double),
        ("weights", C.c_double * 4),
        ("damp", C.c_double),
        ("lambda_accept", C.c_double),
        ("lambda_reject", C.c_double),
        ("lambda_limits", C.c_double * 2),
        ("bounds", (C.c_double * 2) * 9),
        ("fixed", C.c_int * 9),
        ("regul", C.c_int * 9),
        ("regul_value", C.c_double * 9),
        ("regul_weight", C.c_double * 9),
        ("vcoeffi", C.c_double),
        ("lcoeffi", C.c_double),
        ("icriteria", C.c_double),
        ("svd_threshold", C.c_double),
        ("hmi_ref", C.c_int),
        ("nfigures", C.c_int),
        ("step", C.c_double * 9),
        ("output_fit", C.c_int),
        ("precision", C.c_int),
        ("init_mode", C.c_int),
        ("seed", C.c_uint64),
        ("block_size", C.c_int),
        ("waves_per_cu", C.c_int),
        ("learned_lambda_limit", C.c_double),
        ("learned_s_limit", C.c_double),
        ("svd_fp32", C.c_int),
    ]


class MeStats(C.Structure):
    _fields_ = [
        ("iters", C.POINTER(C.c_int32)),
        ("runs", C.POINTER(C.c_int32)),
        ("trials", C.POINTER(C.c_int32)),
        ("status", C.POINTER(C.c_int32)),
        ("n_pixels", C.c_int64),
        ("n_valid", C.c_int64),
        ("n_converged", C.c_int64),
        ("n_runs", C.c_int64),
        ("n_iters", C.c_int64),
        ("n_jac", C.c_int64),
        ("n_trial", C.c_int64),
        ("n_hess", C.c_int64),
        ("n_svd", C.c_int64),
        ("n_bk", C.c_int64),
        ("n_fadd", C.c_int64 * 8),
        ("n_svd_fail", C.c_int64),
        ("n_fallback", C.c_int64),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
    ]


# entry points declared in include/meinv.h: (name, restype, argtypes)
_vp, _dp, _fp, _i64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int64
ENTRY_POINTS = [
    ("me_config_defaults", None, [C.POINTER(MeConfig)]),
    ("me_ctx_create", C.c_int, [C.POINTER(MeConfig), _dp, C.c_int, C.POINTER(_vp)]),
    ("me_ctx_destroy", None, [_vp]),
    ("me_ctx_info", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int), _dp]),
    ("me_ctx_set_region_counts", C.c_int, [_vp, C.c_int]),
    ("me_ctx_set_pixel_base", C.c_int, [_vp, _i64]),
    ("me_ctx_autotune", C.c_int, [_vp, _vp, C.c_int, _i64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
    ("me_invert_batch", C.c_int, [_vp, _fp, _i64, _dp, _dp, _dp, C.POINTER(MeStats)]),
    ("me_invert_batch_f64", C.c_int, [_vp, _dp, _i64, _dp, _dp, _dp, C.POINTER(MeStats)]),
    ("me_invert_device", C.c_int, [_vp, _vp, C.c_int, _i64, _vp, _vp, _vp, C.POINTER(MeStats)]),
    ("me_invert_device_noalloc", C.c_int, [_vp, _vp, C.c_int, _i64, _vp, _vp, C.POINTER(MeStats)]),
    ("me_synth_batch", C.c_int, [_vp, _dp, _i64, _dp, _dp]),
    ("me_gen_map_device", C.c_int, [_vp, C.c_uint64, _dp, C.c_double, C.c_double, C.c_double,
                                    _i64, _vp, _vp]),
    ("me_mlp_train", C.c_int, [_vp, C.c_uint64, _dp, C.c_double, C.c_double, C.c_double, _i64,
                               C.c_int, C.c_int, C.c_double, _dp]),
    ("me_mlp_predict_device", C.c_int, [_vp, _vp, _i64, _vp]),
    ("me_mlp_num_params", C.c_int64, [_vp]),
    ("me_mlp_get_params", C.c_int, [_vp, _fp]),
    ("me_mlp_set_params", C.c_int, [_vp, _fp]),
    ("me_mlp_import", C.c_int, [_vp, _fp, _dp]),
    ("me_mlp_grad", C.c_int, [_vp, _fp, _fp, C.c_int, _fp, _dp]),
    ("me_version", C.c_char_p, []),
]

_LIB = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libmeinv.so (built in-tree by build.py).  Raises when it is missing:
    there is no fallback path."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libmeinv.so not built at {p}; run __graft_entry__.build()")
    # torch ships its own libamdhip64.so with the same soname: when both are
    # used in one process, load torch's first so that one HIP runtime serves
    # both (the other order leaves torch without devices)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, res, args in ENTRY_POINTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def reference_defaults() -> MeConfig:
    """Keyword defaults of the reference control file (Src/RINPUT.c:400-458),
    converted as Get_Keys does (Src/RINPUT.c:242-288), computed without the
    library so that CPU-only callers can use them."""
    cfg = MeConfig()
    pi = 3.14159265358979323846
    cfg.lambda0 = 0.0
    cfg.geff = 0.0
    cfg.num_iter = 60
    cfg.num_run = 4
    cfg.convg_criteria = 1e-3
    cfg.chisq_criteria = 1.0
    for i in range(4):
        cfg.weights[i] = 1.0
    cfg.damp = 1e-3
    cfg.lambda_accept = 5.0
    cfg.lambda_reject = 5.0
    cfg.lambda_limits[0], cfg.lambda_limits[1] = 1e-5, 1e3
    bounds = [(5, 6000), (0 * pi, 1 * pi), (0 * pi, 1 * pi), (-20., 20.), (10, 80), (0.4, 0.8),
              (2, 90), (0, 1e6), (0.1, 0.85)]
    for i, (lo, hi) in enumerate(bounds):
        cfg.bounds[i][0], cfg.bounds[i][1] = lo, hi
    for i in range(9):
        cfg.fixed[i] = 0
        cfg.regul[i] = 0
        cfg.regul_value[i] = 0.0
        cfg.regul_weight[i] = 0.0
    # Btheta_Regularization = YES, 0.5, 1e-4 (value x pi, Src/RINPUT.c:279-281)
    cfg.regul[1] = 1
    cfg.regul_value[1] = 0.5 * pi
    cfg.regul_weight[1] = 1e-4
    # "NO, 20, 0.05" style entries parse three fields but stay off; their
    # numbers are still stored (Src/RINPUT.c:269-272)
    cfg.regul_value[4], cfg.regul_weight[4] = 20.0, 0.05
    cfg.regul_value[5], cfg.regul_weight[5] = 0.5, 80.0
    cfg.regul_value[6], cfg.regul_weight[6] = 2.0, 0.01
    cfg.vcoeffi = 100.0
    cfg.lcoeffi = 250.0
    cfg.icriteria = 100.0
    cfg.svd_threshold = 1e-6
    cfg.hmi_ref = 0
    cfg.nfigures = 6
    for i, s in enumerate([1000., 2., 2., 2., 20., 0.3, 30., 1e5, 0.3]):
        cfg.step[i] = s
    cfg.output_fit = 0
    cfg.precision = ME_FP64
    cfg.init_mode = ME_INIT_HEURISTIC
    cfg.seed = 20261200
    cfg.block_size = 0
    cfg.waves_per_cu = 0
    cfg.learned_lambda_limit = 4e-3
    cfg.learned_s_limit = 2.0
    cfg.svd_fp32 = 0
    return cfg
