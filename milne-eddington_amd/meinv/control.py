"""MEINV control-file keywords -> me_config.

Mirrors the reference's parser for the keywords the inversion loop consumes
(Src/RINPUT.c:361-563 RDINPUT, Src/RINPUT.c:48-357 Get_Keys): ``key = value``
lines, comments introduced by ``#``, ``!`` or ``*`` (Src/STR.c:441-452), the
default table of Src/RINPUT.c:400-458 and the unit conversions of Get_Keys
(theta/phi bounds and regularisation constants times pi).  Errors that the
reference treats as fatal raise ``ControlFileError``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

from .abi import MeConfig, reference_defaults

PI = 3.14159265358979323846


class ControlFileError(ValueError):
    pass


# keyword -> default text, as in the reference's keyword table
DEFAULTS = {
    "verbose": "1", "num_iter": "60", "num_run": "4", "convg_criteria": "1e-3",
    "chisq_criteria": "1.0", "weights": "1., 1., 1., 1.", "sol_box": "-1, -1, -1, -1",
    "output_fit": "NO", "cache_path": "./cache", "result_path": "./res.fits",
    "fit_path": "./fit.fits", "damp": "1e-3", "lambda_accept": "5.0", "lambda_reject": "5.0",
    "lambda_limits": "1e-5, 1e3", "Bounds_Bmod": "5, 6000", "Bounds_Btheta": "0, 1",
    "Bounds_Bphi": "0, 1", "Bounds_Vlos": "-20., 20.", "Bounds_Dopp": "10, 80",
    "Bounds_Damp": "0.4, 0.8", "Bounds_Eta": "2, 90", "Bounds_Continuum": "0, 1e6",
    "Bounds_Beta": "0.1, 0.85", "Bmod_fixed": "NO", "Btheta_fixed": "NO", "Bphi_fixed": "NO",
    "Vlos_fixed": "NO", "Dopp_fixed": "NO", "Damp_fixed": "NO", "Eta_fixed": "NO",
    "Continuum_fixed": "NO", "Beta_fixed": "NO", "Bmod_Regularization": "NO",
    "Btheta_Regularization": "YES, 0.5, 1e-4", "Bphi_Regularization": "NO",
    "Vlos_Regularization": "NO", "Dopp_Regularization": "NO, 20, 0.05",
    "Damp_Regularization": "NO, 0.5, 80", "Eta_Regularization": "NO, 2.0, 0.01",
    "Continuum_Regularization": "NO", "Beta_Regularization": "NO", "VCoeffi": "100.",
    "LCoeffi": "250.", "Icriteria": "100.", "SVDthreshold": "1e-6", "HMI_REF": "No",
    "Cache_prof": "No", "Nthreads": "4", "Nfigures": "6",
    "Step": "1000., 2., 2., 2., 20., 0.3, 30., 1e5, 0.3", "nprofiles": "20", "Cache_inv": "No",
}
REQUIRED = ("lines", "data_path", "wavelength_path")
BOUNDS = ["Bounds_Bmod", "Bounds_Btheta", "Bounds_Bphi", "Bounds_Vlos", "Bounds_Dopp",
          "Bounds_Damp", "Bounds_Eta", "Bounds_Continuum", "Bounds_Beta"]
FIXED = ["Bmod_fixed", "Btheta_fixed", "Bphi_fixed", "Vlos_fixed", "Dopp_fixed", "Damp_fixed",
         "Eta_fixed", "Continuum_fixed", "Beta_fixed"]
REGUL = ["Bmod_Regularization", "Btheta_Regularization", "Bphi_Regularization",
         "Vlos_Regularization", "Dopp_Regularization", "Damp_Regularization",
         "Eta_Regularization", "Continuum_Regularization", "Beta_Regularization"]


@dataclass
class ControlFile:
    values: dict = field(default_factory=dict)   # keyword -> raw text (after defaults)
    lines: list = field(default_factory=list)    # raw `lines` entries
    warnings: list = field(default_factory=list)


def _strip_comment(line: str) -> str:
    # a comment runs from the first '#', '!' or '*' to the end of the line
    cut = len(line)
    for ch in "#!*":
        i = line.find(ch)
        if i >= 0:
            cut = min(cut, i)
    return line[:cut]


def parse_text(text: str) -> ControlFile:
    cf = ControlFile(values=dict(DEFAULTS))
    seen = set()
    for raw in text.splitlines():
        line = _strip_comment(raw)
        if "=" not in line:
            continue
        key, value = line.split("=", 1)
        key = key.strip()
        value = value.strip()
        if not key:
            continue
        if key == "lines":
            cf.lines.append(value)
            seen.add(key)
            continue
        if key in DEFAULTS or key in REQUIRED:
            if key in seen:
                cf.warnings.append(f"keyword {key} redefined; overwriting")
            cf.values[key] = value
            seen.add(key)
        else:
            cf.warnings.append(f"unknown keyword: {key}")
    missing = [k for k in REQUIRED if k not in seen]
    if missing:
        raise ControlFileError(f"required keyword missing: {', '.join(missing)}")
    return cf


def parse_file(path) -> ControlFile:
    p = Path(path)
    if not p.exists():
        raise ControlFileError(f"control file doesn't exist: {p}")
    return parse_text(p.read_text())


def _floats(text: str, n: int, key: str) -> list[float]:
    parts = [t for t in text.replace(",", " ").split() if t]
    try:
        vals = [float(t) for t in parts[:n]]
    except ValueError as e:
        raise ControlFileError(f"error reading {n} numbers for key {key}") from e
    if len(vals) != n:
        raise ControlFileError(f"error reading {n} numbers for key {key}")
    return vals


def _yes(text: str) -> bool:
    return text.strip().upper() == "YES"


def _atoi(text: str) -> int:
    # atoi semantics: leading integer, 0 when none
    t = text.strip()
    num = ""
    for i, ch in enumerate(t):
        if ch.isdigit() or (i == 0 and ch in "+-"):
            num += ch
        else:
            break
    try:
        return int(num)
    except ValueError:
        return 0


def _atof(text: str) -> float:
    t = text.strip().split()[0] if text.strip() else "0"
    try:
        return float(t.rstrip(","))
    except ValueError:
        return 0.0


def to_config(cf: ControlFile, line_index: int = 0) -> MeConfig:
    """Convert parsed keywords to the engine configuration (internal units)."""
    cfg = reference_defaults()
    v = cf.values
    if not cf.lines:
        raise ControlFileError("Error in reading the spectral line info.")
    lam0, geff = _floats(cf.lines[line_index], 2, "lines")
    cfg.lambda0, cfg.geff = lam0, geff
    cfg.num_iter = _atoi(v["num_iter"])
    cfg.num_run = _atoi(v["num_run"])
    cfg.convg_criteria = _atof(v["convg_criteria"])
    cfg.chisq_criteria = _atof(v["chisq_criteria"])
    for i, w in enumerate(_floats(v["weights"], 4, "weights")):
        cfg.weights[i] = w
    cfg.output_fit = int(_yes(v["output_fit"]))
    cfg.damp = _atof(v["damp"])
    cfg.lambda_accept = _atof(v["lambda_accept"])
    cfg.lambda_reject = _atof(v["lambda_reject"])
    if cfg.lambda_accept < 1.5 or cfg.lambda_reject < 1.5:
        raise ControlFileError("Lam_accept and Lam_reject should be larger than 1.5.")
    lo, hi = _floats(v["lambda_limits"], 2, "lambda_limits")
    cfg.lambda_limits[0], cfg.lambda_limits[1] = lo, hi
    for i, key in enumerate(BOUNDS):
        lo, hi = _floats(v[key], 2, key)
        if i in (1, 2):
            lo, hi = lo * PI, hi * PI
        cfg.bounds[i][0], cfg.bounds[i][1] = lo, hi
    for i in range(9):
        cfg.fixed[i] = int(_yes(v[FIXED[i]]))
        parts = [t.strip() for t in v[REGUL[i]].split(",")]
        cfg.regul[i] = 0
        cfg.regul_value[i] = 0.0
        cfg.regul_weight[i] = 0.0
        if len(parts) >= 3:
            try:
                val, wt = float(parts[1]), float(parts[2])
            except ValueError:
                continue
            cfg.regul_value[i], cfg.regul_weight[i] = val, wt
            if parts[0].upper() == "YES":
                cfg.regul[i] = 1
                if i in (1, 2):
                    cfg.regul_value[i] = val * PI
    cfg.vcoeffi = _atof(v["VCoeffi"])
    cfg.lcoeffi = _atof(v["LCoeffi"])
    cfg.icriteria = _atof(v["Icriteria"])
    cfg.svd_threshold = _atof(v["SVDthreshold"])
    cfg.hmi_ref = int(_yes(v["HMI_REF"]))
    if _yes(v["Cache_prof"]):
        cfg.output_fit = 0
    cfg.nfigures = _atoi(v["Nfigures"])
    for i, s in enumerate(_floats(v["Step"], 9, "Step")):
        cfg.step[i] = s
    if not math.isfinite(cfg.lambda0):
        raise ControlFileError("bad line centre")
    return cfg
