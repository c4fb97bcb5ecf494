"""The learned initialiser's network file, shared with the host program's
--save-mlp / --mlp: magic "MEINVMLP", int64 parameter count, the truth ranges
[9][2] (float64) the network was trained on (they decode its outputs), then the
FP32 weights in me_mlp_get_params order; little-endian."""
from __future__ import annotations

from pathlib import Path

import numpy as np

MAGIC = b"MEINVMLP"


def save_network(path, params, ranges) -> None:
    p = np.ascontiguousarray(params, dtype="<f4").ravel()
    r = np.ascontiguousarray(ranges, dtype="<f8").reshape(18)
    Path(path).write_bytes(MAGIC + np.int64(p.size).astype("<i8").tobytes() + r.tobytes() + p.tobytes())


def load_network(path):
    raw = Path(path).read_bytes()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a network file")
    n = int(np.frombuffer(raw[8:16], dtype="<i8")[0])
    if len(raw) != 16 + 144 + 4 * n:
        raise ValueError(f"{path}: size does not match its parameter count")
    ranges = np.frombuffer(raw[16:160], dtype="<f8").reshape(9, 2).copy()
    params = np.frombuffer(raw[160:], dtype="<f4").copy()
    return params, ranges
