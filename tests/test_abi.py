"""libmeinv.so loads without a GPU, exports every entry point declared in
include/meinv.h, and the ctypes mirrors have the C layout."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from meinv import abi

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "meinv.h"


def declared_symbols():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(me_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    if not abi.LIB_PATH.exists():
        pytest.skip("libmeinv.so not built (run __graft_entry__.build())")
    lib = C.CDLL(str(abi.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert {n for n, _, _ in abi.ENTRY_POINTS} == set(names)


def test_ctypes_layout_matches_c(tmp_path):
    fields_c = []
    for struct, cls in (("me_config", abi.MeConfig), ("me_stats", abi.MeStats)):
        for name, _ in cls._fields_:
            fields_c.append(f'printf("{struct}.{name} %zu\\n", offsetof({struct}, {name}));')
        fields_c.append(f'printf("{struct} %zu\\n", sizeof({struct}));')
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "meinv.h"\nint main(void){'
                   + "".join(fields_c) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split("\n") if line)
    for struct, cls in (("me_config", abi.MeConfig), ("me_stats", abi.MeStats)):
        for name, _ in cls._fields_:
            assert int(got[f"{struct}.{name}"]) == getattr(cls, name).offset, (struct, name)
        assert int(got[struct]) == C.sizeof(cls), struct


def test_config_defaults_agree_with_the_library():
    if not abi.LIB_PATH.exists():
        pytest.skip("libmeinv.so not built")
    lib = abi.load_library()
    c = abi.MeConfig()
    lib.me_config_defaults(C.byref(c))
    r = abi.reference_defaults()
    for name, _ in abi.MeConfig._fields_:
        a, b = getattr(c, name), getattr(r, name)
        if hasattr(a, "__len__"):
            a = [list(x) if hasattr(x, "__len__") else x for x in a]
            b = [list(x) if hasattr(x, "__len__") else x for x in b]
        if name in ("lambda0", "geff"):
            continue
        assert a == b, name
