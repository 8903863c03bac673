"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name not in ("if", "defined", "sizeof"):
                names.add(name)
    return sorted(n for n in names if n.startswith(("elfmm", "el3d")))


def test_exports_every_declared_symbol(gpu_lib):
    lib = ctypes.CDLL(gpu_lib.library_path())
    syms = declared_symbols()
    assert len(syms) >= 6
    for s in syms:
        assert hasattr(lib, s), s


def test_version_string(gpu_lib):
    assert b"gfx950" in gpu_lib.library().elfmm_gfx950_version()
