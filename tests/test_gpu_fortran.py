"""Drop-in check from Fortran: tests/fortran/drop_in.f90 calls
elfmm3dtriatarg through an implicit interface with the argument list of
src/elfmm3dtria.f:162-167 and is linked against libelfmm_gfx950.so in place
of the reference's objects. Its outputs must match the reference FMM's on
the same inputs (relative L2 <= 1e-10)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, rel

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "fortran", "_build", "drop_in")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (no flang when build() ran)")
@pytest.mark.parametrize("case", ["fmm_s6", "fmm_c1"])
def test_fortran_caller_matches_reference(tmp_path, case):
    g = load_golden(case)
    n, nt = g["triangle"].shape[0], g["target"].shape[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, nt, int(g["iprec"])], dtype=np.int32).tofile(f)
        np.array([float(g["lam"]), float(g["mu"])]).tofile(f)
        for k in ("triangle", "trinorm", "source", "sigma_sl", "sigma_dl", "target"):
            np.ascontiguousarray(g[k], dtype=np.float64).tofile(f)
    print(f"case {case}: fortran drop_in n={n} nt={nt}", flush=True)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = open(fout, "rb").read()
    ier = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    assert ier == 0
    v = np.frombuffer(raw[4:], dtype=np.float64)
    sizes = [3 * n, 9 * n, 3 * nt, 9 * nt]
    parts = np.split(v, np.cumsum(sizes)[:-1])
    for a, k in zip(parts, ("ptfrc", "strain", "ptfrctarg", "straintarg")):
        assert rel(a, g[k].ravel()) < 1e-10, k
