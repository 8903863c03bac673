"""The library's near-field pair code (elfmm_math.h) built for the host:
the pair quadrature against the reference's pair routines
elust3triadirecttarg_one / eltst3triadirecttarg_one (fixture `pairs`), the
unit-density basis used by stored interactions against the pair itself
(linearity in the density), and its operation count (tests/csrc/nf_opcount.cpp)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

CSRC = os.path.join(ROOT, "tests", "csrc")


def build(tmp_path_factory, src):
    out = str(tmp_path_factory.mktemp("nf") / (os.path.splitext(src)[0] + ".so"))
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, os.path.join(CSRC, src)],
                   check=True)
    return C.CDLL(out)


@pytest.fixture(scope="module")
def nf(tmp_path_factory):
    return build(tmp_path_factory, "nf_host.cpp")


@pytest.fixture(scope="module")
def nfops(tmp_path_factory):
    return build(tmp_path_factory, "nf_opcount.cpp")


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)


def sym6(s):
    return np.array([s[0, 0], s[1, 1], s[2, 2], s[0, 1], s[0, 2], s[1, 2]])


def test_pair_matches_reference(nf):
    g = load_golden("pairs")
    for t in range(len(g["ifself"])):
        lam, mu = float(g["lam"][t]), float(g["mu"][t])
        for ifsl, ifdl, kp, ks in ((1, 0, "pt_sl", "st_sl"), (0, 1, "pt_dl", "st_dl")):
            out = np.zeros(9)
            nf.nf_host_pair(P(g["tri"][t]), P(g["sigma"][t]), P(g["sigma"][t]), P(g["target"][t]),
                            int(g["ifself"][t]), C.c_double(lam), C.c_double(mu), ifsl, ifdl, P(out))
            ref = np.concatenate([g[kp][t], sym6(g[ks][t])])
            assert np.abs(out - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max()), (t, ifsl)


def test_basis_is_linear_in_density(nf):
    rng = np.random.default_rng(4)
    for t in range(50):
        tri = rng.uniform(-1, 1, (3, 3))
        x = tri.mean(0) + rng.uniform(-1, 1, 3) * (0.0 if t % 5 == 0 else 0.8)
        iself = 1 if t % 5 == 0 else 0
        ssl, sdl = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        lam, mu = 1.3, 0.7
        b = np.zeros(54)
        nf.nf_host_basis(P(tri), P(x), iself, C.c_double(lam), C.c_double(mu), 1, 1, P(b))
        b = b.reshape(2, 3, 9)
        out = np.zeros(9)
        nf.nf_host_pair(P(tri), P(ssl), P(sdl), P(x), iself, C.c_double(lam), C.c_double(mu), 1, 1, P(out))
        comb = ssl @ b[0] + sdl @ b[1]
        assert np.abs(comb - out).max() <= 1e-12 * np.abs(out).max()


def test_operation_count(nfops):
    """Counted operations of one generic off-surface pair (both layers, all
    outputs): the per-pair work the near-field roofline uses."""
    cnt = np.zeros(4, dtype=np.int64)
    out = np.zeros(9)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0.3, 0.8, 0]], dtype=float)
    nfops.nf_opcount(P(tri), P([0.2, -0.1, 0.3]), P([0.5, 0.1, -0.2]), P([0.4, 0.3, 0.2]), 0,
                     C.c_double(1.0), C.c_double(1.0), 1, 1, cnt.ctypes.data_as(C.c_void_p), P(out))
    adds, muls, divs, trans = (int(v) for v in cnt)
    print("adds", adds, "muls", muls, "divs", divs, "transcendentals", trans)
    assert 500 < adds < 2000 and 800 < muls < 3000 and divs < 100 and trans < 60
