"""The gather table behind the L2T/M2T field operand (elfmm_ops.cpp
build_field_table): the nine fields it forms from a packed real expansion
(value, gradient, five Hessian entries) are checked against central finite
differences of the value field, with the solid harmonics taken from
scipy's associated Legendre functions, for local (regular) and multipole
(irregular) expansions. The reference evaluates the same fields in
lfmm3dtrilhess's L2T/M2T steps (src/lfmm3dtrilhess.f:1120-1200,
l3dtaevalhessdini / l3dmpevalhessd)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.special import lpmv

from conftest import ROOT

CSRC = os.path.join(ROOT, "elfmmlib3d_amd", "csrc")


@pytest.fixture(scope="module")
def ft_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ft") / "libfield_table.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out,
                    os.path.join(ROOT, "tests", "csrc", "field_table.cpp"),
                    os.path.join(CSRC, "elfmm_ops.cpp")], check=True)
    return C.CDLL(out)


def table(lib, p, irr):
    nmax = 2 * 9 * ((p + 3) * (p + 4) // 2 + 1)
    idx = np.zeros((nmax, 3), dtype=np.int32)
    c = np.zeros((nmax, 3))
    n = lib.field_table(p, int(irr), idx.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p), nmax)
    assert n > 0
    return idx[:n].reshape(-1, 2, 9, 3), c[:n].reshape(-1, 2, 9, 3)


def harmonics(ph, x, irr):
    """V_l^m = (l-m)! r^l P_l^m(cos t) e^{i m phi} (regular) or
    (l-m)! P_l^m(cos t) e^{i m phi} / r^(l+1) (irregular), Condon-Shortley
    phase, in stream order (m outer, l inner)."""
    r = np.linalg.norm(x)
    ct, phi = x[2] / r, math.atan2(x[1], x[0])
    out = []
    for m in range(ph + 1):
        for l in range(m, ph + 1):
            rad = r ** l if not irr else r ** (-l - 1)
            out.append(math.factorial(l - m) * lpmv(m, l, ct) * rad * np.exp(1j * m * phi))
    return np.array(out)


def fields(idx, c, E, ph, x, irr):
    v = harmonics(ph, x, irr)
    npos = idx.shape[0]
    h = np.zeros(npos, dtype=complex)
    h[:len(v)] = v
    g = (c * E[idx]).sum(axis=-1)                  # (pos, part, q)
    return g[:, 0, :].T @ h.real + g[:, 1, :].T @ h.imag


def random_expansion(p, rng):
    E = np.zeros((p + 1) * (p + 2))
    for l in range(p + 1):
        for m in range(l + 1):
            j = 2 * (l * (l + 1) // 2 + m)
            E[j] = rng.standard_normal() / (l + 1)
            E[j + 1] = rng.standard_normal() / (l + 1) if m else 0.0
    return E


@pytest.mark.parametrize("irr", [False, True])
@pytest.mark.parametrize("p", [1, 2, 7])
def test_fields_match_finite_differences(ft_lib, p, irr):
    rng = np.random.default_rng(10 * p + irr)
    idx, c = table(ft_lib, p, irr)
    ph = p + 2 if irr else p
    E = random_expansion(p, rng)
    x0 = np.array([1.3, -0.8, 1.1]) if irr else np.array([0.31, -0.22, 0.27])
    f = fields(idx, c, E, ph, x0, irr)
    val = lambda x: fields(idx, c, E, ph, x, irr)[0]
    I = np.eye(3)
    h1, h2 = 1e-5, 1e-3
    grad = [(val(x0 + h1 * I[a]) - val(x0 - h1 * I[a])) / (2 * h1) for a in range(3)]

    def d2(a, b):
        ea, eb = h2 * I[a], h2 * I[b]
        return (val(x0 + ea + eb) - val(x0 + ea - eb) - val(x0 - ea + eb) + val(x0 - ea - eb)) / (4 * h2 * h2)

    hess = {4: d2(0, 0), 5: d2(1, 1), 6: d2(0, 1), 7: d2(0, 2), 8: d2(1, 2)}
    scale = max(1.0, np.abs(f).max())
    assert np.allclose(f[1:4], grad, atol=1e-7 * scale, rtol=1e-7)
    for q, ref in hess.items():
        assert abs(f[q] - ref) < 1e-5 * scale, (q, f[q], ref)
    # harmonic field: the trace of the Hessian vanishes (zz = -(xx + yy))
    assert abs(f[4] + f[5] + d2(2, 2)) < 1e-5 * scale


def test_table_builds_for_all_orders(ft_lib):
    for p in range(1, 45):
        for irr in (False, True):
            idx, c = table(ft_lib, p, irr)
            ph = p + 2 if irr else p
            assert idx.shape[0] == ((ph + 1) * (ph + 2) // 2 + 1) // 2 * 2
            assert idx.max() < (p + 1) * (p + 2)
