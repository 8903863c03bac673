"""The Mindlin far-field operations of the library (elfmm_mindlin.h, the
code the Mindlin far-field kernels run), built for the host, against the
reference's own routines compiled from its sources
(oracle/_ref/libelfmm_refh.so), on random expansions in the reference's
(0:p, -p:p) layout:

- plane-wave tables: lwtsexp3 + numthetahalf / numthetafour (+2 modes);
- multipole -> physical plane waves: mpoletoexp + ftophys;
- the B scaling of the plane waves: mindlinbpwdiagall;
- the plane-wave shift: xyzshift_add;
- physical plane waves -> local: phystof + exptolocald_add;
- B's proxy charges: mptoslpborder;
- B's remainder: l3dformmp of the proxies, mpdiff, mpdzminus2;
- the proxies' local expansion: l3dformtaintkerbf90sc."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "tests", "csrc")


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    from oracle import ref
    if not ref.available_h():
        pytest.skip("reference half-space library not built (make -C oracle refh)")
    out = str(tmp_path_factory.mktemp("mind") / "mindlin_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, os.path.join(CSRC, "mindlin_host.cpp")],
                   check=True)
    return C.CDLL(out), ref.libh()


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def I(v):
    return C.byref(C.c_int(int(v)))


def D(v):
    return C.byref(C.c_double(float(v)))


ORDER = {1: 14, 2: 26, 3: 39}


class Ex:
    """Packed (l, m >= 0) complex coefficients <-> the reference's full
    (0:p, -p:p) array, a_{l,-m} = conj(a_{l,m})."""

    def __init__(self, p):
        self.p = p

    def rand(self, rng, decay):
        p = self.p
        out = np.zeros((p + 1) * (p + 2))
        for l in range(p + 1):
            for m in range(l + 1):
                k = 2 * (l * (l + 1) // 2 + m)
                out[k] = rng.normal() * decay ** l
                out[k + 1] = 0.0 if m == 0 else rng.normal() * decay ** l
        return out

    def full(self, pk):
        p = self.p
        F = np.zeros((p + 1, 2 * p + 1), complex, order="F")
        for l in range(p + 1):
            for m in range(l + 1):
                v = complex(pk[2 * (l * (l + 1) // 2 + m)], pk[2 * (l * (l + 1) // 2 + m) + 1])
                F[l, m + p] = v
                if m:
                    F[l, p - m] = np.conj(v)
        return F

    def packed(self, F):
        p = self.p
        out = np.zeros((p + 1) * (p + 2))
        for l in range(p + 1):
            for m in range(l + 1):
                out[2 * (l * (l + 1) // 2 + m)] = F[l, m + p].real
                out[2 * (l * (l + 1) // 2 + m) + 1] = F[l, m + p].imag
        return out


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def ref_tables(ref, iprec, p):
    xs, ws, nl = np.zeros(100), np.zeros(100), C.c_int(0)
    ref.lwtsexp3_(I(iprec), P(xs), P(ws), C.byref(nl))
    nl = nl.value
    nf, nph = np.zeros(nl, np.int32), np.zeros(nl, np.int32)
    ref.numthetahalf_(P(nf), I(nl))
    ref.numthetafour_(P(nph), I(nl))
    nf += 2
    nph += 2
    ne, nb = C.c_int(0), C.c_int(0)
    ref.getfexplengths_(I(nl), P(nf), P(nph), C.byref(ne), C.byref(nb))
    fexpe, fexpb = np.zeros(ne.value, complex), np.zeros(nb.value, complex)
    ref.mkfexp_(I(nl), P(nf), P(nph), P(fexpe), I(ne.value), P(fexpb), I(nb.value))
    rlsc = np.zeros(nl * (p + 1) * (p + 1))
    ref.rlscini_(P(rlsc), I(nl), P(xs), I(p))
    return dict(xs=xs, ws=ws, nl=nl, nf=nf, nph=nph, nexptot=int(nf.sum()), nexptotp=int(nph.sum()),
                nthmax=int(nf.max()), nthmaxp=int(nph.max()), fexpe=fexpe, ne=ne.value, fexpb=fexpb, nb=nb.value,
                rlsc=rlsc)


@pytest.mark.parametrize("iprec", [1, 2, 3])
def test_plane_wave_pipeline(libs, iprec):
    h, ref = libs
    p = ORDER[iprec]
    ex = Ex(p)
    t = ref_tables(ref, iprec, p)
    assert h.mh_total_phys(iprec) == t["nexptotp"]
    rng = np.random.default_rng(10 + iprec)
    M = ex.rand(rng, 0.6)
    # multipole -> physical plane waves
    mexpf = np.zeros(t["nexptot"], complex)
    ph_ref = np.zeros(t["nexptotp"], complex)
    ref.mpoletoexp_(P(ex.full(M)), I(p), I(t["nl"]), P(t["nf"]), I(t["nexptot"]), P(mexpf), P(t["rlsc"]))
    ref.ftophys_(P(mexpf), I(t["nexptot"]), I(t["nl"]), P(t["xs"]), P(t["nf"]), P(t["nph"]), I(t["nthmax"]),
                 P(ph_ref), I(t["nexptotp"]), P(t["fexpe"]), I(t["ne"]))
    ph = np.zeros(2 * t["nexptotp"])
    h.mh_pw_from_multipole(iprec, p, P(M), P(ph))
    assert relmax(ph[0::2] + 1j * ph[1::2], ph_ref) < 1e-13
    # B scaling
    lam, mu, rscale = 1.3, 0.7, 1 / 0.37
    rl = np.array([lam, mu])
    b_ref = ph_ref.copy()
    ref.mindlinbpwdiagall_(D(rscale), P(rl), D(1.0), P(t["xs"]), P(t["ws"]), I(t["nl"]), P(t["nph"]),
                           I(t["nthmaxp"]), I(t["nexptotp"]), P(b_ref))
    h.mh_pw_scale_b(iprec, C.c_double(lam), C.c_double(mu), C.c_double(rscale), P(ph))
    assert relmax(ph[0::2] + 1j * ph[1::2], b_ref) < 1e-13
    # shift by (1, -2, -3) boxes
    dst_ref = np.zeros(t["nexptotp"], complex)
    ref.xyzshift_add_(P(ph_ref), P(t["xs"]), I(t["nl"]), P(t["nph"]), P(dst_ref), I(t["nexptotp"]), D(1.0),
                      D(-2.0), D(-3.0))
    src = np.zeros(2 * t["nexptotp"])
    src[0::2], src[1::2] = ph_ref.real, ph_ref.imag
    dst = np.zeros(2 * t["nexptotp"])
    h.mh_pw_shift_add(iprec, P(src), C.c_double(1.0), C.c_double(-2.0), C.c_double(-3.0), P(dst))
    assert relmax(dst[0::2] + 1j * dst[1::2], dst_ref) < 1e-13
    # physical -> local
    side = 0.37
    mexpf2 = np.zeros(t["nexptot"], complex)
    ref.phystof_(P(mexpf2), I(t["nexptot"]), I(t["nl"]), P(t["xs"]), P(t["nf"]), P(t["nph"]), I(t["nthmax"]),
                 P(dst_ref), I(t["nexptotp"]), P(t["fexpb"]), I(t["nb"]))
    loc_ref = np.zeros((p + 1, 2 * p + 1), complex, order="F")
    tmp = np.zeros((p + 1, 2 * p + 1), complex, order="F")
    rlp, facts = np.zeros(p + 1), np.zeros(2 * p + 1)
    zeye = np.zeros(max(t["nthmax"], p) + 1, complex)
    ref.exptolocald_add_(P(loc_ref), I(p), P(t["xs"]), P(t["ws"]), I(t["nl"]), P(t["nf"]), I(t["nthmax"]),
                         I(t["nexptot"]), P(mexpf2), P(tmp), D(side), D(1 / side), P(rlp), P(facts), P(zeye))
    lo = np.zeros((p + 1) * (p + 2))
    d2 = np.zeros(2 * t["nexptotp"])
    d2[0::2], d2[1::2] = dst_ref.real, dst_ref.imag
    h.mh_pw_to_local_add(iprec, p, C.c_double(side), P(d2), P(lo))
    assert relmax(lo, ex.packed(loc_ref)) < 1e-12


@pytest.mark.parametrize("iprec", [1, 2])
def test_b_proxies_and_remainder(libs, iprec):
    h, ref = libs
    p = ORDER[iprec]
    ex = Ex(p)
    rng = np.random.default_rng(20 + iprec)
    M = ex.rand(rng, 0.6)
    side = 0.37
    ctr = np.array([0.3, -0.1, 0.5])
    radius = side * np.sqrt(3) / 2
    nq = 2 * p
    xs_ref = np.zeros((2 * nq, 3))
    ch_ref = np.zeros(2 * nq, complex)
    ref.mptoslpborder_(P(ex.full(M)), I(p), P(ctr), D(1 / side), D(radius), I(nq), P(xs_ref), P(ch_ref))
    q, xs = np.zeros(2 * nq), np.zeros(6 * nq)
    h.mh_skeleton(p, P(M), P(ctr), C.c_double(1 / side), C.c_double(radius), nq, P(q), P(xs))
    assert np.abs(xs.reshape(-1, 3) - xs_ref).max() < 1e-14
    assert np.abs(ch_ref.imag).max() < 1e-12 * np.abs(ch_ref).max()
    assert relmax(q, ch_ref.real) < 1e-12
    # remainder: the proxies' multipole minus B, integrated twice in z
    Ms = ex.rand(rng, 0.6)
    Fm = ex.full(Ms)
    ref.mpdiff_(P(Fm), P(ex.full(M)), I(p))
    ref.mpdzminus2_(P(Fm), I(p))
    mine = M.copy()
    h.mh_b_remainder(p, P(Ms), P(mine))
    assert relmax(mine, ex.packed(Fm)) < 1e-13
    # the proxies' B-kernel local expansion about a neighbouring box centre
    lam, mu = 1.0, 0.8
    rl = np.array([lam, mu])
    ct_t = ctr + np.array([side, 0.0, -side])
    loc = np.zeros((p + 1, 2 * p + 1), complex, order="F")
    ier = C.c_int(0)
    ref.l3dformtaintkerbf90sc_(C.byref(ier), D(1 / side), D(1 / side), P(xs_ref), P(ch_ref), I(2 * nq), P(rl),
                               P(ct_t), D(radius), I(p), P(loc))
    lo = np.zeros((p + 1) * (p + 2))
    h.mh_proxy_local_add(p, 2 * nq, P(q), P(xs), P(ct_t), C.c_double(radius), C.c_double(1 / side),
                         C.c_double(1 / side), C.c_double(lam), C.c_double(mu), P(lo))
    assert relmax(lo, ex.packed(loc)) < 1e-11
