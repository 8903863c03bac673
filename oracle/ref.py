"""ctypes driver for the reference library compiled from its Fortran sources.

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's cpu_baseline
leg may load it to compare against; the product path never does.

``oracle/_ref/libelfmm_ref.so`` is built by ``make -C oracle ref`` from the
reference's sources where they lie (see oracle/Makefile). Every Fortran
argument is passed by reference; arrays are numpy C-contiguous arrays whose
trailing axis is the Fortran leading axis, so ``triangle[k, v, c]`` is the
Fortran ``triangle(c, v, k)``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libelfmm_ref.so")

_lib = None


def available():
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference library not built: {REF_SO}")
        _lib = C.CDLL(REF_SO)
    return _lib


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def elfmm3dtriatarg(iprec, rlam, rmu, triangle, trinorm, source,
                    sigma_sl=None, sigma_dl=None, target=None,
                    ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """Reference FMM (src/elfmm3dtria.f:162). Returns
    (ier, ptfrc[n,3], strain[n,3,3], ptfrctarg[m,3], straintarg[m,3,3])."""
    n = triangle.shape[0]
    tri = _f64(triangle, (n, 3, 3))
    nrm = _f64(trinorm, (n, 3))
    src = _f64(source, (n, 3))
    ifs = 0 if sigma_sl is None else 1
    ifd = 0 if sigma_dl is None else 1
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    ier = C.c_int(0)
    lib().elfmm3dtriatarg_(
        C.byref(ier), _i(iprec), _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n),
        _p(src), _i(ifs), _p(ssl), _i(ifd), _p(sdl),
        _i(ifptfrc), _p(pt), _i(ifstrain), _p(st),
        _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt))
    return ier.value, pt, st, ptt[:m], stt[:m]


def el3dtriadirecttarg(rlam, rmu, triangle, trinorm, source,
                       sigma_sl=None, sigma_dl=None, target=None,
                       ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """Reference O(N^2) direct sum (src/elfmm3dtria.f:1453)."""
    n = triangle.shape[0]
    tri = _f64(triangle, (n, 3, 3))
    nrm = _f64(trinorm, (n, 3))
    src = _f64(source, (n, 3))
    ifs = 0 if sigma_sl is None else 1
    ifd = 0 if sigma_dl is None else 1
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    lib().el3dtriadirecttarg_(
        _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n), _p(src),
        _i(ifs), _p(ssl), _i(ifd), _p(sdl),
        _i(ifptfrc), _p(pt), _i(ifstrain), _p(st),
        _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def pair_sl(rlam, rmu, tri, sigma, target, ifself=0, ifstrain=1):
    """elust3triadirecttarg_one (src/elfmm3dtria.f:1689): one (target, triangle)."""
    t = _f64(tri, (3, 3))
    s = _f64(sigma, (3,))
    x = _f64(target, (3,))
    pt = np.zeros(3)
    st = np.zeros((3, 3))
    lib().elust3triadirecttarg_one_(_d(rlam), _d(rmu), _p(t), _p(s),
                                    _i(ifself), _p(x), _p(pt), _i(ifstrain), _p(st))
    return pt, st


def pair_dl(rlam, rmu, tri, sigma, normal, target, ifself=0, ifstrain=1):
    """eltst3triadirecttarg_one (src/elfmm3dtria.f:2368): one (target, triangle)."""
    t = _f64(tri, (3, 3))
    s = _f64(sigma, (3,))
    nn = _f64(normal, (3,))
    x = _f64(target, (3,))
    pt = np.zeros(3)
    st = np.zeros((3, 3))
    lib().eltst3triadirecttarg_one_(_d(rlam), _d(rmu), _p(t), _p(s), _p(nn),
                                    _i(ifself), _p(x), _p(pt), _i(ifstrain), _p(st))
    return pt, st


def triabtable(ix, iy, iz, iquad, v1, v2, v3, x0, y0, z0):
    """Integral over a plane triangle of a derivative of r (src/triahquad.f:797)."""
    out = C.c_double(0.0)
    a, b, c = (_f64(v, (2,)) for v in (v1, v2, v3))
    lib().triabtable_(_i(ix), _i(iy), _i(iz), _i(iquad), _p(a), _p(b), _p(c),
                      _d(x0), _d(y0), _d(z0), C.byref(out))
    return out.value


def triartable(ix, iy, iz, iquad, v1, v2, v3, x0, y0, z0):
    """Integral over a plane triangle of a derivative of 1/r (src/triahquad.f:680)."""
    out = C.c_double(0.0)
    a, b, c = (_f64(v, (2,)) for v in (v1, v2, v3))
    lib().triartable_(_i(ix), _i(iy), _i(iz), _i(iquad), _p(a), _p(b), _p(c),
                      _d(x0), _d(y0), _d(z0), C.byref(out))
    return out.value


def triasymq(norder):
    """D3-symmetric triangle rule (src/triasymq.f:26) on (0,0),(1,0),(0,1)."""
    v1 = np.array([0.0, 0.0])
    v2 = np.array([1.0, 0.0])
    v3 = np.array([0.0, 1.0])
    nodes = np.zeros((500, 2))
    w = np.zeros(500)
    nn = C.c_int(0)
    lib().triasymq_(_i(norder), _p(v1), _p(v2), _p(v3), _p(nodes), _p(w), C.byref(nn))
    k = nn.value
    return nodes[:k].copy(), w[:k].copy()


def tree(source, target, nbox):
    """Octree + lists from d3tstrcr (src/d3tstrcr.f:31) with the free-space root.

    Returns a dict with boxes[nboxes,20] (Fortran 1-based contents), centers,
    iz/iztarg permutations (1-based), laddr, nlev, center, size and
    lists[itype][ibox] (1-based box numbers, itype 1..5)."""
    n = source.shape[0]
    m = target.shape[0]
    src = _f64(source, (n, 3))
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    lw = 200 * (n + m) + 100000
    for _ in range(8):
        w = np.zeros(lw, dtype=np.int32)
        iz = np.zeros(n, dtype=np.int32)
        izt = np.zeros(max(m, 1), dtype=np.int32)
        laddr = np.zeros((200, 2), dtype=np.int32)
        center = np.zeros(3)
        size = C.c_double(0.0)
        ier = C.c_int(0)
        nboxes = C.c_int(0)
        nlev = C.c_int(0)
        lused = C.c_int(0)
        lib().d3tstrcr_(C.byref(ier), _p(src), _i(n), _i(nbox), C.byref(nboxes), _p(iz),
                        _p(laddr), C.byref(nlev), _p(center), C.byref(size),
                        _p(tgt), _i(m), _p(izt), _p(w), _i(lw), C.byref(lused))
        if ier.value == 0:
            break
        lw *= 2
    else:
        raise RuntimeError(f"d3tstrcr failed ier={ier.value}")
    nb = nboxes.value
    boxes = np.zeros((nb, 20), dtype=np.int32)
    centers = np.zeros((nb, 3))
    box = np.zeros(20, dtype=np.int32)
    c0 = np.zeros(3)
    corners = np.zeros((8, 3))
    lists = {t: [] for t in range(1, 6)}
    lbuf = np.zeros(200000, dtype=np.int32)
    for ib in range(1, nb + 1):
        e = C.c_int(0)
        lib().d3tgetb_(C.byref(e), _i(ib), _p(box), _p(c0), _p(corners), _p(w))
        boxes[ib - 1] = box
        centers[ib - 1] = c0
        for t in range(1, 6):
            nl = C.c_int(0)
            lib().d3tgetl_(C.byref(e), _i(ib), _i(t), _p(lbuf), C.byref(nl), _p(w))
            lists[t].append(lbuf[: nl.value].copy() if e.value == 0 else np.zeros(0, np.int32))
    return dict(boxes=boxes, centers=centers, iz=iz.copy(), iztarg=izt[:m].copy(),
                laddr=laddr.copy(), nlev=nlev.value, center=center.copy(),
                size=size.value, lists=lists)


# ---------------------------------------------------------------------------
# Half-space link set (oracle/_ref/libelfmm_refh.so, `make -C oracle refh`):
# the iter entry points, built as the reference builds them, with the
# half-space root rule of d3tstrcr_mindlin.f.

REFH_SO = os.path.join(HERE, "_ref", "libelfmm_refh.so")
_libh = None


def available_h():
    return os.path.exists(REFH_SO)


def libh():
    global _libh
    if _libh is None:
        if not available_h():
            raise RuntimeError(f"reference library not built: {REFH_SO}")
        _libh = C.CDLL(REFH_SO)
    return _libh


def elfmm3dtriatargiter(iprec, rlam, rmu, triangle, trinorm, source,
                        sigma_sl=None, sigma_dl=None, target=None,
                        ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1,
                        icomp_type=1, wsave=None):
    """Reference iter FMM (src/elfmm3dtria_iter_1.f:160). Returns
    (ier, lused, ptfrc, strain, ptfrctarg, straintarg, wsave); wsave is the
    caller's array (allocated with the returned lused for icomp 2/3)."""
    n = triangle.shape[0]
    tri = _f64(triangle, (n, 3, 3))
    nrm = _f64(trinorm, (n, 3))
    src = _f64(source, (n, 3))
    ifs = 0 if sigma_sl is None else 1
    ifd = 0 if sigma_dl is None else 1
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    if wsave is None:
        wsave = np.zeros(1)
    ier = C.c_int(0)
    lused = C.c_int(0)
    libh().elfmm3dtriatargiter_(
        C.byref(ier), _i(iprec), _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n),
        _p(src), _i(ifs), _p(ssl), _i(ifd), _p(sdl), _i(ifptfrc), _p(pt),
        _i(ifstrain), _p(st), _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt), _i(icomp_type), _p(wsave), _i(wsave.size),
        C.byref(lused))
    return ier.value, lused.value, pt, st, ptt[:m], stt[:m], wsave


def _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl, target):
    n = triangle.shape[0]
    tri = _f64(triangle, (n, 3, 3))
    nrm = _f64(trinorm, (n, 3))
    src = _f64(source, (n, 3))
    ifs = 0 if sigma_sl is None else 1
    ifd = 0 if sigma_dl is None else 1
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    return n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt


def elhfmm3dtriatargiter(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                         target=None, ifptfrctarg=1, ifstraintarg=1, icomp_type=1, wsave=None):
    """Reference half-space FMM (src/elhfmm3dtria_iter.f:46). Returns
    (ier, lused, ptfrctarg, straintarg, wsave)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl,
                                                              target)
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    if wsave is None:
        wsave = np.zeros(1)
    ier = C.c_int(0)
    lused = C.c_int(0)
    libh().elhfmm3dtriatargiter_(
        C.byref(ier), _i(iprec), _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n), _p(src), _i(ifs), _p(ssl),
        _i(ifd), _p(sdl), _i(m), _p(tgt), _i(ifptfrctarg), _p(ptt), _i(ifstraintarg), _p(stt),
        _i(icomp_type), _p(wsave), _i(wsave.size), C.byref(lused))
    return ier.value, lused.value, ptt[:m], stt[:m], wsave


def elhfmm3d_near(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None, target=None,
                  ifptfrctarg=1, ifstraintarg=1, icomp_type=1):
    """The half-space image near field alone, elhfmm3dtriatargiter0
    (src/elhfmm3dtria_iter.f:1387), called as the entry point calls it
    (targets mirrored, :252-258) with a work array large enough for every
    icomp. Returns (ier, lused, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl,
                                                              target)
    tgtm = tgt.copy()
    tgtm[:, 2] *= -1
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    pt = np.zeros(3)
    st = np.zeros(9)
    ier = C.c_int(0)
    lused = C.c_int(0)
    zero = C.c_int(0)
    args = lambda icomp, ws: (C.byref(ier), _i(iprec), _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n), _p(src),
                              _i(ifs), _p(ssl), _i(ifd), _p(sdl), _p(nrm), C.byref(zero), _p(pt),
                              C.byref(zero), _p(st), _i(m), _p(tgtm), _i(ifptfrctarg), _p(ptt),
                              _i(ifstraintarg), _p(stt), _i(icomp), _p(ws), _i(ws.size), C.byref(lused))
    ws = np.zeros(1)
    libh().elhfmm3dtriatargiter0_(*args(4, ws))
    need = lused.value
    ws = np.zeros(max(need, 1) + 16)
    libh().elhfmm3dtriatargiter0_(*args(icomp_type, ws))
    return ier.value, need, ptt[:m].copy(), stt[:m].copy()


def elxfmm3d_anti(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None, target=None,
                  ifptfrctarg=1, ifstraintarg=1):
    """The half-space anti-image far field alone, as elhfmm3dtriatarg_image
    uses it (src/elhfmm3dtria_iter.f:747-777): elxfmm3dtriatarg with
    (lam + 4 mu, -mu), -sigma_dl and the mirrored targets, subtracted.
    Returns (ier, ptfrctarg, straintarg) with the subtraction applied."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl,
                                                              target)
    tgtm = tgt.copy()
    tgtm[:, 2] *= -1
    sdlm = -sdl
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    pt = np.zeros(3)
    st = np.zeros(9)
    ier = C.c_int(0)
    libh().elxfmm3dtriatarg_(
        C.byref(ier), _i(iprec), _d(rlam + 4 * rmu), _d(-rmu), _p(tri), _p(nrm), _i(n), _p(src), _i(ifs),
        _p(ssl), _i(ifd), _p(sdlm), _i(0), _p(pt), _i(0), _p(st), _i(m), _p(tgtm), _i(ifptfrctarg), _p(ptt),
        _i(ifstraintarg), _p(stt))
    return ier.value, -ptt[:m].copy(), -stt[:m].copy()


def lfmm3d_mindlin(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None, target=None):
    """The Mindlin harmonic-image FMM alone, as elhfmm3dtriatarg_image uses
    it (src/elhfmm3dtria_iter.f:795-831): lfmm3dmindlintriatarg on the
    mirrored triangles, its potential / mu added to the displacement and
    -sym(field) / (2 mu) to the strain. Returns (ier, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl,
                                                              target)
    trim = tri.copy()
    trim[:, :, 2] *= -1
    nrmm = nrm.copy()
    nrmm[:, 2] *= -1
    srcm = src.copy()
    srcm[:, 2] *= -1
    pot = np.zeros((max(m, 1), 3), dtype=np.complex128)
    fld = np.zeros((max(m, 1), 3, 3), dtype=np.complex128)
    rlame = np.array([rlam, rmu], dtype=np.float64)
    ier = C.c_int(0)
    libh().lfmm3dmindlintriatarg_(
        C.byref(ier), _i(iprec), _i(n), _p(trim), _p(nrmm), _p(srcm), _p(rlame), _i(ifs), _p(ssl), _i(ifd),
        _p(sdl), _p(nrm), _i(m), _p(tgt), _i(1), _p(pot), _i(1), _p(fld))
    pt = pot[:m].real / rmu
    f = fld[:m].real        # f[t, icomp, k] = Fortran fld(k, icomp, t)
    st = np.zeros((m, 3, 3))
    for ic in range(3):
        for k in range(3):
            # straintarg(k, ic) -= fld(k, ic)/(2mu); straintarg(ic, k) -= fld(k, ic)/(2mu)
            st[:, ic, k] -= f[:, ic, k] / (2 * rmu)
            st[:, k, ic] -= f[:, ic, k] / (2 * rmu)
    return ier.value, pt, st


def elh3dtriadirecttarg(rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None, target=None,
                        ifptfrc=0, ifstrain=0, ifptfrctarg=1, ifstraintarg=1):
    """Reference half-space direct sum (src/elhfmm3dtria_iter.f:3033, the copy
    of src/elh3dtria.f:53). Returns (ptfrc, strain, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _hs_inputs(triangle, trinorm, source, sigma_sl, sigma_dl,
                                                              target)
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    libh().elh3dtriadirecttarg_(
        _d(rlam), _d(rmu), _p(tri), _p(nrm), _i(n), _p(src), _i(ifs), _p(ssl), _i(ifd), _p(sdl), _i(ifptfrc),
        _p(pt), _i(ifstrain), _p(st), _i(m), _p(tgt), _i(ifptfrctarg), _p(ptt), _i(ifstraintarg), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def _part_inputs(source, sigma_sl, sigma_dl, sigma_dv, target):
    n = source.shape[0]
    src = _f64(source, (n, 3))
    ifs = 0 if sigma_sl is None else 1
    ifd = 0 if sigma_dl is None else 1
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    sdv = _f64(sigma_dv if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    return n, m, src, ifs, ssl, ifd, sdl, sdv, tgt


def elfmm3dparttarg(iprec, rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                    ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """Reference point-source FMM (src/elfmm3dpart.f:104-215). Returns
    (ier, ptfrc[n,3], strain[n,3,3], ptfrctarg[m,3], straintarg[m,3,3])."""
    n, m, src, ifs, ssl, ifd, sdl, sdv, tgt = _part_inputs(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    ier = C.c_int(0)
    lib().elfmm3dparttarg_(
        C.byref(ier), _i(iprec), _d(rlam), _d(rmu), _i(n), _p(src),
        _i(ifs), _p(ssl), _i(ifd), _p(sdl), _p(sdv),
        _i(ifptfrc), _p(pt), _i(ifstrain), _p(st),
        _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt))
    return ier.value, pt, st, ptt[:m], stt[:m]


def el3dpartdirecttarg(rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                       ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """Reference O(N^2) point-source sum without self pairs
    (src/elfmm3dpart.f:506-712)."""
    n, m, src, ifs, ssl, ifd, sdl, sdv, tgt = _part_inputs(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    lib().el3dpartdirecttarg_(
        _d(rlam), _d(rmu), _i(n), _p(src),
        _i(ifs), _p(ssl), _i(ifd), _p(sdl), _p(sdv),
        _i(ifptfrc), _p(pt), _i(ifstrain), _p(st),
        _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def elh3dpartdirecttarg(rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                        ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """Reference half-space point-source direct sum (src/elhfmm3dpart.f:961-1178):
    the Mindlin Green's function, the free-space part of the self pair
    excluded. Points in z < 0."""
    n, m, src, ifs, ssl, ifd, sdl, sdv, tgt = _part_inputs(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt = np.zeros((n, 3))
    st = np.zeros((n, 3, 3))
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    libh().elh3dpartdirecttarg_(
        _d(rlam), _d(rmu), _i(n), _p(src),
        _i(ifs), _p(ssl), _i(ifd), _p(sdl), _p(sdv),
        _i(ifptfrc), _p(pt), _i(ifstrain), _p(st),
        _i(m), _p(tgt), _i(ifptfrctarg if m else 0), _p(ptt),
        _i(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]
