"""ctypes bindings of libelfmm_gfx950.so (see include/elfmm_gfx950.h)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._build import LIB

_lib = None


def library_path():
    return LIB


def library():
    """Load the in-tree library; raises if it has not been built.

    PyTorch-ROCm ships its own copy of the HIP runtime (same soname). Loading
    torch first makes this library bind to that already-loaded runtime, so
    one process never holds two HIP runtimes (the second one would find no
    GPU) and torch tensors can be passed to the device entry points."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run elfmmlib3d_amd.build()")
        try:
            import torch  # noqa: F401  (loads torch's HIP runtime first)
        except ImportError:
            pass
        _lib = C.CDLL(LIB)
        _lib.elfmm_gfx950_triatarg_dev.restype = C.c_int
        _lib.elfmm_gfx950_triatargiter_dev.restype = C.c_int
        _lib.elfmm_gfx950_htriatargiter_dev.restype = C.c_int
        _lib.elfmm_gfx950_parttarg_dev.restype = C.c_int
        _lib.elfmm_gfx950_version.restype = C.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, shape):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(shape)


def _iref(v):
    return C.byref(C.c_int(int(v)))


def _dref(v):
    return C.byref(C.c_double(float(v)))


def _prep(triangle, trinorm, source, sigma_sl, sigma_dl, target):
    n = triangle.shape[0]
    tri = _f64(triangle, (n, 3, 3))
    nrm = _f64(trinorm, (n, 3))
    src = _f64(source, (n, 3))
    ifs = int(sigma_sl is not None)
    ifd = int(sigma_dl is not None)
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    return n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt


def elfmm3dtriatarg(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                    target=None, ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """FMM displacement and strain at centroids and targets
    (src/elfmm3dtria.f:162). Returns (ier, ptfrc, strain, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _prep(triangle, trinorm, source, sigma_sl,
                                                          sigma_dl, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    ier = C.c_int(0)
    library().elfmm3dtriatarg_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _p(tri), _p(nrm),
                               _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd), _p(sdl),
                               _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st), _iref(m), _p(tgt),
                               _iref(ifptfrctarg if m else 0), _p(ptt),
                               _iref(ifstraintarg if m else 0), _p(stt))
    return ier.value, pt, st, ptt[:m], stt[:m]


def elfmm3dtria(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                ifptfrc=1, ifstrain=1):
    """On-surface FMM (src/elfmm3dtria.f:71). Returns (ier, ptfrc, strain)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, _, _ = _prep(triangle, trinorm, source, sigma_sl,
                                                        sigma_dl, None)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ier = C.c_int(0)
    library().elfmm3dtria_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _p(tri), _p(nrm),
                           _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd), _p(sdl),
                           _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st))
    return ier.value, pt, st


def el3dtriadirecttarg(rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                       target=None, ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """O(N^2) direct sum on the GPU (src/elfmm3dtria.f:1453).
    Returns (ptfrc, strain, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _prep(triangle, trinorm, source, sigma_sl,
                                                          sigma_dl, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    library().el3dtriadirecttarg_(_dref(rlam), _dref(rmu), _p(tri), _p(nrm), _iref(n), _p(src),
                                  _iref(ifs), _p(ssl), _iref(ifd), _p(sdl), _iref(ifptfrc), _p(pt),
                                  _iref(ifstrain), _p(st), _iref(m), _p(tgt),
                                  _iref(ifptfrctarg if m else 0), _p(ptt),
                                  _iref(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def elfmm3dtriatargiter(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                        target=None, ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1,
                        icomp_type=1, wsave=None, ptfrc=None, strain=None, ptfrctarg=None,
                        straintarg=None):
    """Iter FMM (src/elfmm3dtria_iter_1.f:160): icomp_type 1 compute, 2 compute
    and keep the near-field interactions, 3 replay them with new densities,
    4 plan only (returns lused), 5 store and add the near field only. The
    stored interactions live in device memory; wsave (float64, at least the
    lused words icomp 4 returns) carries their handle between calls.
    Output arrays may be passed in (icomp 5 adds to them). Returns
    (ier, lused, ptfrc, strain, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _prep(triangle, trinorm, source, sigma_sl,
                                                          sigma_dl, target)
    pt = np.zeros((n, 3)) if ptfrc is None else np.ascontiguousarray(ptfrc, dtype=np.float64)
    st = np.zeros((n, 3, 3)) if strain is None else np.ascontiguousarray(strain, dtype=np.float64)
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    if ptfrctarg is not None and m:
        ptt[:m] = ptfrctarg
    if straintarg is not None and m:
        stt[:m] = straintarg
    if wsave is None:
        wsave = np.zeros(1)
    ier, lused = C.c_int(0), C.c_int(0)
    library().elfmm3dtriatargiter_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _p(tri), _p(nrm),
                                   _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd), _p(sdl),
                                   _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st), _iref(m), _p(tgt),
                                   _iref(ifptfrctarg if m else 0), _p(ptt),
                                   _iref(ifstraintarg if m else 0), _p(stt), _iref(icomp_type),
                                   _p(wsave), _iref(wsave.size), C.byref(lused))
    return ier.value, lused.value, pt, st, ptt[:m], stt[:m]


def elhfmm3dtriatargiter(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                         target=None, ifptfrctarg=1, ifstraintarg=1, icomp_type=1, wsave=None):
    """Half-space (Mindlin) FMM, image part, at the targets
    (src/elhfmm3dtria_iter.f:46): icomp_type as elfmm3dtriatargiter (the
    image near field is what 2 keeps and 3 replays). Returns
    (ier, lused, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _prep(triangle, trinorm, source, sigma_sl,
                                                          sigma_dl, target)
    ptt = np.zeros((max(m, 1), 3))
    stt = np.zeros((max(m, 1), 3, 3))
    if wsave is None:
        wsave = np.zeros(1)
    ier, lused = C.c_int(0), C.c_int(0)
    library().elhfmm3dtriatargiter_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _p(tri), _p(nrm),
                                    _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd), _p(sdl), _iref(m),
                                    _p(tgt), _iref(ifptfrctarg if m else 0), _p(ptt),
                                    _iref(ifstraintarg if m else 0), _p(stt), _iref(icomp_type),
                                    _p(wsave), _iref(wsave.size), C.byref(lused))
    return ier.value, lused.value, ptt[:m], stt[:m]


def elh3dtriadirecttarg(rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
                        target=None, ifptfrc=0, ifstrain=0, ifptfrctarg=1, ifstraintarg=1):
    """O(N^2) half-space direct sum on the GPU (src/elh3dtria.f:53): free space
    plus the adaptive image integral of every pair.
    Returns (ptfrc, strain, ptfrctarg, straintarg)."""
    n, tri, nrm, src, ifs, ssl, ifd, sdl, m, tgt = _prep(triangle, trinorm, source, sigma_sl,
                                                          sigma_dl, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    library().elh3dtriadirecttarg_(_dref(rlam), _dref(rmu), _p(tri), _p(nrm), _iref(n), _p(src),
                                   _iref(ifs), _p(ssl), _iref(ifd), _p(sdl), _iref(ifptfrc), _p(pt),
                                   _iref(ifstrain), _p(st), _iref(m), _p(tgt),
                                   _iref(ifptfrctarg if m else 0), _p(ptt),
                                   _iref(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def half_parts(mask=7):
    """Diagnostic: which parts of the half-space apply run (bit 0 image near
    field, bit 1 anti-image far field, bit 2 Mindlin far field)."""
    library().elfmm_gfx950_half_parts(C.c_int(mask))


def expansion_order(iprec):
    """Expansion order the library uses at iprec (-1 outside -2..3)."""
    return library().elfmm_gfx950_expansion_order(C.c_int(iprec))


def iter_release(wsave):
    """Free the device-resident interactions named by wsave. Returns 1 if
    a store was freed."""
    return library().elfmm_gfx950_iter_release(_p(wsave), C.c_int(wsave.size))


def _prep_part(source, sigma_sl, sigma_dl, sigma_dv, target):
    n = source.shape[0]
    src = _f64(source, (n, 3))
    ifs = int(sigma_sl is not None)
    ifd = int(sigma_dl is not None)
    if ifd and sigma_dv is None:
        raise ValueError("sigma_dl needs sigma_dv (dipole orientations)")
    ssl = _f64(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f64(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    sdv = _f64(sigma_dv if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f64(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    return n, src, ifs, ssl, ifd, sdl, sdv, m, tgt


def elfmm3dparttarg(iprec, rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                    ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """FMM displacement and strain of point sources at the sources (self
    pair excluded) and at targets (src/elfmm3dpart.f:104). Returns
    (ier, ptfrc, strain, ptfrctarg, straintarg)."""
    n, src, ifs, ssl, ifd, sdl, sdv, m, tgt = _prep_part(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    ier = C.c_int(0)
    library().elfmm3dparttarg_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _iref(n), _p(src),
                               _iref(ifs), _p(ssl), _iref(ifd), _p(sdl), _p(sdv), _iref(ifptfrc), _p(pt),
                               _iref(ifstrain), _p(st), _iref(m), _p(tgt), _iref(ifptfrctarg if m else 0),
                               _p(ptt), _iref(ifstraintarg if m else 0), _p(stt))
    return ier.value, pt, st, ptt[:m], stt[:m]


def elfmm3dpart(iprec, rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, ifptfrc=1, ifstrain=1):
    """elfmm3dparttarg without targets (src/elfmm3dpart.f:52). Returns
    (ier, ptfrc, strain)."""
    n, src, ifs, ssl, ifd, sdl, sdv, _, _ = _prep_part(source, sigma_sl, sigma_dl, sigma_dv, None)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ier = C.c_int(0)
    library().elfmm3dpart_(C.byref(ier), _iref(iprec), _dref(rlam), _dref(rmu), _iref(n), _p(src), _iref(ifs),
                           _p(ssl), _iref(ifd), _p(sdl), _p(sdv), _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st))
    return ier.value, pt, st


def el3dpartdirecttarg(rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                       ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """O(N^2) point-source sum without self pairs (src/elfmm3dpart.f:506).
    Returns (ptfrc, strain, ptfrctarg, straintarg)."""
    n, src, ifs, ssl, ifd, sdl, sdv, m, tgt = _prep_part(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    library().el3dpartdirecttarg_(_dref(rlam), _dref(rmu), _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd),
                                  _p(sdl), _p(sdv), _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st), _iref(m),
                                  _p(tgt), _iref(ifptfrctarg if m else 0), _p(ptt),
                                  _iref(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def elh3dpartdirecttarg(rlam, rmu, source, sigma_sl=None, sigma_dl=None, sigma_dv=None, target=None,
                        ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """O(N^2) point-source sum with the Mindlin half-space Green's function,
    points in z < 0 (src/elhfmm3dpart.f:961). Returns (ptfrc, strain,
    ptfrctarg, straintarg)."""
    n, src, ifs, ssl, ifd, sdl, sdv, m, tgt = _prep_part(source, sigma_sl, sigma_dl, sigma_dv, target)
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    library().elh3dpartdirecttarg_(_dref(rlam), _dref(rmu), _iref(n), _p(src), _iref(ifs), _p(ssl), _iref(ifd),
                                   _p(sdl), _p(sdv), _iref(ifptfrc), _p(pt), _iref(ifstrain), _p(st), _iref(m),
                                   _p(tgt), _iref(ifptfrctarg if m else 0), _p(ptt),
                                   _iref(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def elfmm3dparttarg_dev(iprec, rlam, rmu, source, sigma_sl, sigma_dl, sigma_dv, target, ptfrc, strain, ptfrctarg,
                        straintarg, stream=None):
    """Device-resident point-source apply on torch tensors (float64,
    contiguous); any of sigma_sl / sigma_dl (with sigma_dv) / target /
    outputs may be None. Returns ier."""
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())
    n = source.shape[0]
    m = 0 if target is None else target.shape[0]
    s = C.c_void_p(0 if stream is None else stream.cuda_stream)
    return library().elfmm_gfx950_parttarg_dev(
        C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(n), ptr(source),
        C.c_int(sigma_sl is not None), ptr(sigma_sl), C.c_int(sigma_dl is not None), ptr(sigma_dl), ptr(sigma_dv),
        C.c_int(ptfrc is not None), ptr(ptfrc), C.c_int(strain is not None), ptr(strain), C.c_int(m), ptr(target),
        C.c_int(ptfrctarg is not None), ptr(ptfrctarg), C.c_int(straintarg is not None), ptr(straintarg), s)


def elfmm3dtriatarg_dev(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl, sigma_dl, target,
                        ptfrc, strain, ptfrctarg, straintarg, stream=None):
    """Device-resident apply on torch CUDA/HIP tensors (float64, contiguous).
    Any of sigma_sl/sigma_dl/target/outputs may be None. Returns ier."""
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())
    n = triangle.shape[0]
    m = 0 if target is None else target.shape[0]
    s = C.c_void_p(0 if stream is None else stream.cuda_stream)
    return library().elfmm_gfx950_triatarg_dev(
        C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(n), ptr(triangle), ptr(trinorm),
        ptr(source), C.c_int(sigma_sl is not None), ptr(sigma_sl), C.c_int(sigma_dl is not None),
        ptr(sigma_dl), C.c_int(ptfrc is not None), ptr(ptfrc), C.c_int(strain is not None),
        ptr(strain), C.c_int(m), ptr(target), C.c_int(ptfrctarg is not None), ptr(ptfrctarg),
        C.c_int(straintarg is not None), ptr(straintarg), s)


def elfmm3dtriatargiter_dev(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl, sigma_dl, target,
                            ptfrc, strain, ptfrctarg, straintarg, icomp_type, wsave, stream=None):
    """Device-resident iter apply (elfmm_gfx950_triatargiter_dev) on torch
    tensors (float64, contiguous); wsave is a numpy array (>= 4 words for a
    store). Any of sigma_sl/sigma_dl/target/outputs may be None. Returns
    (ier, lused)."""
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())
    n = triangle.shape[0]
    m = 0 if target is None else target.shape[0]
    s = C.c_void_p(0 if stream is None else stream.cuda_stream)
    lused = C.c_int(0)
    ier = library().elfmm_gfx950_triatargiter_dev(
        C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(n), ptr(triangle), ptr(trinorm),
        ptr(source), C.c_int(sigma_sl is not None), ptr(sigma_sl), C.c_int(sigma_dl is not None),
        ptr(sigma_dl), C.c_int(ptfrc is not None), ptr(ptfrc), C.c_int(strain is not None),
        ptr(strain), C.c_int(m), ptr(target), C.c_int(ptfrctarg is not None), ptr(ptfrctarg),
        C.c_int(straintarg is not None), ptr(straintarg), C.c_int(icomp_type), _p(wsave), C.c_int(wsave.size),
        C.byref(lused), s)
    return ier, lused.value


def elhfmm3dtriatargiter_dev(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl, sigma_dl, target,
                             ptfrctarg, straintarg, icomp_type, wsave, stream=None):
    """Device-resident half-space image part (elfmm_gfx950_htriatargiter_dev)
    on torch tensors. Returns (ier, lused)."""
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())
    n = triangle.shape[0]
    m = target.shape[0]
    s = C.c_void_p(0 if stream is None else stream.cuda_stream)
    lused = C.c_int(0)
    ier = library().elfmm_gfx950_htriatargiter_dev(
        C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(n), ptr(triangle), ptr(trinorm),
        ptr(source), C.c_int(sigma_sl is not None), ptr(sigma_sl), C.c_int(sigma_dl is not None),
        ptr(sigma_dl), C.c_int(m), ptr(target), C.c_int(ptfrctarg is not None), ptr(ptfrctarg),
        C.c_int(straintarg is not None), ptr(straintarg), C.c_int(icomp_type), _p(wsave), C.c_int(wsave.size),
        C.byref(lused), s)
    return ier, lused.value


def triatargiter_multi_dev(iprec, rlam, rmu, triangle, trinorm, source, sigma_sl, sigma_dl, target, ptfrc,
                           strain, ptfrctarg, straintarg, wsave, half=False, stream=None):
    """Replays of a store for several density sets in one call
    (elfmm_gfx950_triatargiter_multi_dev, or the half-space image part with
    half=True): torch tensors, densities (nrhs, n, 3), outputs (nrhs, ...).
    Returns ier."""
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())
    n = triangle.shape[0]
    m = 0 if target is None else target.shape[0]
    nrhs = (sigma_dl if sigma_dl is not None else sigma_sl).shape[0]
    s = C.c_void_p(0 if stream is None else stream.cuda_stream)
    common = [C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(n), ptr(triangle), ptr(trinorm),
              ptr(source), C.c_int(nrhs), C.c_int(sigma_sl is not None), ptr(sigma_sl),
              C.c_int(sigma_dl is not None), ptr(sigma_dl)]
    if half:
        return library().elfmm_gfx950_htriatargiter_multi_dev(
            *common, C.c_int(m), ptr(target), C.c_int(ptfrctarg is not None), ptr(ptfrctarg),
            C.c_int(straintarg is not None), ptr(straintarg), _p(wsave), C.c_int(wsave.size), s)
    return library().elfmm_gfx950_triatargiter_multi_dev(
        *common, C.c_int(ptfrc is not None), ptr(ptfrc), C.c_int(strain is not None), ptr(strain), C.c_int(m),
        ptr(target), C.c_int(ptfrctarg is not None), ptr(ptfrctarg), C.c_int(straintarg is not None),
        ptr(straintarg), _p(wsave), C.c_int(wsave.size), s)


class MeshBatch:
    """Many free-space meshes applied together (elfmm_gfx950_batch_*): each
    mesh's result is its elfmm3dtriatarg. ``meshes`` is a list of
    (triangle (n,3,3), trinorm (n,3), source (n,3), target (m,3) or None) as
    torch float64 tensors on the current GPU; the geometry is copied at
    creation. width 0 lets the library's tuner choose the schedule.
    apply() takes packed densities (sum n, 3) and returns packed outputs
    (ptfrc (sum n, 3), strain (sum n, 3, 3), ptfrctarg (sum m, 3),
    straintarg (sum m, 3, 3); None where off); split() cuts a packed array
    into per-mesh pieces."""

    def __init__(self, iprec, rlam, rmu, meshes, ifsingle=True, ifdouble=True, ifptfrc=True, ifstrain=True,
                 ifptfrctarg=True, ifstraintarg=True, width=0):
        import torch
        cat = lambda ts: torch.cat([t.reshape(t.shape[0], -1) for t in ts]).contiguous()
        self.n = [int(m[0].shape[0]) for m in meshes]
        self.m = [0 if m[3] is None else int(m[3].shape[0]) for m in meshes]
        self.flags = (bool(ifsingle), bool(ifdouble), bool(ifptfrc), bool(ifstrain), bool(ifptfrctarg) and
                      sum(self.m) > 0, bool(ifstraintarg) and sum(self.m) > 0)
        tri, nrm, src = cat([m[0] for m in meshes]), cat([m[1] for m in meshes]), cat([m[2] for m in meshes])
        tg = cat([m[3] for m in meshes if m[3] is not None]) if sum(self.m) else None
        nparts = (C.c_int * len(meshes))(*self.n)
        ntargs = (C.c_int * len(meshes))(*self.m)
        h = C.c_void_p()
        f = self.flags
        ier = library().elfmm_gfx950_batch_create(
            C.c_int(iprec), C.c_double(rlam), C.c_double(rmu), C.c_int(len(meshes)), nparts,
            C.c_void_p(tri.data_ptr()), C.c_void_p(nrm.data_ptr()), C.c_void_p(src.data_ptr()), C.c_int(f[0]),
            C.c_int(f[1]), C.c_int(f[2]), C.c_int(f[3]), ntargs, C.c_void_p(0 if tg is None else tg.data_ptr()),
            C.c_int(f[4]), C.c_int(f[5]), C.c_int(width), C.byref(h))
        if ier:
            raise RuntimeError(f"elfmm_gfx950_batch_create failed: ier={ier}")
        self.handle = h

    def apply(self, sigma_sl=None, sigma_dl=None, stream=None):
        import torch
        N, M = sum(self.n), sum(self.m)
        z = lambda *sh: torch.empty(sh, dtype=torch.float64, device="cuda")
        f = self.flags
        outs = [z(N, 3) if f[2] else None, z(N, 3, 3) if f[3] else None, z(M, 3) if f[4] else None,
                z(M, 3, 3) if f[5] else None]
        ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
        s = C.c_void_p(0 if stream is None else stream.cuda_stream)
        ier = library().elfmm_gfx950_batch_apply(self.handle, ptr(sigma_sl), ptr(sigma_dl), *map(ptr, outs), s)
        if ier:
            raise RuntimeError(f"elfmm_gfx950_batch_apply failed: ier={ier}")
        return outs

    def split(self, a, targets=False):
        counts = self.m if targets else self.n
        return list(a.split(counts)) if a is not None else None

    def schedule(self):
        """(width, graph, [(graph, width, ms), ...]) of the tuner."""
        w, g = C.c_int(), C.c_int()
        tune = (C.c_double * 96)()
        k = library().elfmm_gfx950_batch_schedule(self.handle, C.byref(w), C.byref(g), 32, tune)
        return w.value, g.value, [(int(tune[3 * i]), int(tune[3 * i + 1]), tune[3 * i + 2]) for i in range(k)]

    def close(self):
        if getattr(self, "handle", None):
            library().elfmm_gfx950_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def last_stats():
    """Work counts of the last apply on this thread (see include/elfmm_gfx950.h)."""
    out = np.zeros(16)
    library().elfmm_gfx950_last_stats(_p(out))
    keys = ["nboxes", "nlev", "p", "m2l_pairs", "m2l_flops_dense", "m2l_flops_alg", "near_pairs",
            "m2l_tiles", "m2m_pairs", "l2l_pairs", "p2m_boxes", "p2l_pairs", "m2t_pairs", "l2t_boxes",
            "iter_ref_store_words", "iter_store_words"]
    return dict(zip(keys, out.tolist()))


def last_timings():
    """Per-phase times (ms) of the last apply on this thread (include/elfmm_gfx950.h)."""
    out = np.zeros(16)
    library().elfmm_gfx950_last_timings(_p(out))
    keys = ["plan", "upload", "p2m", "m2m", "m2l", "p2l", "l2l", "l2t", "m2t", "near",
            "assemble", "total", "tree", "plan_tasks", "mindlin", "half_total"]
    return dict(zip(keys, out.tolist()))
