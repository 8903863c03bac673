"""MATLAB-gateway calling convention in Python: the triangle functions of
the reference's mwrap gateway (matlab/elfmm3d_r2013a.mw: elfmm3dtria :19,
el3dtriadirect :145, elhfmm3dtria :922, elh3dtriadirect :1072) with its
argument lists, optional trailing arguments and column-major shapes:
``triaflat`` (3, 3, nsource) = (coordinate, vertex, triangle), ``trianorm``,
``source``, ``charge``, ``dipstr`` (3, nsource), ``target`` (3, ntarget).
Each returns ``U`` with ``pot`` (3, nsource), ``strain`` (3, 3, nsource),
``pottarg`` (3, ntarget), ``straintarg`` (3, 3, ntarget) for the flags set,
and ``ier``; the work runs on the GPU through the library's entry points.

Defaults follow the gateway: 11 arguments -> source outputs only; 13 ->
no targets; 15 -> both target outputs. ``elhfmm3dtria`` returns the whole
half-space field, the image part (elhfmm3dtriatargiter, sources passed as
targets for the source outputs, as the gateway does) plus the free-space
part (elfmm3dtria); the gateway's own code overwrites the free-space part
with the image part before adding them (elfmm3d_r2013a.mw:1060-1063),
which this function does not reproduce."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from . import _lib


def _tri(nsource, triaflat):
    return np.ascontiguousarray(np.asarray(triaflat, dtype=np.float64).reshape(3, 3, nsource).transpose(2, 1, 0))


def _vec(n, a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(3, n).T)


def _flags(nargs, ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg):
    if nargs <= 11:
        ifpot, ifstrain = 1, 1
    if nargs <= 13:
        ntarget, target, ifpottarg, ifstraintarg = 0, None, 0, 0
    elif nargs <= 15:
        ifpottarg, ifstraintarg = 1, 1
    return ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg


def _out(ier, n, m, ifpot, ifstrain, ifpottarg, ifstraintarg, pt, st, ptt, stt):
    U = SimpleNamespace(ier=ier)
    if ifpot:
        U.pot = np.asarray(pt).reshape(n, 3).T.copy()
    if ifstrain:
        U.strain = np.asarray(st).reshape(n, 3, 3).transpose(2, 1, 0).copy()
    if ifpottarg:
        U.pottarg = np.asarray(ptt).reshape(m, 3).T.copy()
    if ifstraintarg:
        U.straintarg = np.asarray(stt).reshape(m, 3, 3).transpose(2, 1, 0).copy()
    return U


def _args(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr):
    return (_tri(nsource, triaflat), _vec(nsource, trianorm), _vec(nsource, source),
            _vec(nsource, charge) if ifcharge else None, _vec(nsource, dipstr) if ifdipole else None)


def elfmm3dtria(iprec, rlam, rmu, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                *rest):
    """[U] = elfmm3dtria(iprec,rlam,rmu,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,
    ifpot,ifstrain,ntarget,target,ifpottarg,ifstraintarg) (matlab/elfmm3d_r2013a.mw:19)."""
    rest = list(rest) + [None] * (6 - len(rest))
    ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg = _flags(11 + len([r for r in rest if r is not None]),
                                                                       *rest)
    tri, nrm, src, ssl, sdl = _args(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr)
    tg = _vec(ntarget, target) if ntarget else np.zeros((0, 3))
    ier, pt, st, ptt, stt = _lib.elfmm3dtriatarg(iprec, rlam, rmu, tri, nrm, src, ssl, sdl, tg, ifpot, ifstrain,
                                                 ifpottarg, ifstraintarg)
    return _out(ier, nsource, ntarget, ifpot, ifstrain, ifpottarg and ntarget, ifstraintarg and ntarget, pt, st,
                ptt, stt)


def el3dtriadirect(rlam, rmu, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr, *rest):
    """[U] = el3dtriadirect(rlam,rmu,nsource,triaflat,trianorm,source,ifcharge,charge,ifdipole,dipstr,
    ifpot,ifstrain,ntarget,target,ifpottarg,ifstraintarg) (matlab/elfmm3d_r2013a.mw:145)."""
    rest = list(rest) + [None] * (6 - len(rest))
    ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg = _flags(11 + len([r for r in rest if r is not None]),
                                                                       *rest)
    tri, nrm, src, ssl, sdl = _args(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr)
    tg = _vec(ntarget, target) if ntarget else np.zeros((0, 3))
    pt, st, ptt, stt = _lib.el3dtriadirecttarg(rlam, rmu, tri, nrm, src, ssl, sdl, tg, ifpot, ifstrain,
                                               ifpottarg, ifstraintarg)
    return _out(0, nsource, ntarget, ifpot, ifstrain, ifpottarg and ntarget, ifstraintarg and ntarget, pt, st,
                ptt, stt)


def elhfmm3dtria(iprec, rlam, rmu, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr,
                 *rest):
    """[U] = elhfmm3dtria(...) (matlab/elfmm3d_r2013a.mw:922): half-space
    field = image part + free-space part; U.ier = [ier_image, ier_free]."""
    rest = list(rest) + [None] * (6 - len(rest))
    ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg = _flags(11 + len([r for r in rest if r is not None]),
                                                                       *rest)
    tri, nrm, src, ssl, sdl = _args(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr)
    tg = _vec(ntarget, target) if ntarget else np.zeros((0, 3))
    ier_h = 0
    ph = np.zeros((nsource, 3))
    sh = np.zeros((nsource, 3, 3))
    if ifpot or ifstrain:
        ier_h, _, ph, sh = _lib.elhfmm3dtriatargiter(iprec, rlam, rmu, tri, nrm, src, ssl, sdl, src, ifpot, ifstrain)
    pth = np.zeros((ntarget, 3))
    sth = np.zeros((ntarget, 3, 3))
    if ntarget:
        ier2, _, pth, sth = _lib.elhfmm3dtriatargiter(iprec, rlam, rmu, tri, nrm, src, ssl, sdl, tg, ifpottarg,
                                                      ifstraintarg)
        ier_h = ier_h or ier2
    ier_f, pt, st, ptt, stt = _lib.elfmm3dtriatarg(iprec, rlam, rmu, tri, nrm, src, ssl, sdl, tg, ifpot, ifstrain,
                                                   ifpottarg, ifstraintarg)
    U = _out([ier_h, ier_f], nsource, ntarget, ifpot, ifstrain, ifpottarg and ntarget, ifstraintarg and ntarget,
             pt + ph, st + sh, ptt + pth, stt + sth)
    return U


def elh3dtriadirect(rlam, rmu, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr, *rest):
    """[U] = elh3dtriadirect(...) (matlab/elfmm3d_r2013a.mw:1072): the
    half-space direct sum (elh3dtriadirecttarg)."""
    rest = list(rest) + [None] * (6 - len(rest))
    ifpot, ifstrain, ntarget, target, ifpottarg, ifstraintarg = _flags(11 + len([r for r in rest if r is not None]),
                                                                       *rest)
    tri, nrm, src, ssl, sdl = _args(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr)
    tg = _vec(ntarget, target) if ntarget else np.zeros((0, 3))
    pt, st, ptt, stt = _lib.elh3dtriadirecttarg(rlam, rmu, tri, nrm, src, ssl, sdl, tg, ifpot, ifstrain, ifpottarg,
                                                ifstraintarg)
    return _out(0, nsource, ntarget, ifpot, ifstrain, ifpottarg and ntarget, ifstraintarg and ntarget, pt, st,
                ptt, stt)
