"""ctypes driver for the C restatement (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's
cpu_baseline leg use it to compare against; the product never imports it.
Build with ``make -C oracle oracle``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            raise RuntimeError(f"oracle not built: {SO} (make -C oracle oracle)")
        _lib = C.CDLL(SO)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f(a, shape):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(shape)


def tables(a, b, c, x0, y0, z0, iself=0):
    bt = np.zeros(125)
    rt = np.zeros(27)
    lib().oracle_tables(C.c_double(a), C.c_double(b), C.c_double(c), C.c_double(x0),
                        C.c_double(y0), C.c_double(z0), C.c_int(iself), _p(bt), _p(rt))
    return bt.reshape(5, 5, 5), rt.reshape(3, 3, 3)


def pair_sl(rlam, rmu, tri, sigma, target, ifself=0, ifstrain=1):
    t, s, x = _f(tri, (3, 3)), _f(sigma, (3,)), _f(target, (3,))
    pt, st = np.zeros(3), np.zeros((3, 3))
    lib().oracle_pair_sl(C.c_double(rlam), C.c_double(rmu), _p(t), _p(s), C.c_int(ifself),
                         _p(x), _p(pt), C.c_int(ifstrain), _p(st))
    return pt, st


def pair_dl(rlam, rmu, tri, sigma, target, ifself=0, ifstrain=1):
    t, s, x = _f(tri, (3, 3)), _f(sigma, (3,)), _f(target, (3,))
    pt, st = np.zeros(3), np.zeros((3, 3))
    lib().oracle_pair_dl(C.c_double(rlam), C.c_double(rmu), _p(t), _p(s), C.c_int(ifself),
                         _p(x), _p(pt), C.c_int(ifstrain), _p(st))
    return pt, st


def direct(rlam, rmu, triangle, trinorm, source, sigma_sl=None, sigma_dl=None,
           target=None, ifptfrc=1, ifstrain=1, ifptfrctarg=1, ifstraintarg=1):
    """el3dtriadirecttarg restated; returns (ptfrc, strain, ptfrctarg, straintarg)."""
    n = triangle.shape[0]
    tri, nrm, src = _f(triangle, (n, 3, 3)), _f(trinorm, (n, 3)), _f(source, (n, 3))
    ifs, ifd = int(sigma_sl is not None), int(sigma_dl is not None)
    ssl = _f(sigma_sl if ifs else np.zeros((n, 3)), (n, 3))
    sdl = _f(sigma_dl if ifd else np.zeros((n, 3)), (n, 3))
    if target is None:
        target = np.zeros((0, 3))
    m = target.shape[0]
    tgt = _f(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ptt, stt = np.zeros((max(m, 1), 3)), np.zeros((max(m, 1), 3, 3))
    lib().oracle_direct(C.c_double(rlam), C.c_double(rmu), _p(tri), _p(nrm), C.c_int(n), _p(src),
                        C.c_int(ifs), _p(ssl), C.c_int(ifd), _p(sdl), C.c_int(ifptfrc), _p(pt),
                        C.c_int(ifstrain), _p(st), C.c_int(m), _p(tgt),
                        C.c_int(ifptfrctarg if m else 0), _p(ptt),
                        C.c_int(ifstraintarg if m else 0), _p(stt))
    return pt, st, ptt[:m], stt[:m]


def tree(source, target, nbox, imindlin=0):
    """Octree + lists of d3tstrcr restated; same dict layout as ref.tree()."""
    n = source.shape[0]
    m = target.shape[0]
    src = _f(source, (n, 3))
    tgt = _f(target if m else np.zeros((1, 3)), (max(m, 1), 3))
    maxb = 16 * (n + m) // max(nbox, 1) + 4096
    nboxes = C.c_int(0)
    nlev = C.c_int(0)
    size = C.c_double(0)
    nitems = C.c_long(0)
    boxes = np.zeros((maxb, 20), dtype=np.int32)
    centers = np.zeros((maxb, 3))
    iz = np.zeros(n, dtype=np.int32)
    izt = np.zeros(max(m, 1), dtype=np.int32)
    laddr = np.zeros((200, 2), dtype=np.int32)
    center0 = np.zeros(3)
    ier = lib().oracle_tree(_p(src), C.c_int(n), _p(tgt), C.c_int(m), C.c_int(nbox), C.c_int(imindlin),
                            C.c_int(maxb), C.byref(nboxes), _p(boxes), _p(centers), _p(iz), _p(izt),
                            _p(laddr), C.byref(nlev), _p(center0), C.byref(size), None, None,
                            C.byref(nitems))
    if ier:
        raise RuntimeError(f"oracle_tree ier={ier}")
    nb = nboxes.value
    lstart = np.zeros(5 * (nb + 1), dtype=np.int32)
    litems = np.zeros(max(nitems.value, 1), dtype=np.int32)
    lib().oracle_tree(_p(src), C.c_int(n), _p(tgt), C.c_int(m), C.c_int(nbox), C.c_int(imindlin),
                      C.c_int(maxb), C.byref(nboxes), _p(boxes), _p(centers), _p(iz), _p(izt),
                      _p(laddr), C.byref(nlev), _p(center0), C.byref(size), _p(lstart), _p(litems),
                      C.byref(nitems))
    lists = {}
    for t in range(1, 6):
        base = (t - 1) * (nb + 1)
        lists[t] = [litems[lstart[base + b]:lstart[base + b + 1]].copy() for b in range(nb)]
    return dict(boxes=boxes[:nb].copy(), centers=centers[:nb].copy(), iz=iz.copy(),
                iztarg=izt[:m].copy(), laddr=laddr.copy(), nlev=nlev.value,
                center=center0.copy(), size=size.value, lists=lists)
