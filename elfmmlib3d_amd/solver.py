"""Iterative solvers driving repeated applies of the iter engine.

gmres_simple follows the reference's matlab/gmres_simple.m: a GMRES in
GCR form that keeps the orthonormalised images A e_i and the directions
e_i, re-orthogonalises once, recomputes the residual every 10 steps and
stops with ier 8 when the residual stops decreasing, ier 4 at maxit.
bicgstab_simple follows matlab/bicgstab_simple.m. Both work on any vectors
that support the torch/numpy operations used (torch tensors on the GPU keep
the whole solve in HBM).

IterOperator is the matrix-vector product of the elastic layer potentials
at the triangle centroids through elfmm3dtriatargiter: the first product
stores the near-field interactions on the device (icomp 2), every later one
replays them with the new density (icomp 3), which is what the reference's
iter protocol is for (src/elfmm3dtria_iter_1.f:97-113).
"""
from __future__ import annotations

import numpy as np


def _dot(a, b):
    return (a.conj() * b).sum() if hasattr(a, "conj") else np.vdot(a, b)


def _norm(a):
    return float(abs(_dot(a, a)) ** 0.5)


def gmres_simple(A, b, tol=1e-6, maxit=None, x0=None, log=None):
    """Solve A x = b (A: callable or matrix). Returns (x, ier, history);
    history holds norm(r)/norm(b) after each step."""
    if not callable(A):
        M = A
        A = (lambda x: M @ x)
    n = b.shape[0]
    if maxit is None:
        maxit = min(n, 20)
    Ae, E = [], []
    if x0 is not None:
        x = x0.clone() if hasattr(x0, "clone") else x0.copy()
        r = b - A(x)
    else:
        x = b * 0
        r = b.clone() if hasattr(b, "clone") else b.copy()
    er0 = _norm(r)
    nb = _norm(b)
    hist = []
    for i in range(1, maxit + 1):
        if i % 10 == 0:
            r = b - A(x)
        w = A(r)
        wp, rp = w, r
        for aj, ej in zip(Ae, E):
            d = _dot(aj, wp)
            wp = wp - d * aj
            rp = rp - d * ej
        d = 1.0 / _norm(wp)
        ai, ei = d * wp, d * rp
        for aj, ej in zip(Ae, E):
            d = _dot(aj, ai)
            ai = ai - d * aj
            ei = ei - d * ej
        d = 1.0 / _norm(ai)
        ai, ei = ai * d, ei * d
        Ae.append(ai)
        E.append(ei)
        d = _dot(ai, r)
        r = r - d * ai
        er1 = _norm(r)
        if er1 >= er0:
            return x, 8, hist
        er0 = er1
        x = x + d * ei
        hist.append(er1 / nb)
        if log:
            log(f"iter: {i}, rel=norm(r,2)/norm(b,2): {er1 / nb:17.12f}")
        if er1 < tol * nb:
            return x, 0, hist
    return x, 4, hist


def _sqrt(v):
    """Square root of a scalar as MATLAB takes it: real for a real v >= 0,
    complex otherwise."""
    import cmath
    import math
    v = complex(v) if not isinstance(v, (float, int)) else v
    if isinstance(v, complex) and v.imag == 0.0:
        v = v.real
    if isinstance(v, complex) or v < 0:
        return cmath.sqrt(v)
    return math.sqrt(v)


def _scalar(v):
    v = v.item() if hasattr(v, "item") else v
    return v.real if isinstance(v, complex) and v.imag == 0.0 else v


def bicgstab_simple(A, b, tol=1e-6, maxit=None, log=None):
    """Solve A x = b by the reference's BiCGSTAB (matlab/bicgstab_simple.m:
    36-108). Returns (x, ier, history); history holds norm(r)/norm(b) after
    each step.

    The recurrence is the reference's: x0 = 0, the shadow residual fixed to
    the first residual, two products per step (steps i = 2..maxit, so
    maxit - 1 of them), and beta = (sqrt(rho_i) / sqrt(rho_{i-1})) *
    (alpha / omega) with rho = <r0, r>, the ratio of square roots the
    listing takes (:88-90); it stops when norm(r) < tol * norm(r0)
    (:58,101-102; norm(r0) = norm(b) for a linear A). Two differences, both in
    what the listing returns rather than in the steps: ier is 0 when the
    tolerance is met and 4 when maxit is reached (the header's meaning;
    the listing sets ier = 4 after the loop either way, :108), and a
    1 x 1 system is solved as b / A(1) (the listing divides b by A(0),
    :42-45, which is 0 for a linear A)."""
    if not callable(A):
        M = A
        A = (lambda x: M @ x)
    n = b.shape[0]
    if maxit is None:
        maxit = min(n, 20)
    if n == 1:
        one = b * 0 + 1
        return b / A(one), 0, []
    x = b * 0
    r = b - A(x)
    p = r
    r0 = r
    norm1 = _norm(r)
    nb = _norm(b)
    er1 = _sqrt(_scalar(_dot(r0, r)))
    hist = []
    for i in range(2, maxit + 1):
        v = A(p)
        alpha = _scalar(_dot(r0, r)) / _scalar(_dot(r0, v))
        s = r - alpha * v
        t = A(s)
        omega = _scalar(_dot(t, s)) / _scalar(_dot(t, t))
        x = x + alpha * p + omega * s
        r = s - omega * t
        er = _norm(r)
        er2 = _sqrt(_scalar(_dot(r0, r)))
        beta = (er2 / er1) * (alpha / omega)
        er1 = er2
        p = r + beta * (p - omega * v)
        hist.append(er / nb)
        if log:
            log(f"iter: {i}, rel=norm(r,2)/norm(b,2): {er / nb:17.12f}")
        if er < tol * norm1:
            return x, 0, hist
    return x, 4, hist


class IterOperator:
    """x (n*3 vector) -> displacement (n*3) at the centroids of the elastic
    single (layer="single") or double (layer="double") layer potential with
    density x, through the iter engine: the first call stores the
    near-field interactions (icomp 2), later calls replay them (icomp 3).
    half=True adds the Mindlin half-space image part
    (elhfmm3dtriatargiter at the centroids), as the reference's half-space
    driver sums the two. device="cuda" keeps the geometry, the vectors and
    the products on the GPU (torch tensors in, torch tensors out; only the
    coordinates the host planner reads and the solver's scalars cross
    PCIe); device=None works on numpy arrays through the host entry points.
    Call release() to free the device stores."""

    def __init__(self, iprec, rlam, rmu, triangle, trinorm, source, layer="double", half=False, device=None):
        from . import _lib
        self._lib = _lib
        self.layer = layer
        self.half = half
        self.device = device
        self.n = triangle.shape[0]
        self.iprec, self.rlam, self.rmu = iprec, rlam, rmu
        self.wsave = np.zeros(4)
        self.wsave_h = np.zeros(4)
        self.calls = 0
        if device is not None:
            import torch
            self._torch = torch
            T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
            self.geo = (T(triangle), T(trinorm), T(source))
            self.out = torch.zeros((self.n, 3), dtype=torch.float64, device=device)
            self.out_h = torch.zeros((self.n, 3), dtype=torch.float64, device=device)
        else:
            self.geo = (triangle, trinorm, source)

    def _dens(self, s):
        return (s, None) if self.layer == "single" else (None, s)

    def __call__(self, x):
        icomp = 2 if self.calls == 0 else 3
        if self.device is None:
            s = np.asarray(x, dtype=np.float64).reshape(self.n, 3)
            args = (self.iprec, self.rlam, self.rmu, *self.geo, *self._dens(s))
            ier, _, pt, *_ = self._lib.elfmm3dtriatargiter(*args, None, ifstrain=0, icomp_type=icomp,
                                                           wsave=self.wsave)
            if ier:
                raise RuntimeError(f"elfmm3dtriatargiter ier={ier}")
            if self.half:
                ier, _, ph, _ = self._lib.elhfmm3dtriatargiter(*args, self.geo[2], 1, 0, icomp_type=icomp,
                                                               wsave=self.wsave_h)
                if ier:
                    raise RuntimeError(f"elhfmm3dtriatargiter ier={ier}")
                pt = pt + ph
            self.calls += 1
            return pt.reshape(-1)
        torch = self._torch
        s = torch.as_tensor(x, dtype=torch.float64, device=self.device).reshape(self.n, 3).contiguous()
        args = (self.iprec, self.rlam, self.rmu, *self.geo, *self._dens(s))
        ier, _ = self._lib.elfmm3dtriatargiter_dev(*args, None, self.out, None, None, None, icomp, self.wsave)
        if ier:
            raise RuntimeError(f"elfmm3dtriatargiter ier={ier}")
        y = self.out.clone()
        if self.half:
            ier, _ = self._lib.elhfmm3dtriatargiter_dev(*args, self.geo[2], self.out_h, None, icomp, self.wsave_h)
            if ier:
                raise RuntimeError(f"elhfmm3dtriatargiter ier={ier}")
            y += self.out_h
        self.calls += 1
        return y.reshape(-1)

    def release(self):
        r = self._lib.iter_release(self.wsave)
        if self.half:
            r += self._lib.iter_release(self.wsave_h)
        return r
