"""Serving: continuous batching of FMM applies on one GPU.

Requests (one elfmm3dtriatarg or elfmm3dtriatargiter apply each, of any
size) are admitted while a worker slot and device-memory budget are free;
each worker is a host thread with its own HIP stream and its own library
workspace, so one request's host planning (octree, lists, task lists)
overlaps other requests' kernels. The library runs the far fields of
different applies in the order the applies finish planning (an apply's
first far-field kernel waits for the last far-field kernel of the one
before); near fields run on their own streams beside them. A request's result does not depend on what runs beside it (the
library keeps no shared mutable state between applies except the cached
translation operators and the iter stores a request names itself).

    srv = Server(workers=3)
    fut = srv.submit(Request(iprec=2, rlam=1.0, rmu=1.0, triangle=..., ...))
    res = fut.result()          # Result(ier, ptfrc, strain, ptfrctarg, straintarg, latency_ms, ...)
    srv.close()

Inputs may be numpy arrays (host) or float64 torch tensors on the GPU;
outputs are torch tensors on the GPU (``Result.numpy()`` copies them out).
"""
from __future__ import annotations

import ctypes as C
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Request:
    iprec: int
    rlam: float
    rmu: float
    triangle: object
    trinorm: object
    source: object
    sigma_sl: object = None
    sigma_dl: object = None
    target: object = None
    ifptfrc: int = 1
    ifstrain: int = 1
    ifptfrctarg: int = 1
    ifstraintarg: int = 1
    # iter engine: icomp_type 1/2/3/5 with a caller-owned wsave (numpy,
    # >= 4 words) that names the device-resident near-field store; without
    # a wsave, 2 and 5 compute without keeping a store. Requests that share
    # a wsave run one at a time, in the order they were submitted.
    icomp_type: Optional[int] = None
    wsave: Optional[np.ndarray] = None
    # Mindlin half space z < 0 (targets only, as the reference's half-space
    # driver: the free-space iter apply plus elhfmm3dtriatargiter's image
    # part, summed). An iter half-space request's wsave holds two handles
    # (>= 8 words: the free-space store, then the image store).
    half: bool = False


@dataclass
class Result:
    ier: int
    ptfrc: object
    strain: object
    ptfrctarg: object
    straintarg: object
    latency_ms: float
    queue_ms: float
    timings: dict = field(default_factory=dict)
    wsave: Optional[np.ndarray] = None        # the request's wsave (iter requests)

    def numpy(self):
        f = lambda t: None if t is None else t.cpu().numpy()
        return f(self.ptfrc), f(self.strain), f(self.ptfrctarg), f(self.straintarg)


def estimate_bytes(n, nt, p=26, store_words=0):
    """Device memory of one apply's workspace (expansions, quadrature nodes,
    near-field records, per-point fields, inputs/outputs) at expansion order
    p, plus the iter store it keeps (store_words doubles), a conservative
    estimate used for admission. The L2T/M2T field operands are formed in
    LDS and take no device memory."""
    nboxes = 2 * (n + nt) // 40 + 64
    nc = (p + 1) * (p + 2)
    exps = 2 * nboxes * 4 * nc * 8
    pointwise = (n + nt) * (40 + 9 + 3 + 24) * 8 + n * (12 * 20 + 18 + 24) * 8
    io = (n * 27 + nt * 15) * 8
    return int(exps + pointwise + io + 8 * store_words) + (64 << 20)


class Server:
    def __init__(self, workers=3, hbm_budget=None, device=0):
        import torch
        from . import _lib
        self._torch = torch
        self._lib = _lib
        _lib.library()
        self.device = torch.device("cuda", device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.8)
        self.budget = hbm_budget
        self._inflight = 0
        self._cv = threading.Condition()
        self._q: "queue.Queue" = queue.Queue()
        self._closed = False
        self._ws_locks = {}                     # id(wsave) -> lock: one request per store at a time
        self._ws_mu = threading.Lock()
        self._threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(workers)]
        for t in self._threads:
            t.start()

    def submit(self, req: Request) -> Future:
        fut: Future = Future()
        lock = None
        if req.icomp_type is not None and req.wsave is not None:
            with self._ws_mu:
                lock = self._ws_locks.setdefault(id(req.wsave), _Turnstile())
            lock.ticket(fut)
        self._q.put((req, fut, time.perf_counter(), lock))
        return fut

    def close(self):
        self._closed = True
        for _ in self._threads:
            self._q.put(None)
        for t in self._threads:
            t.join()

    # -- worker side ------------------------------------------------------
    def _admit(self, need):
        with self._cv:
            # a request larger than the whole budget runs alone
            self._cv.wait_for(lambda: self._inflight == 0 or self._inflight + need <= self.budget)
            self._inflight += need

    def _release(self, need):
        with self._cv:
            self._inflight -= need
            self._cv.notify_all()

    def _worker(self):
        torch = self._torch
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            item = self._q.get()
            if item is None:
                return
            req, fut, t_submit, lock = item
            if lock is not None:
                lock.wait_turn(fut)
            need = 0
            try:
                n = req.triangle.shape[0]
                nt = 0 if req.target is None else req.target.shape[0]
                ins = self._inputs(req, stream)
                store_words = 0
                if req.icomp_type in (2, 5) and req.wsave is not None:
                    store_words = self._plan_store_words(req, ins, stream)
                need = estimate_bytes(n, nt, max(self._lib.expansion_order(req.iprec), 0), store_words)
                self._admit(need)
                t_start = time.perf_counter()
                res = self._run_half(req, ins, stream) if req.half else self._run(req, ins, stream)
                res.queue_ms = (t_start - t_submit) * 1e3
                res.latency_ms = (time.perf_counter() - t_submit) * 1e3
                fut.set_result(res)
            except Exception as e:                       # noqa: BLE001 (reported to the caller)
                fut.set_exception(e)
            finally:
                if need:
                    self._release(need)
                if lock is not None:
                    lock.done(fut)

    def _inputs(self, req: Request, stream):
        """The request's arrays on the device and its zeroed outputs."""
        torch = self._torch
        dev = self.device

        def dt(a, shape):
            if a is None:
                return None
            if isinstance(a, torch.Tensor):
                return a.to(device=dev, dtype=torch.float64).contiguous().reshape(shape)
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(shape)).to(
                dev, non_blocking=False)

        n = req.triangle.shape[0]
        nt = 0 if req.target is None else req.target.shape[0]
        with torch.cuda.stream(stream):
            tri, nrm, src = dt(req.triangle, (n, 3, 3)), dt(req.trinorm, (n, 3)), dt(req.source, (n, 3))
            ssl, sdl = dt(req.sigma_sl, (n, 3)), dt(req.sigma_dl, (n, 3))
            tgt = dt(req.target, (nt, 3)) if nt else None
            z = lambda shape, flag: torch.zeros(shape, dtype=torch.float64, device=dev) if flag else None
            pt, st = z((n, 3), req.ifptfrc and not req.half), z((n, 3, 3), req.ifstrain and not req.half)
            ptt, stt = z((nt, 3), req.ifptfrctarg and nt), z((nt, 3, 3), req.ifstraintarg and nt)
            stream.synchronize()
        return tri, nrm, src, ssl, sdl, tgt, pt, st, ptt, stt

    def _args(self, req, ins):
        tri, nrm, src, ssl, sdl, tgt, pt, st, ptt, stt = ins
        n = req.triangle.shape[0]
        nt = 0 if req.target is None else req.target.shape[0]
        ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
        return [C.c_int(req.iprec), C.c_double(req.rlam), C.c_double(req.rmu), C.c_int(n), ptr(tri), ptr(nrm),
                ptr(src), C.c_int(ssl is not None), ptr(ssl), C.c_int(sdl is not None), ptr(sdl),
                C.c_int(pt is not None), ptr(pt), C.c_int(st is not None), ptr(st), C.c_int(nt), ptr(tgt),
                C.c_int(ptt is not None), ptr(ptt), C.c_int(stt is not None), ptr(stt)]

    def _plan_store_words(self, req, ins, stream):
        """Words of the store(s) an icomp 2/5 request keeps (icomp-4 plans)."""
        lib = self._lib.library()
        ws = np.zeros(1)
        lused = C.c_int(0)
        s = C.c_void_p(stream.cuda_stream)
        if not req.half:
            ier = lib.elfmm_gfx950_triatargiter_dev(*self._args(req, ins), C.c_int(4), ws.ctypes.data_as(C.c_void_p),
                                                    C.c_int(1), C.byref(lused), s)
            return int(self._lib.last_stats()["iter_store_words"]) if ier == 0 else 0
        tri, nrm, src, ssl, sdl, tgt, pt, st, ptt, stt = ins
        n, nt = req.triangle.shape[0], tgt.shape[0]
        ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
        common = [C.c_int(req.iprec), C.c_double(req.rlam), C.c_double(req.rmu), C.c_int(n), ptr(tri), ptr(nrm),
                  ptr(src), C.c_int(ssl is not None), ptr(ssl), C.c_int(sdl is not None), ptr(sdl)]
        words = 0
        ier = lib.elfmm_gfx950_triatargiter_dev(
            *common, C.c_int(0), ptr(None), C.c_int(0), ptr(None), C.c_int(nt), ptr(tgt), C.c_int(ptt is not None),
            ptr(ptt), C.c_int(stt is not None), ptr(stt), C.c_int(4), ws.ctypes.data_as(C.c_void_p), C.c_int(1),
            C.byref(lused), s)
        words += int(self._lib.last_stats()["iter_store_words"]) if ier == 0 else 0
        ier = lib.elfmm_gfx950_htriatargiter_dev(
            *common, C.c_int(nt), ptr(tgt), C.c_int(ptt is not None), ptr(ptt), C.c_int(stt is not None), ptr(stt),
            C.c_int(4), ws.ctypes.data_as(C.c_void_p), C.c_int(1), C.byref(lused), s)
        words += int(self._lib.last_stats()["iter_store_words"]) if ier == 0 else 0
        return words

    def _run(self, req: Request, ins, stream):
        lib = self._lib.library()
        tri, nrm, src, ssl, sdl, tgt, pt, st, ptt, stt = ins
        args = self._args(req, ins)
        s = C.c_void_p(stream.cuda_stream)
        ws = None
        if req.icomp_type is None:
            ier = lib.elfmm_gfx950_triatarg_dev(*args, s)
        else:
            # without a wsave there is no room for a handle: no store is kept
            ws = req.wsave if req.wsave is not None else np.zeros(1)
            lused = C.c_int(0)
            ier = lib.elfmm_gfx950_triatargiter_dev(*args, C.c_int(req.icomp_type),
                                                    ws.ctypes.data_as(C.c_void_p), C.c_int(ws.size),
                                                    C.byref(lused), s)
        timings = self._lib.last_timings()           # same host thread: this request's phases
        return Result(ier, pt, st, ptt, stt, 0.0, 0.0, timings, req.wsave)

    def _run_half(self, req: Request, ins, stream):
        """Half-space request: free-space iter apply + image part at the
        targets (examples/elhfmm3dtria_self_iter_dr.f:219-373 sums them)."""
        torch = self._torch
        lib = self._lib.library()
        tri, nrm, src, ssl, sdl, tgt, pt, st, ptt, stt = ins
        n = req.triangle.shape[0]
        nt = 0 if req.target is None else req.target.shape[0]
        ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
        s = C.c_void_p(stream.cuda_stream)
        icomp = 1 if req.icomp_type is None else req.icomp_type
        if req.wsave is not None and req.wsave.size >= 8:
            wa, wb = req.wsave[0:4], req.wsave[4:8]
        else:
            wa, wb = np.zeros(1), np.zeros(1)
        lused = C.c_int(0)
        common = [C.c_int(req.iprec), C.c_double(req.rlam), C.c_double(req.rmu), C.c_int(n), ptr(tri), ptr(nrm),
                  ptr(src), C.c_int(ssl is not None), ptr(ssl), C.c_int(sdl is not None), ptr(sdl)]
        ier = lib.elfmm_gfx950_triatargiter_dev(
            *common, C.c_int(0), ptr(None), C.c_int(0), ptr(None), C.c_int(nt), ptr(tgt),
            C.c_int(ptt is not None), ptr(ptt), C.c_int(stt is not None), ptr(stt), C.c_int(icomp),
            wa.ctypes.data_as(C.c_void_p), C.c_int(wa.size), C.byref(lused), s)
        ta = self._lib.last_timings()
        if ier not in (0, 64):
            return Result(ier, None, None, ptt, stt, 0.0, 0.0, ta, req.wsave)
        with torch.cuda.stream(stream):
            pth = torch.zeros_like(ptt) if ptt is not None else None
            sth = torch.zeros_like(stt) if stt is not None else None
        ierh = lib.elfmm_gfx950_htriatargiter_dev(
            *common, C.c_int(nt), ptr(tgt), C.c_int(pth is not None), ptr(pth), C.c_int(sth is not None), ptr(sth),
            C.c_int(icomp), wb.ctypes.data_as(C.c_void_p), C.c_int(wb.size), C.byref(lused), s)
        tb = self._lib.last_timings()
        with torch.cuda.stream(stream):
            if ptt is not None:
                ptt += pth
            if stt is not None:
                stt += sth
            stream.synchronize()
        timings = {"arrival_" + k: v for k, v in ta.items()}
        timings.update({"image_" + k: v for k, v in tb.items()})
        return Result(ierh if ierh not in (0,) else ier, None, None, ptt, stt, 0.0, 0.0, timings, req.wsave)


class _Turnstile:
    """Runs the requests that share one wsave one at a time, in submission
    order."""

    def __init__(self):
        self._cv = threading.Condition()
        self._order = []

    def ticket(self, key):
        with self._cv:
            self._order.append(key)

    def wait_turn(self, key):
        with self._cv:
            self._cv.wait_for(lambda: self._order and self._order[0] is key)

    def done(self, key):
        with self._cv:
            self._order.remove(key)
            self._cv.notify_all()
