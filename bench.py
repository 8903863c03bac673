"""FMM apply throughput on one MI355X: elastostatic single+double layer on
flat triangles (elfmm3dtriatarg), workload C2:
  icosphere 64x64 per face (81,920 triangles), densities uniform in [-1,1]^3,
  evaluation at all centroids plus 100,000 targets uniform in [-2,2]^3 minus
  the unit ball, lam = mu = 1, FP64, iprec 2 (tolerance 0.5e-6).
A step is one complete apply (tree + far field + near field + assembly) with
inputs already resident in HBM. The K timed steps are served as a stream of
requests with --workers applies in flight (library calls on separate host
threads and streams, so host planning and kernels of different requests
overlap); value = (triangles + targets) x K / wall time, latency = per
request from submission to result. The same K applies back to back on one
stream are reported as single_request. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workers 3] [--iprec 2] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time
from concurrent import futures

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12        # MI355X FP64 vector / matrix dense peak, FLOP/s
PEAK_HBM = 8.0e12          # bytes/s
# operations of one near-field pair (single + double layer, displacement and
# strain), counted by tests/csrc/nf_opcount.cpp on the library's pair code:
# 862 adds, 1441 muls, 31 divs, 25 sqrt/log/atan2
NEAR_OPS_PER_PAIR = 862 + 1441 + 31 + 25
# HBM bytes per launch on this workload, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, each scaled by what the counter reads for a 512 MiB device copy
# in the same run (profiles/c2_apply_kernels.md)
TRAFFIC = {"k_near": 211.5e6, "k_m2l": 8148.0e6}


def workload(nsplit=64, ntarg=100000, seed=2024):
    from elfmmlib3d_amd import mesh
    rng = np.random.default_rng(seed)
    tri = mesh.icosphere(nsplit)
    nrm = mesh.triangle_normals(tri)
    src = mesh.centroids(tri)
    ssl = rng.uniform(-1, 1, (tri.shape[0], 3))
    sdl = rng.uniform(-1, 1, (tri.shape[0], 3))
    tg = mesh.cube_targets(ntarg, 2.0, 1.0, rng)
    return tri, nrm, src, ssl, sdl, tg


def cpu_baseline(seconds_budget=30.0):
    """The reference's own FMM (oracle/_ref, OpenMP) on a bounded sample of
    the same workload: icosphere 16x16 per face (5,120 triangles) with 6,250
    targets in the same cube (same target/triangle ratio as C2), iprec 2."""
    from oracle import ref
    if not ref.available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    tri, nrm, src, ssl, sdl, tg = workload(16, 6250, 7)
    pts = tri.shape[0] + tg.shape[0]
    times = []
    t_end = time.time() + seconds_budget
    while True:
        t0 = time.time()
        ier, *_ = ref.elfmm3dtriatarg(2, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
        times.append(time.time() - t0)
        if time.time() > t_end or len(times) >= 3:
            break
    t = float(np.median(times))
    return {"value": pts / t / 1e6, "unit": "M tris+targets/s", "cores": threads,
            "kind": "reference",
            "sample": f"reference elfmm3dtriatarg (flang -O1 -fopenmp), icosphere 16x16/face "
                      f"(5120 tris) + 6250 cube targets, iprec 2, median of {len(times)} applies "
                      f"({t:.2f} s each)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--iprec", type=int, default=2)
    ap.add_argument("--nsplit", type=int, default=64)
    ap.add_argument("--ntarg", type=int, default=100000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", type=int, default=2000, help="targets checked against the direct sum")
    ap.add_argument("--workers", type=int, default=3,
                    help="requests in flight in the served measurement (1: single-request path only)")
    ap.add_argument("--no-iter-replay", dest="iter_replay", action="store_false",
                    help="skip the stored near-field replay measurement")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB first (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU per process; --gpus must be 1"}))
        sys.exit(2)

    import torch
    import elfmmlib3d_amd as E

    dev = torch.device("cuda:0")
    tri, nrm, src, ssl, sdl, tg = workload(args.nsplit, args.ntarg)
    n, m = tri.shape[0], tg.shape[0]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d = dict(triangle=T(tri), trinorm=T(nrm), source=T(src), sigma_sl=T(ssl), sigma_dl=T(sdl), target=T(tg))
    out = dict(ptfrc=torch.empty((n, 3), dtype=torch.float64, device=dev),
               strain=torch.empty((n, 3, 3), dtype=torch.float64, device=dev),
               ptfrctarg=torch.empty((m, 3), dtype=torch.float64, device=dev),
               straintarg=torch.empty((m, 3, 3), dtype=torch.float64, device=dev))

    def step():
        ier = E.elfmm3dtriatarg_dev(args.iprec, 1.0, 1.0, d["triangle"], d["trinorm"], d["source"],
                                    d["sigma_sl"], d["sigma_dl"], d["target"], out["ptfrc"], out["strain"],
                                    out["ptfrctarg"], out["straintarg"])
        if ier != 0:
            raise RuntimeError(f"ier={ier}")

    if args.calib_copy:
        a = torch.ones(args.calib_copy << 17, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # single-request path: K applies back to back on the default stream
    lat = []
    phases = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        lat.append(time.perf_counter() - ts)
        phases.append(E.last_timings())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = E.last_stats()
    single_total = t1 - t0
    single_value = (n + m) * args.steps / single_total / 1e6
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    single_lat = np.array(lat) * 1e3

    # served path: the same apply as a stream of requests with `workers` in
    # flight (each worker a host thread with its own stream); K requests timed
    served = None
    if args.workers > 1:
        from elfmmlib3d_amd.serve import Request, Server
        req = Request(args.iprec, 1.0, 1.0, d["triangle"], d["trinorm"], d["source"], d["sigma_sl"],
                      d["sigma_dl"], d["target"])
        srv = Server(workers=args.workers)
        for f in [srv.submit(req) for _ in range(max(args.warmup, args.workers))]:
            assert f.result().ier == 0
        torch.cuda.synchronize()
        # closed loop: `workers` requests in flight, a slot refilled as soon
        # as its request completes (in whatever order they complete)
        t0 = time.perf_counter()
        pending, results, submitted = set(), [], 0
        while submitted < min(args.workers, args.steps):
            pending.add(srv.submit(req))
            submitted += 1
        while pending:
            done, pending = futures.wait(pending, return_when=futures.FIRST_COMPLETED)
            for f in done:
                results.append(f.result())
                if submitted < args.steps:
                    pending.add(srv.submit(req))
                    submitted += 1
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        srv.close()
        assert all(r.ier == 0 for r in results)
        last = results[-1].numpy()
        diff = max(float(np.linalg.norm(a - b.cpu().numpy()) / np.linalg.norm(b.cpu().numpy()))
                   for a, b in zip(last, (out["ptfrc"], out["strain"], out["ptfrctarg"], out["straintarg"])))
        served = {"total_s": t1 - t0, "lat_ms": np.array([r.latency_ms for r in results]),
                  "vs_single_rel": diff,
                  "plan_ms": np.array([r.timings["plan"] for r in results])}

    if served:
        total = served["total_s"]
        lat_ms = served["lat_ms"]
    else:
        total = single_total
        lat_ms = single_lat
    ms_per_step = 1e3 * total / args.steps
    value = (n + m) * args.steps / total / 1e6

    # kernel times for the rooflines: a few more applies with the near field
    # on the caller's stream, so that every phase time is its kernels alone
    # (in the timed applies the near field overlaps the far field)
    os.environ["ELFMM_SERIAL_NEAR"] = "1"
    kph = []
    for _ in range(3):
        step()
        torch.cuda.synchronize()
        kph.append(E.last_timings())
    del os.environ["ELFMM_SERIAL_NEAR"]
    kph = {k: float(np.mean([p[k] for p in kph])) for k in kph[0]}

    # accuracy against the direct sum at sampled targets
    err = {}
    if args.check:
        rng = np.random.default_rng(5)
        idx = rng.choice(m, size=min(args.check, m), replace=False)
        dp, ds_, dpt, dst = E.el3dtriadirecttarg(1.0, 1.0, tri, nrm, src, ssl, sdl, tg[idx], 0, 0, 1, 1)
        fp = out["ptfrctarg"].cpu().numpy()[idx]
        fs = out["straintarg"].cpu().numpy()[idx]
        err = {"disp_rel_l2": float(np.linalg.norm(fp - dpt) / np.linalg.norm(dpt)),
               "strain_rel_l2": float(np.linalg.norm(fs - dst) / np.linalg.norm(dst)),
               "checked_targets": int(len(idx))}

    # iter engine (elfmm3dtriatargiter): the near field stored once (icomp 2)
    # and replayed with the densities (icomp 3), the matvec of an iterative
    # solve. The replay streams the store from HBM once per apply: bytes per
    # launch = stored words x 8 over the replay kernel's time (near field on
    # the apply's own stream). Workload: the mesh above at its centroids,
    # SL + DL displacement. Outside the timed region.
    replay = None
    if args.iter_replay:
        os.environ["ELFMM_SERIAL_NEAR"] = "1"
        ws = None
        try:
            kw = dict(ifptfrc=1, ifstrain=0)
            ier, lused, *_ = E.elfmm3dtriatargiter(args.iprec, 1.0, 1.0, tri, nrm, src, ssl, sdl, None,
                                                   icomp_type=4, **kw)
            ws = np.zeros(max(lused, 4))
            ier, *_ = E.elfmm3dtriatargiter(args.iprec, 1.0, 1.0, tri, nrm, src, ssl, sdl, None,
                                            icomp_type=2, wsave=ws, **kw)
            words = E.last_stats()["iter_store_words"]
            tms, tap = [], []
            for _ in range(3):
                ier, *_ = E.elfmm3dtriatargiter(args.iprec, 1.0, 1.0, tri, nrm, src, ssl, sdl, None,
                                                icomp_type=3, wsave=ws, **kw)
                if ier != 0:
                    raise RuntimeError(f"iter replay ier={ier}")
                tm = E.last_timings()
                tms.append(tm["near"])
                tap.append((tm["total"], tm["plan"]))
            ms = float(np.median(tms))
            bts = words * 8.0
            replay = {"kernel": "k_near_replay", "bound": "hbm", "achieved": bts / ms / 1e6,
                      "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bts / (ms / 1e3) / PEAK_HBM,
                      "traffic": None, "algorithmic_bytes_per_launch": bts, "ms_per_launch": ms,
                      "product_ms": float(np.median([x[0] for x in tap])),
                      "product_host_plan_ms": float(np.median([x[1] for x in tap])),
                      "workload": f"{n} triangles at their centroids, SL+DL displacement, iprec {args.iprec}"}
        finally:
            if ws is not None:
                E.iter_release(ws)
            del os.environ["ELFMM_SERIAL_NEAR"]

    # dominant kernel and its roofline. Algorithmic work per launch:
    #  near field: near (point, triangle) pairs x NEAR_OPS_PER_PAIR, the
    #    counted operations of one pair of the library's pair quadrature
    #    (tests/csrc/nf_opcount.cpp: adds, muls, divs, 1 per sqrt/log/atan2)
    #  M2L: per pair and right-hand side the reference's rotate/shift/rotate
    #    count 2 R(p_t) + Z(p_t) (elfmm_gfx950_last_stats()[5])
    kernels = {k: kph[k] for k in ("p2m", "m2m", "m2l", "p2l", "l2l", "l2t", "m2t", "near")}
    work = {"near": (stats["near_pairs"] * NEAR_OPS_PER_PAIR, "k_near"),
            "m2l": (stats["m2l_flops_alg"], "k_m2l")}
    rooflines = {}
    for k, (alg, name) in work.items():
        t = kernels[k] / 1e3
        rooflines[k] = {"kernel": name, "bound": "fp64", "achieved": alg / t / 1e12, "peak": PEAK_FP64 / 1e12,
                        "unit": "TFLOP/s", "frac": alg / t / PEAK_FP64, "traffic": None,
                        "algorithmic_flop_per_launch": alg, "ms_per_launch": kernels[k]}
    rooflines["m2l"]["executed_tflops"] = stats["m2l_flops_dense"] / (kernels["m2l"] / 1e3) / 1e12
    dom = max(kernels, key=kernels.get)
    roof = rooflines.get(dom, {"kernel": dom, "bound": "fp64", "achieved": None, "peak": PEAK_FP64 / 1e12,
                               "unit": "TFLOP/s", "frac": None, "traffic": None,
                               "ms_per_launch": kernels[dom]})
    if TRAFFIC.get(roof["kernel"]) is not None:
        roof["traffic"] = TRAFFIC[roof["kernel"]]

    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "FMM apply throughput (M tris+targets/s) @1e-6 err; p50/p99 served latency",
        "value": value, "unit": "M tris+targets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C2: icosphere 64x64/face (81,920 tris) SL+DL, centroids + 100k cube "
                               "targets, lam=mu=1", "iprec": args.iprec, "triangles": n, "targets": m,
                   "tolerance": {1: 0.5e-3, 2: 0.5e-6, 3: 0.5e-9}.get(args.iprec)},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "served": ({"workers": args.workers, "requests": args.steps,
                    "result_vs_single_request_rel": served["vs_single_rel"],
                    "host_plan_ms": {"p50": float(np.percentile(served["plan_ms"], 50)),
                                     "max": float(served["plan_ms"].max())}} if served else None),
        "single_request": {"value": single_value, "ms_per_step": 1e3 * single_total / args.steps,
                           "latency_ms": {"p50": float(np.percentile(single_lat, 50)),
                                          "p99": float(np.percentile(single_lat, 99))}},
        "accuracy_vs_direct": err,
        "phases_ms": ph, "kernel_phases_ms": kph, "work": stats,
        "roofline": roof, "rooflines": dict(rooflines, **({"iter_replay": replay} if replay else {})),
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_reference"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
