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
stream are reported as single_request. Beside the C2 line: the C3 half-space
apply and its 64-slip-vector call (half_space_c3), the C4 batch of 256
meshes (batched_c4), the C5 served stream (served_c5), the point-source FMM
on 1 M points (points_fmm), the iter replay's HBM roofline, the per-kernel
HBM traffic (rocprofv3 child runs) and the reference's CPU time on the host
cores (cpu_baseline). Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workers 3] [--iprec 2] [--no-cpu]
                  [--no-c3] [--no-c4] [--no-c5] [--no-part] [--no-traffic]
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
CALIB_MIB = 512            # device copy whose counters scale FETCH_SIZE / WRITE_SIZE to bytes


def workload(nsplit=64, ntarg=100000, seed=2024):
    from elfmmlib3d_amd import mesh
    return mesh.sphere_workload(nsplit, ntarg, seed)


def cpu_baseline(points_sample=None):
    """The reference's own FMM (oracle/_ref: flang -O1 -fopenmp, OpenMP over
    the host cores this process may use, OMP_STACKSIZE 1024M) on this host:
    C1 (icosphere 12x12/face, 2,880 triangles + 1,000 shell targets) as a
    warm-up and its own figure (median of 5 applies), then the full C2
    workload (81,920 triangles + 100,000 targets), one apply."""
    from oracle import ref
    if not ref.available():
        return None
    from elfmmlib3d_amd import mesh
    os.environ.setdefault("OMP_STACKSIZE", "1024M")
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or ncpu
    os.environ["OMP_NUM_THREADS"] = str(threads)
    rng = np.random.default_rng(1)
    tri = mesh.icosphere(12)
    c1 = (tri, mesh.triangle_normals(tri), mesh.centroids(tri), rng.uniform(-1, 1, (tri.shape[0], 3)),
          rng.uniform(-1, 1, (tri.shape[0], 3)), mesh.shell_targets(1000, 1.5, 2.0, rng))
    t1 = []
    for _ in range(5):
        t0 = time.time()
        ier, *_ = ref.elfmm3dtriatarg(2, 1.0, 1.0, *c1)
        t1.append(time.time() - t0)
    c1_t = float(np.median(t1))
    c2 = workload()
    t0 = time.time()
    ier, *_ = ref.elfmm3dtriatarg(2, 1.0, 1.0, *c2)
    c2_t = time.time() - t0
    pts = c2[0].shape[0] + c2[5].shape[0]
    part = None
    if points_sample is not None:
        arrs, pg = points_sample
        t0 = time.time()
        ier, pr, *_ = ref.elfmm3dparttarg(2, 1.0, 1.0, *arrs, None, 1, 1, 0, 0)
        tp = time.time() - t0
        part = {"sample": "reference elfmm3dparttarg, 100,000 points uniform in the unit cube, SL + DL, "
                          "iprec 2, one apply", "s_per_apply": tp, "value": 1e5 / tp / 1e6, "unit": "M points/s",
                "gpu_vs_reference_rel": float(np.linalg.norm(pg - pr) / np.linalg.norm(pr))}
    return {"value": pts / c2_t / 1e6, "unit": "M tris+targets/s", "cores": threads, "kind": "reference",
            "points": part,
            "sample": f"reference elfmm3dtriatarg (flang -O1 -fopenmp, OMP_STACKSIZE 1024M) on the full C2 "
                      f"workload, one apply after 5 C1 applies: {c2_t:.1f} s; {threads} threads of "
                      f"{os.cpu_count()} CPUs",
            "c1": {"value": (c1[0].shape[0] + c1[5].shape[0]) / c1_t / 1e6, "s_per_apply": c1_t,
                   "applies": len(t1)},
            "c2_s_per_apply": c2_t}


def kernel_pass(args):
    """Child of the traffic measurement: a 512 MiB device copy (counter
    calibration), then two applies with the near field on the apply's own
    stream, under rocprofv3 --pmc."""
    import torch
    import elfmmlib3d_amd as E
    os.environ["ELFMM_SERIAL_NEAR"] = "1"
    dev = torch.device("cuda:0")
    tri, nrm, src, ssl, sdl, tg = workload(args.nsplit, args.ntarg)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    n, m = tri.shape[0], tg.shape[0]
    z = lambda *sh: torch.empty(sh, dtype=torch.float64, device=dev)
    a = torch.ones(CALIB_MIB << 17, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    ins = [T(x) for x in (tri, nrm, src, ssl, sdl, tg)]
    outs = [z(n, 3), z(n, 3, 3), z(m, 3), z(m, 3, 3)]
    for _ in range(2):
        ier = E.elfmm3dtriatarg_dev(args.iprec, 1.0, 1.0, *ins, *outs)
        assert ier == 0
    torch.cuda.synchronize()


def measure_traffic(args):
    """HBM bytes per dispatch of every kernel of the apply: two rocprofv3
    --pmc runs of kernel_pass (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the counter reads for the 512 MiB device
    copy of the same run. Returns ({kernel: bytes}, factors) or (None, why)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    kib = {}
    factors = {}
    with tempfile.TemporaryDirectory() as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "-f", "csv", "-d", out, "-o", counter, "--", sys.executable,
                   os.path.abspath(__file__), "--kernel-pass", "--iprec", str(args.iprec), "--nsplit",
                   str(args.nsplit), "--ntarg", str(args.ntarg)]
            env = dict(os.environ, TMPDIR=tmp)
            try:
                r = subprocess.run(cmd, cwd=tmp, env=env, capture_output=True, text=True, timeout=300)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 {counter} pass timed out"
            files = glob.glob(os.path.join(out, "**", "*counter_collection*.csv"), recursive=True)
            if r.returncode != 0 and not files:
                return None, f"rocprofv3 {counter} pass failed ({r.returncode})"
            per = {}
            for f in files:
                with open(f) as fh:
                    for row in csv.DictReader(fh):
                        name = row.get("Kernel_Name", "")
                        if row.get("Counter_Name", counter) != counter:
                            continue
                        per.setdefault(name, []).append(float(row["Counter_Value"]))
            copies = [v for k, vs in per.items() if "copyBuffer" in k for v in vs]
            if not copies:
                return None, f"no device copy in the {counter} pass"
            factors[counter] = (CALIB_MIB << 20) / (max(copies) * 1024.0)
            for name, vs in per.items():
                short = name.split("(")[0].split("::")[-1].split("<")[0]
                kib.setdefault(short, {})[counter] = float(np.mean(vs))
    traffic = {k: v.get("FETCH_SIZE", 0.0) * 1024 * factors["FETCH_SIZE"] +
               v.get("WRITE_SIZE", 0.0) * 1024 * factors["WRITE_SIZE"] for k, v in kib.items()
               if k.startswith("k_")}
    return traffic, factors


def c3_half(args):
    """Config C3 on one GPU: 8 planar faults in z < 0 (mesh.fault_mesh, seed
    33) meshed into about args.c3_tri right triangles, slip (double layer)
    uniform in [-1,1]^3, displacement and strain at the centroids, iprec 2:
    the reference driver's sequence (icomp 4, 2, then 3 with new slip) for
    the direct arrival (elfmm3dtriatargiter) and the image part
    (elhfmm3dtriatargiter). An apply = both replays; value = triangles per
    second of an apply. The image-part kernels' times come from the
    library's phase timers of the same calls."""
    import torch
    import elfmmlib3d_amd as E
    from elfmmlib3d_amd import mesh
    rng = np.random.default_rng(33)
    tri = mesh.fault_mesh(8, args.c3_tri, rng)
    nrm, src = mesh.triangle_normals(tri), mesh.centroids(tri)
    n = tri.shape[0]
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_tri, d_nrm, d_src = T(tri), T(nrm), T(src)
    z = lambda *sh: torch.zeros(sh, dtype=torch.float64, device=dev)
    pa, sa, ph, sh = z(n, 3), z(n, 3, 3), z(n, 3), z(n, 3, 3)
    lib = E.library()
    import ctypes as C
    ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
    stream = C.c_void_p(0)
    wa, wb = np.zeros(4), np.zeros(4)
    lused = C.c_int(0)

    def arrival(icomp, slip):
        ier = lib.elfmm_gfx950_triatargiter_dev(
            C.c_int(args.iprec), C.c_double(1.0), C.c_double(1.0), C.c_int(n), ptr(d_tri), ptr(d_nrm), ptr(d_src),
            C.c_int(0), ptr(None), C.c_int(1), ptr(slip), C.c_int(0), ptr(None), C.c_int(0), ptr(None), C.c_int(n),
            ptr(d_src), C.c_int(1), ptr(pa), C.c_int(1), ptr(sa), C.c_int(icomp), wa.ctypes.data_as(C.c_void_p),
            C.c_int(4), C.byref(lused), stream)
        return ier, E.last_stats()["iter_store_words"]

    def image(icomp, slip):
        ier = lib.elfmm_gfx950_htriatargiter_dev(
            C.c_int(args.iprec), C.c_double(1.0), C.c_double(1.0), C.c_int(n), ptr(d_tri), ptr(d_nrm), ptr(d_src),
            C.c_int(0), ptr(None), C.c_int(1), ptr(slip), C.c_int(n), ptr(d_src), C.c_int(1), ptr(ph), C.c_int(1),
            ptr(sh), C.c_int(icomp), wb.ctypes.data_as(C.c_void_p), C.c_int(4), C.byref(lused), stream)
        return ier, E.last_timings()

    try:
        slip = T(rng.uniform(-1, 1, (n, 3)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ier_a, words_a = arrival(2, slip)
        torch.cuda.synchronize()
        t_store_a = time.perf_counter() - t0
        t0 = time.perf_counter()
        ier_b, _ = image(2, slip)
        words_b = E.last_stats()["iter_store_words"]
        torch.cuda.synchronize()
        t_store_b = time.perf_counter() - t0
        assert ier_a == 0 and ier_b == 0, (ier_a, ier_b)
        ta, tb, tim = [], [], []
        for _ in range(args.c3_replays):
            slip = T(rng.uniform(-1, 1, (n, 3)))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ier, _ = arrival(3, slip)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ier2, tm = image(3, slip)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            assert ier == 0 and ier2 == 0, (ier, ier2)
            ta.append(t1 - t0)
            tb.append(t2 - t1)
            tim.append(tm)
        apply_s = float(np.median(np.array(ta) + np.array(tb)))
        multi = c3_multi_rhs(args, E, n, d_tri, d_nrm, d_src, wa, wb, arrival, image, pa, sa, ph, sh, rng)
        return {"workload": f"C3: 8 faults, {n} triangles, slip U[-1,1]^3, centroids, iprec {args.iprec}",
                "value": n / apply_s / 1e6, "unit": "M tris/s", "apply_ms": apply_s * 1e3,
                "arrival_replay_ms": float(np.median(ta)) * 1e3, "image_replay_ms": float(np.median(tb)) * 1e3,
                "image_mindlin_ms": float(np.median([x["mindlin"] for x in tim])),
                "store_s": {"arrival": t_store_a, "image": t_store_b},
                "store_gb": {"arrival": words_a * 8 / 1e9, "image": words_b * 8 / 1e9},
                "replays": args.c3_replays, "multi_rhs": multi}
    finally:
        E.iter_release(wa)
        E.iter_release(wb)


def c3_multi_rhs(args, E, n, d_tri, d_nrm, d_src, wa, wb, arrival, image, pa, sa, ph, sh, rng):
    """C4's first half: args.c3_rhs slip vectors on the C3 mesh in one call
    per part (triatargiter_multi_dev on both stores: near field of 4 sets
    per pass over the store, far field per set), against the same sets
    replayed one call at a time. Set 0 is checked against its single
    replay."""
    import torch
    R = args.c3_rhs
    if R < 2:
        return None
    dev = d_src.device
    slips = torch.empty((R, n, 3), dtype=torch.float64, device=dev).uniform_(-1, 1)
    z = lambda *sh_: torch.zeros(sh_, dtype=torch.float64, device=dev)
    oa, osa, oh, osh = z(R, n, 3), z(R, n, 3, 3), z(R, n, 3), z(R, n, 3, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ier = E.triatargiter_multi_dev(args.iprec, 1.0, 1.0, d_tri, d_nrm, d_src, None, slips, d_src, None, None,
                                   oa, osa, wa)
    ier2 = E.triatargiter_multi_dev(args.iprec, 1.0, 1.0, d_tri, d_nrm, d_src, None, slips, d_src, None, None,
                                    oh, osh, wb, half=True)
    torch.cuda.synchronize()
    t_multi = time.perf_counter() - t0
    assert ier == 0 and ier2 == 0, (ier, ier2)
    # the same sets one replay at a time (a sample of them: the median time
    # per set, times the number of sets)
    ns = min(R, 8)
    per = []
    for r in range(ns):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert arrival(3, slips[r].contiguous())[0] == 0
        assert image(3, slips[r].contiguous())[0] == 0
        torch.cuda.synchronize()
        per.append(time.perf_counter() - t0)
        if r == 0:
            e = max(float(torch.linalg.norm(a - b[0]) / torch.linalg.norm(a))
                    for a, b in ((pa, oa), (sa, osa), (ph, oh), (sh, osh)))
    t_single = float(np.median(per)) * R
    return {"rhs": R, "multi_s": t_multi, "single_calls_s": t_single, "speedup": t_single / t_multi,
            "value": R * n / t_multi / 1e6, "unit": "M tris x rhs / s", "set0_vs_single_rel": e,
            "single_sampled_sets": ns}


def c4_batched(args):
    """C4's second half: args.c4_meshes independent icospheres (triangle
    counts log-uniform in 500-20,000, radii 0.5-3, random centres; SL + DL
    densities U[-1,1]^3; displacement and strain at the centroids; iprec
    2) as one batch (tuned schedule, graph replay), against the plain path
    (one elfmm3dtriatarg_dev call per mesh, back to back). Accuracy: a few
    meshes against the direct sum."""
    import torch
    import elfmmlib3d_amd as E
    from elfmmlib3d_amd import mesh
    ws = mesh.sphere_batch(args.c4_meshes, 500, 20000, 2025)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    meshes = [(T(w[0]), T(w[1]), T(w[2]), None) for w in ws]
    ssl = T(np.concatenate([w[3] for w in ws]))
    sdl = T(np.concatenate([w[4] for w in ws]))
    counts = [w[0].shape[0] for w in ws]
    N = sum(counts)
    t0 = time.perf_counter()
    B = E.MeshBatch(args.iprec, 1.0, 1.0, meshes, ifptfrctarg=False, ifstraintarg=False, width=0)
    t_create = time.perf_counter() - t0
    width, graph, tune = B.schedule()
    outs = B.apply(ssl, sdl)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        outs = B.apply(ssl, sdl)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    # plain path: one call per mesh
    z = lambda *sh_: torch.empty(sh_, dtype=torch.float64, device=dev)
    po, so = z(N, 3), z(N, 3, 3)
    sl, dl = ssl.split(counts), sdl.split(counts)
    pos, sos = po.split(counts), so.split(counts)

    def plain():
        for k, (tri, nrm, src, _) in enumerate(meshes):
            ier = E.elfmm3dtriatarg_dev(args.iprec, 1.0, 1.0, tri, nrm, src, sl[k], dl[k], None, pos[k], sos[k],
                                        None, None)
            assert ier == 0
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    same = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in ((outs[0], po), (outs[1], so)))
    # accuracy: the largest, the smallest and two random meshes vs the direct sum
    order = np.argsort(counts)
    pick = sorted({int(order[0]), int(order[-1]), int(order[len(order) // 3]), int(order[2 * len(order) // 3])})
    off = np.concatenate([[0], np.cumsum(counts)])
    pt, st = outs[0].cpu().numpy(), outs[1].cpu().numpy()
    err = 0.0
    for k in pick:
        tri, nrm, src, s1, s2 = ws[k]
        dp, ds_, _, _ = E.el3dtriadirecttarg(1.0, 1.0, tri, nrm, src, s1, s2, np.zeros((0, 3)), 1, 1, 0, 0)
        a, b = pt[off[k]:off[k + 1]], st[off[k]:off[k + 1]]
        err = max(err, float(np.linalg.norm(a - dp) / np.linalg.norm(dp)))
    B.close()
    return {"workload": f"C4 meshes: {len(ws)} icospheres, {N} triangles ({min(counts)}-{max(counts)} each), "
                        f"SL+DL, centroids, iprec {args.iprec}",
            "value": N / t_batch / 1e6, "unit": "M tris/s", "apply_ms": t_batch * 1e3,
            "plain_ms": t_plain * 1e3, "speedup_vs_plain": t_plain / t_batch, "create_s": t_create,
            "schedule": {"width": width, "graph": graph,
                         "tuned": [{"graph": g, "width": w, "ms": ms} for g, w, ms in tune]},
            "batch_vs_plain_rel": same, "disp_err_vs_direct": err, "checked_meshes": pick}


def c4_forest(args):
    """The C4 meshes of c4_batched under the forest schedule only (direct
    launches): one warm-up and three timed applies, for kernel profiles."""
    import torch
    import elfmmlib3d_amd as E
    from elfmmlib3d_amd import mesh
    ws = mesh.sphere_batch(args.c4_meshes, 500, 20000, 2025)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    meshes = [(T(w[0]), T(w[1]), T(w[2]), None) for w in ws]
    ssl = T(np.concatenate([w[3] for w in ws]))
    sdl = T(np.concatenate([w[4] for w in ws]))
    B = E.MeshBatch(args.iprec, 1.0, 1.0, meshes, ifptfrctarg=False, ifstraintarg=False, width=-100)
    B.apply(ssl, sdl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        B.apply(ssl, sdl)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / 3 * 1e3
    B.close()
    return {"workload": f"C4 meshes as one forest: {len(ws)} icospheres, {sum(w[0].shape[0] for w in ws)} triangles",
            "apply_ms": ms}


def points_fmm(args):
    """Point sources (elfmm3dparttarg): args.part_n points uniform in the unit
    cube, SL + DL with random unit dipole orientations, displacement and
    strain at the points, iprec 2 (400-point leaves), lam=mu=1; median of 4
    applies after one warm-up. Also returns a 100,000-point sample of the
    same kind with its GPU result, which the CPU-baseline leg runs through
    the reference."""
    import torch
    import elfmmlib3d_amd as E
    rng = np.random.default_rng(5)
    dev = torch.device("cuda:0")

    def draw(n):
        src = rng.uniform(0, 1, (n, 3))
        sdv = rng.normal(size=(n, 3))
        sdv /= np.linalg.norm(sdv, axis=1, keepdims=True)
        return src, rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3)), sdv

    def gpu_apply(arrs, reps):
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        d = [T(a) for a in arrs]
        n = arrs[0].shape[0]
        out = [torch.empty((n, 3), dtype=torch.float64, device=dev),
               torch.empty((n, 3, 3), dtype=torch.float64, device=dev)]
        ts = []
        for i in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ier = E.elfmm3dparttarg_dev(2, 1.0, 1.0, d[0], d[1], d[2], d[3], None, out[0], out[1], None, None)
            torch.cuda.synchronize()
            if ier != 0:
                raise RuntimeError(f"elfmm3dparttarg ier={ier}")
            if i:
                ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), out[0].cpu().numpy()

    n = args.part_n
    t, _ = gpu_apply(draw(n), 4)
    stats = E.last_stats()
    res = {"workload": f"{n} points uniform in the unit cube, SL + DL, displacement and strain at the points, "
                       "iprec 2, lam=mu=1", "value": n / t / 1e6, "unit": "M points/s", "apply_ms": 1e3 * t,
           "near_pairs": stats["near_pairs"], "boxes": stats["nboxes"]}
    sample = draw(100000)
    ts, pg = gpu_apply(sample, 2)
    res["sample_100k_gpu_ms"] = 1e3 * ts
    return res, (sample, pg)


def c5_served(args):
    """Config C5 on one GPU: a Poisson stream of requests at 0.5x, 1x and 2x
    the measured capacity; 60% half-space fault meshes (8 faults, targets =
    centroids) and 40% free-space spheres (centroids + ntri/2 cube targets),
    triangle counts log-uniform in [1k, args.c5_max]; 30% of the requests
    reuse one of three resident geometries with new densities (iter mode:
    its near fields stored before the stream, replayed by the request). A
    pool of geometries is generated beforehand (request = geometry + fresh
    densities; a non-resident request plans from scratch). Latency = result
    time - arrival time. Returns the per-rate figures."""
    import torch
    from elfmmlib3d_amd import mesh
    from elfmmlib3d_amd.serve import Request, Server
    rng = np.random.default_rng(11)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    def geometry(kind, ntri, seed):
        g = np.random.default_rng(seed)
        if kind == "half":
            tri = mesh.fault_mesh(8, ntri, g)
            tg = mesh.centroids(tri)
        else:
            k = max(1, int(round(np.sqrt(ntri / 20.0))))
            tri = mesh.icosphere(k, radius=g.uniform(0.5, 3.0))
            tg = mesh.cube_targets(max(1, tri.shape[0] // 2), 4.0, 3.1, g)
        return {"kind": kind, "tri": T(tri), "nrm": T(mesh.triangle_normals(tri)), "src": T(mesh.centroids(tri)),
                "tg": T(tg), "n": tri.shape[0], "m": tg.shape[0]}

    lo, hi = np.log(1000.0), np.log(float(args.c5_max))
    pool = [geometry("half" if i % 5 < 3 else "sphere", int(np.exp(rng.uniform(lo, hi))), 100 + i)
            for i in range(10)]
    resident = [geometry("half", 20000, 901), geometry("sphere", 40000, 902), geometry("half", 60000, 903)]

    def request(geo, icomp=None, wsave=None):
        n = geo["n"]
        dl = torch.from_numpy(rng.uniform(-1, 1, (n, 3))).to(dev)
        sl = None if geo["kind"] == "half" else torch.from_numpy(rng.uniform(-1, 1, (n, 3))).to(dev)
        return Request(args.iprec, 1.0, 1.0, geo["tri"], geo["nrm"], geo["src"], sl, dl, geo["tg"],
                       ifptfrc=0 if geo["kind"] == "half" else 1, ifstrain=0 if geo["kind"] == "half" else 1,
                       icomp_type=icomp, wsave=wsave, half=geo["kind"] == "half")

    srv = Server(workers=args.workers)
    wss = [np.zeros(8) for _ in resident]
    for geo, ws in zip(resident, wss):
        r = srv.submit(request(geo, 2, ws)).result()
        assert r.ier == 0, r.ier

    # one sequence of (geometry, resident slot) for every phase, fresh
    # densities per request
    seq = []
    for _ in range(args.c5_requests):
        if rng.uniform() < 0.3:
            k = int(rng.integers(len(resident)))
            seq.append((resident[k], k))
        else:
            seq.append((pool[int(rng.integers(len(pool)))], -1))

    def requests():
        return [(request(geo, 3, wss[k]) if k >= 0 else request(geo), geo) for geo, k in seq]

    def points(geo):
        return geo["n"] + geo["m"] if geo["kind"] == "sphere" else geo["m"]

    # capacity: the sequence submitted at once (the server keeps `workers` in flight)
    reqs = requests()
    t0 = time.perf_counter()
    futs = [srv.submit(r) for r, _ in reqs]
    res = [f.result() for f in futs]
    cap_s = time.perf_counter() - t0
    assert all(r.ier in (0, 64) for r in res), [r.ier for r in res]
    capacity = len(reqs) / cap_s
    out = {"capacity_req_per_s": capacity, "requests_per_rate": args.c5_requests,
           "pool_triangles": sorted(int(g["n"]) for g in pool),
           "resident_triangles": [int(g["n"]) for g in resident]}
    rates = []
    for mult in (0.5, 1.0, 2.0):
        rate = capacity * mult
        reqs = requests()
        gaps = rng.exponential(1.0 / rate, len(reqs))
        t_start = time.perf_counter()
        futs = []
        t_next = t_start
        for (r, geo), gap in zip(reqs, gaps):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            futs.append(srv.submit(r))
        pts = 0
        for (r, geo), f in zip(reqs, futs):
            res = f.result()
            assert res.ier in (0, 64), res.ier
            pts += points(geo)
        t_end = time.perf_counter()
        # latency: submission (at the arrival time) to result, as the server measured it
        lat_srv = np.array([f.result().latency_ms for f in futs])
        rates.append({"offered_req_per_s": rate, "x_capacity": mult,
                      "throughput_M_points_per_s": pts / (t_end - t_start) / 1e6,
                      "completed_req_per_s": len(reqs) / (t_end - t_start),
                      "latency_ms": {"p50": float(np.percentile(lat_srv, 50)),
                                     "p99": float(np.percentile(lat_srv, 99))}})
    srv.close()
    from elfmmlib3d_amd import _lib
    for ws in wss:
        _lib.iter_release(ws[0:4])
        _lib.iter_release(ws[4:8])
    out["rates"] = rates
    return out


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
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the rocprofv3 counter passes that measure each kernel's HBM traffic")
    ap.add_argument("--kernel-pass", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-c5", dest="c5", action="store_false", help="skip the C5 served-stream measurement")
    ap.add_argument("--no-c3", dest="c3", action="store_false", help="skip the C3 half-space measurement")
    ap.add_argument("--c3-tri", type=int, default=200000, help="triangles of the C3 fault mesh")
    ap.add_argument("--c3-replays", type=int, default=3, help="timed C3 applies (replays)")
    ap.add_argument("--c3-only", action="store_true", help="run only the C3 measurement (for profiling)")
    ap.add_argument("--c5-requests", type=int, default=30, help="requests per offered rate in C5")
    ap.add_argument("--c3-rhs", type=int, default=64, help="slip vectors of the C3 multi-set call (0: skip)")
    ap.add_argument("--no-c4", dest="c4", action="store_false", help="skip the C4 many-mesh measurement")
    ap.add_argument("--c4-meshes", type=int, default=256, help="meshes of the C4 batch")
    ap.add_argument("--c4-only", action="store_true", help="run only the C4 many-mesh measurement")
    ap.add_argument("--c4-forest", action="store_true",
                    help="C4 meshes as the forest schedule only, three applies (for profiling)")
    ap.add_argument("--c5-max", type=int, default=500000, help="largest triangle count of a C5 request")
    ap.add_argument("--part-n", type=int, default=1000000, help="points of the point-source FMM measurement")
    ap.add_argument("--no-part", dest="part", action="store_false", help="skip the point-source FMM measurement")
    ap.add_argument("--part-only", action="store_true", help="run only the point-source FMM measurement")
    args = ap.parse_args()
    if args.kernel_pass:
        kernel_pass(args)
        return
    if args.c3_only:
        print(json.dumps(c3_half(args)), flush=True)
        return
    if args.c4_only:
        print(json.dumps(c4_batched(args)), flush=True)
        return
    if args.c4_forest:
        print(json.dumps(c4_forest(args)), flush=True)
        return
    if args.part_only:
        print(json.dumps(points_fmm(args)[0]), flush=True)
        return
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
    traffic, factors = measure_traffic(args) if args.traffic else (None, "not measured (--no-traffic)")
    if traffic:
        for r in list(rooflines.values()) + [roof]:
            if r.get("kernel") in traffic:
                r["traffic"] = traffic[r["kernel"]]

    c3 = c3_half(args) if args.c3 else None
    c4 = c4_batched(args) if args.c4 else None
    c5 = c5_served(args) if args.c5 else None
    pts, pts_sample = points_fmm(args) if args.part else (None, None)
    cpu = None if args.no_cpu else cpu_baseline(pts_sample)
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
        "half_space_c3": c3,
        "batched_c4": c4,
        "served_c5": c5,
        "points_fmm": pts,
        "traffic": ({"bytes_per_dispatch": traffic, "counter_scale": factors,
                     "method": f"rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over a child run of two applies; "
                               f"KiB per dispatch scaled by the counters of a {CALIB_MIB} MiB device copy"}
                    if traffic else {"error": factors}),
    }
    if cpu:
        line["speedup_vs_cpu_reference"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
