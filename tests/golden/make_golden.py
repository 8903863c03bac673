"""Write the golden fixtures from the reference compiled from its sources.

Run in a container that has the reference tree (make -C oracle ref first):
    python tests/golden/make_golden.py [case ...]
(no arguments: every case).
Each fixture holds its inputs and the reference's outputs, so the tests
never need the reference itself. Cases (all FP64):
  fmm_s6        icosphere 6x6/face (720 tris) + 300 shell targets, iprec 2,
                lam=1.3 mu=0.9, single+double layer, all outputs
  fmm_c1        config C1: icosphere 12x12/face (2880 tris) + 1000 targets in
                1.5<r<2, lam=mu=1, iprec 2
  fmm_c1_p1     C1 geometry at iprec 1
  fmm_sl_only   fmm_s6 geometry, single layer only, targets only outputs
  fmm_dl_only   fmm_s6 geometry, double layer only, centroid outputs only
  fmm_tiny      one-level tree (40 tris, all near field)
  fmm_s6_p3     fmm_s6 geometry at iprec 3 (p = 39)
  fmm_s6_m1     fmm_s6 geometry at iprec -1
  fmm_fault     4 planar faults (~3000 right triangles, open surfaces) with
                1000 targets around them, lam=1, mu=0.8, iprec 2
  iter_c1_p2    elfmm3dtriatargiter (half-space link set, d3tstrcr_mindlin
                root) on the C1 sphere moved to centre (0,0,-2) with 1000
                shell targets, lam=1.2 mu=0.7, iprec 2: icomp 4 (lused),
                icomp 1 outputs for two density sets (sigma, sigma2)
  iter_c1_p1    the same at iprec 1
  pairs         500 random (target, triangle) pairs incl. self and
                near-coplanar targets: elust3/eltst3triadirecttarg_one
  tables        random plane-triangle tables (triabtable / triartable)
  tree_c1       boxes, permutations and lists of d3tstrcr on C1 points
  direct_s4     el3dtriadirecttarg on a 320-triangle sphere + 50 targets
  hs_fault_p2   half space: 4 planar faults in z<0 (~2000 right triangles,
                top depths from 0.05), targets = the centroids + 150 points
                around the faults (z<0), lam=1, mu=0.8, iprec 2, single +
                double layer: elhfmm3dtriatargiter icomp 1 (lwsave >= lused)
                and icomp 4 (lused); each part alone: image near field
                (elhfmm3dtriatargiter0, icomp 1 and 2, the second also for a
                second density set), anti-image far field (elxfmm3dtriatarg,
                subtracted), Mindlin far field (lfmm3dmindlintriatarg);
                elh3dtriadirecttarg and the free-space iter FMM (direct
                arrival) at 24 of the targets
  hs_fault_p1   the same geometry at iprec 1
  part_s        point sources (src/elfmm3dpart.f): 2,400 points in the unit
                cube plus a dense cluster of 1,200, 500 targets in [-0.5,1.5]^3,
                lam=1.2 mu=0.8, iprec 2, single + double layer (random
                dipole orientations), all outputs: elfmm3dparttarg
  part_p1       the same points at iprec 1, double layer only, targets only
  part_m1       the same points at iprec -1, single layer only, sources only
  part_direct   el3dpartdirecttarg on 500 of the points + 80 targets
  part_tiny     elfmm3dparttarg with 2 points + 3 targets and with 1 point
                (iprec 2, single + double layer)
  hpart_direct  elh3dpartdirecttarg (half space) on 400 points in z < 0
                (depths 0.02-2) + 60 targets in z < 0, single + double layer
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from elfmmlib3d_amd import mesh  # noqa: E402
from oracle import ref  # noqa: E402


def sphere_case(nsplit, ntarg, seed, rmin=1.5, rmax=2.0):
    rng = np.random.default_rng(seed)
    tri = mesh.icosphere(nsplit)
    nrm = mesh.triangle_normals(tri)
    src = mesh.centroids(tri)
    ssl = rng.uniform(-1, 1, (tri.shape[0], 3))
    sdl = rng.uniform(-1, 1, (tri.shape[0], 3))
    tg = mesh.shell_targets(ntarg, rmin, rmax, rng)
    return tri, nrm, src, ssl, sdl, tg


def save(name, **kw):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **kw)
    print("wrote", path, os.path.getsize(path))


def fmm_case(name, iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, flags=(1, 1, 1, 1)):
    ier, pt, st, ptt, stt = ref.elfmm3dtriatarg(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, *flags)
    assert ier == 0
    save(name, iprec=iprec, lam=lam, mu=mu, triangle=tri, trinorm=nrm, source=src,
         sigma_sl=ssl if ssl is not None else np.zeros(0), sigma_dl=sdl if sdl is not None else np.zeros(0),
         target=tg, flags=np.array(flags), ptfrc=pt, strain=st, ptfrctarg=ptt, straintarg=stt)


def fault_case(seed=13, ntri=3000, ntarg=1000):
    rng = np.random.default_rng(seed)
    tri = mesh.fault_mesh(4, ntri, rng, box=8.0)
    nrm = mesh.triangle_normals(tri)
    src = mesh.centroids(tri)
    ssl = rng.uniform(-1, 1, (tri.shape[0], 3))
    sdl = rng.uniform(-1, 1, (tri.shape[0], 3))
    lo, hi = tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)
    tg = rng.uniform(lo - 0.5, hi + 0.5, (ntarg, 3))
    return tri, nrm, src, ssl, sdl, tg


def iter_case(name, iprec, lam=1.2, mu=0.7):
    tri, nrm, src, ssl, sdl, tg = sphere_case(12, 1000, 23)
    shift = np.array([0.0, 0.0, -2.0])
    tri, src, tg = tri + shift, src + shift, tg + shift
    rng = np.random.default_rng(29)
    ssl2, sdl2 = rng.uniform(-1, 1, ssl.shape), rng.uniform(-1, 1, sdl.shape)
    ier, lused, *_ = ref.elfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, icomp_type=4)
    assert ier == 0
    # the reference computes its near field only when lwsave >= lused
    w = np.zeros(lused)
    out = {}
    for tag, (a, b) in (("", (ssl, sdl)), ("2", (ssl2, sdl2))):
        ier, _, pt, st, ptt, stt, _ = ref.elfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, a, b, tg,
                                                              icomp_type=1, wsave=w)
        assert ier == 0
        out.update({"ptfrc" + tag: pt, "strain" + tag: st, "ptfrctarg" + tag: ptt, "straintarg" + tag: stt})
    save(name, iprec=iprec, lam=lam, mu=mu, triangle=tri, trinorm=nrm, source=src, sigma_sl=ssl,
         sigma_dl=sdl, sigma_sl2=ssl2, sigma_dl2=sdl2, target=tg, lused=lused, **out)


def hs_geometry(seed=21, ntri=2000, ntarg=150):
    rng = np.random.default_rng(seed)
    tri = mesh.fault_mesh(4, ntri, rng, box=10.0, top=(0.05, 2.0))
    nrm = mesh.triangle_normals(tri)
    src = mesh.centroids(tri)
    n = tri.shape[0]
    ssl = rng.uniform(-1, 1, (n, 3))
    sdl = rng.uniform(-1, 1, (n, 3))
    ssl2 = rng.uniform(-1, 1, (n, 3))
    sdl2 = rng.uniform(-1, 1, (n, 3))
    pick = rng.integers(0, n, ntarg)
    off = src[pick] + rng.normal(0, 0.4, (ntarg, 3))
    off[:, 2] = -np.abs(off[:, 2]) - 0.01
    tg = np.concatenate([src, off])
    return tri, nrm, src, ssl, sdl, ssl2, sdl2, tg


def hs_case(name, iprec, lam=1.0, mu=0.8):
    tri, nrm, src, ssl, sdl, ssl2, sdl2, tg = hs_geometry()
    out = {}
    ier, lused, _, _, _ = ref.elhfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, icomp_type=4)
    assert ier == 0
    out["lused"] = lused
    ws = np.zeros(lused + 16)
    ier, _, out["full_pt"], out["full_st"], _ = ref.elhfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl,
                                                                         tg, icomp_type=1, wsave=ws)
    assert ier == 0
    del ws
    for ic in (1, 2):
        ier, need, out[f"near{ic}_pt"], out[f"near{ic}_st"] = ref.elhfmm3d_near(iprec, lam, mu, tri, nrm, src, ssl,
                                                                                 sdl, tg, icomp_type=ic)
        assert ier == 0 and need == lused
    _, _, out["near2b_pt"], out["near2b_st"] = ref.elhfmm3d_near(iprec, lam, mu, tri, nrm, src, ssl2, sdl2, tg,
                                                                 icomp_type=2)
    ier, out["anti_pt"], out["anti_st"] = ref.elxfmm3d_anti(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg)
    assert ier == 0
    ier, out["mind_pt"], out["mind_st"] = ref.lfmm3d_mindlin(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg)
    assert ier == 0
    sub = np.linspace(0, tg.shape[0] - 1, 24).astype(int)
    _, _, out["direct_pt"], out["direct_st"] = ref.elh3dtriadirecttarg(lam, mu, tri, nrm, src, ssl, sdl, tg[sub])
    r = ref.elfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg[sub], ifptfrc=0, ifstrain=0,
                                icomp_type=4)
    ws = np.zeros(r[1] + 16)
    r = ref.elfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg[sub], ifptfrc=0, ifstrain=0,
                                icomp_type=1, wsave=ws)
    out["arrival_pt"], out["arrival_st"] = r[4], r[5]
    save(name, iprec=iprec, lam=lam, mu=mu, triangle=tri, trinorm=nrm, source=src, sigma_sl=ssl, sigma_dl=sdl,
         sigma_sl2=ssl2, sigma_dl2=sdl2, target=tg, direct_idx=sub, **out)


def hs_low_case(name, lam=1.1, mu=0.9):
    """The half-space apply (icomp 1) and its Mindlin part at iprec 0 and -1,
    where the Mindlin P2M uses the 6- and 3-node triangle rules."""
    tri, nrm, src, ssl, sdl, _, _, tg = hs_geometry(seed=5, ntri=900, ntarg=80)
    out = {}
    for tag, iprec in (("p0", 0), ("m1", -1)):
        ier, lused, _, _, _ = ref.elhfmm3dtriatargiter(iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, icomp_type=4)
        assert ier == 0
        ws = np.zeros(lused + 16)
        ier, _, out[f"full_{tag}_pt"], out[f"full_{tag}_st"], _ = ref.elhfmm3dtriatargiter(
            iprec, lam, mu, tri, nrm, src, ssl, sdl, tg, icomp_type=1, wsave=ws)
        assert ier == 0
        ier, out[f"mind_{tag}_pt"], out[f"mind_{tag}_st"] = ref.lfmm3d_mindlin(iprec, lam, mu, tri, nrm, src, ssl,
                                                                              sdl, tg)
        assert ier == 0
    save(name, lam=lam, mu=mu, triangle=tri, trinorm=nrm, source=src, sigma_sl=ssl, sigma_dl=sdl, target=tg,
         **out)


def c2_case(name, iprecs=(2, 3), nsample=3000, nproj=6):
    """C2 (icosphere 64x64/face, 81,920 triangles, 100,000 targets, seed
    2024; mesh.sphere_workload) through the reference FMM at each iprec:
    outputs at seeded sample indices (sources and targets) and the full
    outputs' dot products with seeded Gaussian vectors. The inputs are not
    stored: mesh.sphere_workload regenerates them."""
    tri, nrm, src, ssl, sdl, tg = mesh.sphere_workload(64, 100000, 2024)
    n, m = tri.shape[0], tg.shape[0]
    rng = np.random.default_rng(77)
    si = np.sort(rng.choice(n, nsample, replace=False))
    ti = np.sort(rng.choice(m, nsample, replace=False))
    out = dict(nsplit=64, ntarg=100000, seed=2024, lam=1.0, mu=1.0, src_idx=si, tgt_idx=ti,
               proj_seed=78, nproj=nproj)
    for ip in iprecs:
        ier, pt, st, ptt, stt = ref.elfmm3dtriatarg(ip, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
        assert ier == 0
        prng = np.random.default_rng(78)
        for k, a in (("ptfrc", pt), ("strain", st), ("ptfrctarg", ptt), ("straintarg", stt)):
            v = a.reshape(a.shape[0], -1)
            out[f"p{ip}_{k}_sample"] = v[si if k in ("ptfrc", "strain") else ti]
            R = prng.standard_normal((nproj, v.size))
            out[f"p{ip}_{k}_proj"] = R @ v.ravel()
            out[f"p{ip}_{k}_norm"] = np.linalg.norm(v)
        print("c2 iprec", ip, "done", flush=True)
    save(name, **out)


def part_geometry(seed=41, nu=2400, ncl=1200, ntarg=500):
    rng = np.random.default_rng(seed)
    cl = rng.normal(0.0, 0.04, (ncl, 3)) + np.array([0.7, 0.3, 0.6])
    src = np.concatenate([rng.uniform(0, 1, (nu, 3)), cl])
    n = src.shape[0]
    ssl = rng.uniform(-1, 1, (n, 3))
    sdl = rng.uniform(-1, 1, (n, 3))
    sdv = rng.normal(size=(n, 3))
    sdv /= np.linalg.norm(sdv, axis=1, keepdims=True)
    tg = rng.uniform(-0.5, 1.5, (ntarg, 3))
    return src, ssl, sdl, sdv, tg


def part_case(name, iprec, lam, mu, src, ssl, sdl, sdv, tg, flags=(1, 1, 1, 1)):
    ier, pt, st, ptt, stt = ref.elfmm3dparttarg(iprec, lam, mu, src, ssl, sdl, sdv, tg, *flags)
    assert ier == 0
    z = np.zeros(0)
    save(name, iprec=iprec, lam=lam, mu=mu, source=src, sigma_sl=ssl if ssl is not None else z,
         sigma_dl=sdl if sdl is not None else z, sigma_dv=sdv if sdl is not None else z, target=tg,
         flags=np.array(flags), ptfrc=pt, strain=st, ptfrctarg=ptt, straintarg=stt)


def main(only=None):
    want = lambda name: only is None or name in only
    tri, nrm, src, ssl, sdl, tg = sphere_case(6, 300, 11)
    if want("fmm_s6"):
        fmm_case("fmm_s6", 2, 1.3, 0.9, tri, nrm, src, ssl, sdl, tg)
    if want("fmm_sl_only"):
        fmm_case("fmm_sl_only", 2, 1.3, 0.9, tri, nrm, src, ssl, None, tg, (0, 0, 1, 1))
    if want("fmm_dl_only"):
        fmm_case("fmm_dl_only", 2, 1.3, 0.9, tri, nrm, src, None, sdl, np.zeros((0, 3)), (1, 1, 0, 0))
    if want("fmm_s6_p3"):
        fmm_case("fmm_s6_p3", 3, 1.3, 0.9, tri, nrm, src, ssl, sdl, tg)
    if want("fmm_s6_m1"):
        fmm_case("fmm_s6_m1", -1, 1.3, 0.9, tri, nrm, src, ssl, sdl, tg)
    tri, nrm, src, ssl, sdl, tg = sphere_case(12, 1000, 1)
    if want("fmm_c1"):
        fmm_case("fmm_c1", 2, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
    if want("fmm_c1_p1"):
        fmm_case("fmm_c1_p1", 1, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
    if want("fmm_tiny"):
        t2, n2, s2, sl2, dl2, tg2 = sphere_case(2, 10, 5)
        fmm_case("fmm_tiny", 2, 2.0, 0.7, t2[:40], n2[:40], s2[:40], sl2[:40], dl2[:40], tg2)
    if want("hs_fault_p2"):
        hs_case("hs_fault_p2", 2)
    if want("hs_fault_p1"):
        hs_case("hs_fault_p1", 1)
    if only is not None and "c2_ref" in only:
        c2_case("c2_ref")
    if want("hs_fault_low"):
        hs_low_case("hs_fault_low")
    if want("fmm_fault"):
        fmm_case("fmm_fault", 2, 1.0, 0.8, *fault_case())
    for name, iprec in (("iter_c1_p2", 2), ("iter_c1_p1", 1)):
        if want(name):
            iter_case(name, iprec)
    src, ssl, sdl, sdv, tg = part_geometry()
    if want("part_s"):
        part_case("part_s", 2, 1.2, 0.8, src, ssl, sdl, sdv, tg)
    if want("part_p1"):
        part_case("part_p1", 1, 1.2, 0.8, src, None, sdl, sdv, tg, (0, 0, 1, 1))
    if want("part_m1"):
        part_case("part_m1", -1, 1.2, 0.8, src, ssl, None, None, np.zeros((0, 3)), (1, 1, 0, 0))
    if want("part_tiny"):
        rng = np.random.default_rng(47)
        src2, ssl2, sdl2 = rng.uniform(-1, 1, (2, 3)), rng.uniform(-1, 1, (2, 3)), rng.uniform(-1, 1, (2, 3))
        sdv2 = rng.normal(size=(2, 3))
        sdv2 /= np.linalg.norm(sdv2, axis=1, keepdims=True)
        tg3 = rng.uniform(-2, 2, (3, 3))
        ier2, pt2, st2, ptt2, stt2 = ref.elfmm3dparttarg(2, 1.1, 0.9, src2, ssl2, sdl2, sdv2, tg3)
        ier1, pt1, st1, ptt1, stt1 = ref.elfmm3dparttarg(2, 1.1, 0.9, src2[:1], ssl2[:1], sdl2[:1], sdv2[:1], tg3)
        assert ier2 == 0 and ier1 == 0
        save("part_tiny", lam=1.1, mu=0.9, source=src2, sigma_sl=ssl2, sigma_dl=sdl2, sigma_dv=sdv2, target=tg3,
             ptfrc2=pt2, strain2=st2, ptfrctarg2=ptt2, straintarg2=stt2,
             ptfrc1=pt1, strain1=st1, ptfrctarg1=ptt1, straintarg1=stt1)
    if want("hpart_direct"):
        rng = np.random.default_rng(53)
        src = np.column_stack([rng.uniform(-1, 1, (400, 2)), -rng.uniform(0.02, 2.0, 400)])
        tgh = np.column_stack([rng.uniform(-1.5, 1.5, (60, 2)), -rng.uniform(0.01, 2.5, 60)])
        ssl, sdl = rng.uniform(-1, 1, (400, 3)), rng.uniform(-1, 1, (400, 3))
        sdv = rng.normal(size=(400, 3))
        sdv /= np.linalg.norm(sdv, axis=1, keepdims=True)
        pt, st, ptt, stt = ref.elh3dpartdirecttarg(1.3, 0.7, src, ssl, sdl, sdv, tgh)
        save("hpart_direct", lam=1.3, mu=0.7, source=src, sigma_sl=ssl, sigma_dl=sdl, sigma_dv=sdv, target=tgh,
             ptfrc=pt, strain=st, ptfrctarg=ptt, straintarg=stt)
    if want("part_direct"):
        k = np.random.default_rng(43).choice(src.shape[0], 500, replace=False)
        pt, st, ptt, stt = ref.el3dpartdirecttarg(1.2, 0.8, src[k], ssl[k], sdl[k], sdv[k], tg[:80])
        save("part_direct", lam=1.2, mu=0.8, source=src[k], sigma_sl=ssl[k], sigma_dl=sdl[k], sigma_dv=sdv[k],
             target=tg[:80], ptfrc=pt, strain=st, ptfrctarg=ptt, straintarg=stt)
    if only is not None:
        return

    # pair quadratures
    rng = np.random.default_rng(7)
    P = {k: [] for k in ("tri", "sigma", "target", "ifself", "lam", "mu", "pt_sl", "st_sl",
                         "pt_dl", "st_dl")}
    for t in range(500):
        tr = rng.uniform(-1, 1, (3, 3))
        nv = np.cross(tr[1] - tr[0], tr[2] - tr[0])
        nv /= np.linalg.norm(nv)
        sg = rng.uniform(-1, 1, 3)
        kind = t % 4
        ifself = 0
        if kind == 0:
            x = tr.mean(0)
            ifself = 1
        elif kind == 1:
            x = rng.uniform(-2, 2, 3)
        elif kind == 2:
            x = tr.mean(0) + rng.uniform(-1, 1, 3) * 0.7
            x = x - nv * np.dot(x - tr[0], nv) + nv * rng.choice([1e-3, 1e-6, 0.0])
        else:
            x = tr[0] + rng.uniform(-0.05, 0.05, 3)
        lam, mu = rng.uniform(0.5, 3), rng.uniform(0.5, 3)
        a, b = ref.pair_sl(lam, mu, tr, sg, x, ifself)
        c, d = ref.pair_dl(lam, mu, tr, sg, nv, x, ifself)
        for k, v in zip(P, (tr, sg, x, ifself, lam, mu, a, b, c, d)):
            P[k].append(v)
    save("pairs", **{k: np.array(v) for k, v in P.items()})

    # plane-triangle tables
    T = {k: [] for k in ("geom", "iself", "btab", "rtab")}
    for t in range(200):
        a, b, c = rng.uniform(0.5, 2), rng.uniform(-1, 2), rng.uniform(0.3, 2)
        if t % 3 == 0:
            x0, y0, z0, iself = (a + b) / 3, c / 3, 0.0, 1
        else:
            x0, y0 = rng.uniform(-3, 4, 2)
            z0 = rng.uniform(-2, 2) * rng.choice([1, 1e-3])
            iself = 0
        iquad = 0 if iself else (1 if z0 > 0 else (-1 if z0 < 0 else 0))
        v1, v2, v3 = np.array([0.0, 0.0]), np.array([a, 0.0]), np.array([b, c])
        bt = np.zeros((5, 5, 5))
        rt = np.zeros((3, 3, 3))
        for i in range(5):
            for j in range(5 - i):
                for k in range(5 - i - j):
                    if i + j + k >= 2:
                        bt[i, j, k] = ref.triabtable(i, j, k, iquad, v1, v2, v3, x0, y0, z0)
        for i in range(3):
            for j in range(3 - i):
                for k in range(3 - i - j):
                    rt[i, j, k] = ref.triartable(i, j, k, iquad, v1, v2, v3, x0, y0, z0)
        T["geom"].append((a, b, c, x0, y0, z0))
        T["iself"].append(iself)
        T["btab"].append(bt)
        T["rtab"].append(rt)
    save("tables", **{k: np.array(v) for k, v in T.items()})

    # tree and lists
    tri, nrm, src, ssl, sdl, tg = sphere_case(12, 1000, 1)
    tr = ref.tree(src, tg, 40)
    nb = tr["boxes"].shape[0]
    lists = {}
    for t in range(1, 6):
        start = np.zeros(nb + 1, dtype=np.int32)
        start[1:] = np.cumsum([len(x) for x in tr["lists"][t]])
        lists[f"lstart{t}"] = start
        lists[f"litems{t}"] = np.concatenate(tr["lists"][t]).astype(np.int32) if start[-1] else np.zeros(0, np.int32)
    save("tree_c1", source=src, target=tg, nbox=40, boxes=tr["boxes"], centers=tr["centers"],
         iz=tr["iz"], iztarg=tr["iztarg"], laddr=tr["laddr"], nlev=tr["nlev"], center=tr["center"],
         size=tr["size"], **lists)

    # direct sum
    tri, nrm, src, ssl, sdl, tg = sphere_case(4, 50, 3, 0.5, 3.0)
    pt, st, ptt, stt = ref.el3dtriadirecttarg(1.1, 0.8, tri, nrm, src, ssl, sdl, tg)
    save("direct_s4", lam=1.1, mu=0.8, triangle=tri, trinorm=nrm, source=src, sigma_sl=ssl,
         sigma_dl=sdl, target=tg, ptfrc=pt, strain=st, ptfrctarg=ptt, straintarg=stt)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
