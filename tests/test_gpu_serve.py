"""Serving: requests of mixed sizes run concurrently (several host threads,
one stream each) give the same results as each request run alone."""
import numpy as np
import pytest

from conftest import load_golden, rel
from elfmmlib3d_amd import mesh

pytestmark = pytest.mark.gpu


def requests():
    from elfmmlib3d_amd.serve import Request
    out = []
    for case in ("fmm_s6", "fmm_c1", "fmm_fault", "fmm_c1_p1"):
        g = load_golden(case)
        out.append(Request(int(g["iprec"]), float(g["lam"]), float(g["mu"]), g["triangle"], g["trinorm"],
                           g["source"], g["sigma_sl"] if g["sigma_sl"].size else None,
                           g["sigma_dl"] if g["sigma_dl"].size else None, g["target"]))
    rng = np.random.default_rng(5)
    tri = mesh.icosphere(20, radius=1.2)
    out.append(Request(2, 1.0, 0.6, tri, mesh.triangle_normals(tri), mesh.centroids(tri),
                       rng.uniform(-1, 1, (tri.shape[0], 3)), rng.uniform(-1, 1, (tri.shape[0], 3)),
                       mesh.cube_targets(5000, 2.5, 1.2, rng)))
    return out


def test_results_same_alone_and_in_the_mix(gpu_lib):
    from elfmmlib3d_amd.serve import Server
    reqs = requests()
    print("alone", flush=True)
    srv = Server(workers=1)
    alone = [srv.submit(r).result() for r in reqs]
    srv.close()
    print("mixed", flush=True)
    srv = Server(workers=3)
    futs = [srv.submit(r) for r in reqs * 3]
    mixed = [f.result() for f in futs]
    srv.close()
    for i, m in enumerate(mixed):
        a = alone[i % len(reqs)]
        assert a.ier == 0 and m.ier == 0
        for x, y in zip(m.numpy(), a.numpy()):
            if x is not None and x.size:
                assert rel(x, y) < 1e-13
    # and the alone results are the reference's (fixture) within 1e-10
    g = load_golden("fmm_c1")
    pt, st, ptt, stt = alone[1].numpy()
    assert rel(pt, g["ptfrc"]) < 1e-10 and rel(stt, g["straintarg"]) < 1e-10


def test_iter_and_half_space_requests(gpu_lib):
    """Iter requests on one wsave (a store, then replays with new densities
    submitted together: they run one at a time) and half-space requests,
    beside plain requests; each equals the same call made alone, and the
    store is released afterwards. An icomp-2 request without a wsave keeps
    no store."""
    from elfmmlib3d_amd.serve import Request, Server
    E = gpu_lib
    g = load_golden("iter_c1_p1")
    lam, mu = float(g["lam"]), float(g["mu"])
    base = (int(g["iprec"]), lam, mu, g["triangle"], g["trinorm"], g["source"])
    rng = np.random.default_rng(8)
    n = g["triangle"].shape[0]
    dens = [(rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))) for _ in range(3)]
    ws = np.zeros(4)
    h = load_golden("hs_fault_p2")
    hs = (int(h["iprec"]), float(h["lam"]), float(h["mu"]), h["triangle"], h["trinorm"], h["source"],
          h["sigma_sl"], h["sigma_dl"], h["target"])
    wh = np.zeros(8)
    srv = Server(workers=3)
    futs = [srv.submit(Request(*base, g["sigma_sl"], g["sigma_dl"], g["target"], icomp_type=2, wsave=ws))]
    futs += [srv.submit(Request(*base, a, b, g["target"], icomp_type=3, wsave=ws)) for a, b in dens]
    futs += [srv.submit(r) for r in requests()[:2]]
    futs.append(srv.submit(Request(*base, g["sigma_sl"], g["sigma_dl"], g["target"], icomp_type=2)))
    futs.append(srv.submit(Request(*hs[:6], hs[6], hs[7], hs[8], icomp_type=2, wsave=wh, half=True)))
    futs.append(srv.submit(Request(*hs[:6], hs[6], hs[7], hs[8], half=True)))
    res = [f.result() for f in futs]
    srv.close()
    assert all(r.ier == 0 for r in res), [r.ier for r in res]
    for r, (a, b) in zip(res[1:4], dens):
        _, _, pt, st, ptt, stt = E.elfmm3dtriatargiter(*base, a, b, g["target"], icomp_type=1)
        for x, y in zip(r.numpy(), (pt, st, ptt, stt)):
            assert rel(x, y) < 1e-12
    assert E.iter_release(ws) == 1
    assert res[6].wsave is None
    # half space: free-space iter apply + image part at the targets
    _, _, _, _, pa, sa = E.elfmm3dtriatargiter(*hs[:8], hs[8], ifptfrc=0, ifstrain=0, icomp_type=1)
    for r, want in ((res[7], (h["anti_pt"] + h["mind_pt"] + h["near2_pt"], h["anti_st"] + h["mind_st"] +
                                h["near2_st"])), (res[8], (h["full_pt"], h["full_st"]))):
        _, _, ptt, stt = r.numpy()
        assert rel(ptt - pa, want[0]) < 1e-10 and rel(stt - sa, want[1]) < 1e-10
    assert E.iter_release(wh[0:4]) == 1 and E.iter_release(wh[4:8]) == 1
