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
