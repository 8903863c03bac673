"""Batch of meshes (elfmm_gfx950_batch_*): one apply of many independent
free-space meshes, as one FMM over the forest of their octrees or as one
apply per mesh over branch streams, launched directly or as the replay of a
graph built for the batch, must give each mesh's elfmm3dtriatarg. Meshes of uneven sizes (icospheres of 20
to ~3,000 triangles, some with targets and some without) cross the tuner's
schedules; each graph replay is checked against the single-mesh applies,
with other device work (single-mesh applies, torch kernels) between the
replays."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

TOL = 1e-12      # same plans and kernels; only the order of atomic sums differs


def workload(nmesh=9, seed=31):
    from elfmmlib3d_amd import mesh
    rng = np.random.default_rng(seed + 1)
    meshes = mesh.sphere_batch(nmesh, 20, 3000, seed, box=6.0)
    out = []
    for k, (tri, nrm, src, ssl, sdl) in enumerate(meshes):
        tg = None
        if k % 3 != 1:
            c = src.mean(axis=0)
            r = np.linalg.norm(src - c, axis=1).max()
            tg = mesh.shell_targets(37 * (k + 1), 1.2 * r, 2.0 * r, rng) + c
        out.append((tri, nrm, src, ssl, sdl, tg))
    return out


def torch_meshes(ws):
    import torch
    T = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    return [(T(w[0]), T(w[1]), T(w[2]), T(w[5])) for w in ws]


def singles(E, iprec, ws, ssl, sdl):
    """Each mesh through the single-mesh entry point (packed densities)."""
    res, o = [], 0
    for (tri, nrm, src, _, _, tg) in ws:
        n = tri.shape[0]
        ier, pt, st, ptt, stt = E.elfmm3dtriatarg(iprec, 1.0, 1.0, tri, nrm, src, ssl[o:o + n], sdl[o:o + n],
                                                  tg if tg is not None else np.zeros((0, 3)))
        assert ier == 0
        res.append((pt, st, ptt, stt))
        o += n
    return res


def check(B, outs, ref, ws):
    pt, st, ptt, stt = [None if a is None else a.cpu().numpy() for a in outs]
    n = np.cumsum([0] + [w[0].shape[0] for w in ws])
    m = np.cumsum([0] + [0 if w[5] is None else w[5].shape[0] for w in ws])
    for k, r in enumerate(ref):
        assert rel(pt[n[k]:n[k + 1]], r[0]) < TOL, (k, "ptfrc")
        assert rel(st[n[k]:n[k + 1]].reshape(-1, 3, 3), r[1]) < TOL, (k, "strain")
        if ws[k][5] is not None:
            assert rel(ptt[m[k]:m[k + 1]], r[2]) < TOL, (k, "ptfrctarg")
            assert rel(stt[m[k]:m[k + 1]].reshape(-1, 3, 3), r[3]) < TOL, (k, "straintarg")


@pytest.fixture(scope="module")
def case(gpu_lib):
    ws = workload()
    rng = np.random.default_rng(5)
    N = sum(w[0].shape[0] for w in ws)
    dens = [(np.concatenate([w[3] for w in ws]), np.concatenate([w[4] for w in ws])),
            (rng.uniform(-1, 1, (N, 3)), rng.uniform(-1, 1, (N, 3)))]
    print(f"batch of {len(ws)} meshes: {[w[0].shape[0] for w in ws]} triangles", flush=True)
    ref = {(ip, d): singles(gpu_lib, ip, ws, *dens[d]) for ip in (1, 2) for d in (0, 1)}
    return ws, dens, ref


@pytest.mark.parametrize("width", [-3, -100])
def test_direct_launches_match_single_meshes(gpu_lib, case, width):
    """Branches (-3) and the forest (-100: all meshes' octrees as one FMM)."""
    import torch
    ws, dens, ref = case
    B = gpu_lib.MeshBatch(2, 1.0, 1.0, torch_meshes(ws), width=width)
    assert B.schedule()[:2] == (3 if width == -3 else 0, 0)
    T = lambda a: torch.as_tensor(a, device="cuda")
    for d in (0, 1):
        outs = B.apply(T(dens[d][0]), T(dens[d][1]))
        torch.cuda.synchronize()
        check(B, outs, ref[(2, d)], ws)
    B.close()


@pytest.mark.parametrize("width", [1, 4, 100])
def test_graph_replays_between_other_work(gpu_lib, case, width):
    import torch
    ws, dens, ref = case
    B = gpu_lib.MeshBatch(2, 1.0, 1.0, torch_meshes(ws), width=width)
    w, graph, _ = B.schedule()
    assert (w, graph) == (0 if width == 100 else min(width, len(ws)), 1)
    T = lambda a: torch.as_tensor(a, device="cuda")
    a = torch.randn(512, 512, dtype=torch.float64, device="cuda")
    for rep in range(4):
        d = rep % 2
        outs = B.apply(T(dens[d][0]), T(dens[d][1]))
        torch.cuda.synchronize()
        check(B, outs, ref[(2, d)], ws)
        # eager work between replays: torch kernels and a single-mesh apply
        a = (a @ a.T).sin_()
        one = singles(gpu_lib, 2, ws[:1], dens[d][0][:ws[0][0].shape[0]], dens[d][1][:ws[0][0].shape[0]])
        assert rel(one[0][0], ref[(2, d)][0][0]) < TOL
    B.close()


@pytest.mark.parametrize("iprec", [1, 2])
def test_tuned_batch_matches_single_meshes(gpu_lib, case, iprec):
    import torch
    ws, dens, ref = case
    B = gpu_lib.MeshBatch(iprec, 1.0, 1.0, torch_meshes(ws), width=0)
    width, graph, tune = B.schedule()
    assert width in (0, 1, 2, 4, 8, 16) and graph in (0, 1)
    assert len(tune) == 2 * (1 + len([w for w in (1, 2, 4, 8, 16) if w == 1 or w <= len(ws)]))
    T = lambda a: torch.as_tensor(a, device="cuda")
    for d in (0, 1, 0):
        outs = B.apply(T(dens[d][0]), T(dens[d][1]))
        torch.cuda.synchronize()
        check(B, outs, ref[(iprec, d)], ws)
    B.close()


def test_batch_without_targets_or_single_layer(gpu_lib, case):
    import torch
    ws, dens, _ = case
    ws1 = [(w[0], w[1], w[2], w[3], w[4], None) for w in ws[:4]]
    n = sum(w[0].shape[0] for w in ws1)
    ssl = dens[1][0][:n]
    res, o = [], 0
    for (tri, nrm, src, _, _, _) in ws1:
        k = tri.shape[0]
        ier, pt, st, _, _ = gpu_lib.elfmm3dtriatarg(2, 1.0, 1.0, tri, nrm, src, ssl[o:o + k], None,
                                                    np.zeros((0, 3)))
        assert ier == 0
        res.append((pt, st))
        o += k
    B = gpu_lib.MeshBatch(2, 1.0, 1.0, torch_meshes(ws1), ifdouble=False, width=2)
    pt, st, ptt, stt = B.apply(torch.as_tensor(ssl, device="cuda"), None)
    torch.cuda.synchronize()
    assert ptt is None and stt is None
    pt, st = pt.cpu().numpy(), st.cpu().numpy()
    o = 0
    for k, (tri, *_rest) in enumerate(ws1):
        c = tri.shape[0]
        assert rel(pt[o:o + c], res[k][0]) < TOL and rel(st[o:o + c], res[k][1]) < TOL, k
        o += c
    B.close()
