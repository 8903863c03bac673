"""Parity of the gfx950 library with the reference: FMM outputs against the
reference FMM (golden fixtures, relative L2 <= 1e-10), the direct sum against
the reference and the oracle, and the FMM against the direct sum within the
requested tolerance."""
import numpy as np
import pytest

from conftest import load_golden, rel
from elfmmlib3d_amd import mesh

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def run_golden(gpu_lib, g):
    ssl = g["sigma_sl"] if g["sigma_sl"].size else None
    sdl = g["sigma_dl"] if g["sigma_dl"].size else None
    f = [int(x) for x in g["flags"]]
    return gpu_lib.elfmm3dtriatarg(int(g["iprec"]), float(g["lam"]), float(g["mu"]), g["triangle"],
                                   g["trinorm"], g["source"], ssl, sdl, g["target"], *f)


@pytest.mark.parametrize("case", ["fmm_tiny", "fmm_s6", "fmm_c1", "fmm_c1_p1", "fmm_sl_only", "fmm_dl_only",
                                  "fmm_s6_p3", "fmm_s6_m1", "fmm_fault"])
def test_fmm_matches_reference_fmm(gpu_lib, case):
    g = load_golden(case)
    print(f"case {case}", flush=True)
    ier, pt, st, ptt, stt = run_golden(gpu_lib, g)
    assert ier == 0
    f = [int(x) for x in g["flags"]]
    if f[0]:
        assert rel(pt, g["ptfrc"]) < PARITY
    if f[1]:
        assert rel(st, g["strain"]) < PARITY
    if f[2] and g["target"].shape[0]:
        assert rel(ptt, g["ptfrctarg"]) < PARITY
    if f[3] and g["target"].shape[0]:
        assert rel(stt, g["straintarg"]) < PARITY


def test_direct_matches_reference(gpu_lib):
    g = load_golden("direct_s4")
    out = gpu_lib.el3dtriadirecttarg(float(g["lam"]), float(g["mu"]), g["triangle"], g["trinorm"],
                                     g["source"], g["sigma_sl"], g["sigma_dl"], g["target"])
    for a, k in zip(out, ("ptfrc", "strain", "ptfrctarg", "straintarg")):
        assert rel(a, g[k]) < 1e-11, k


def test_direct_matches_oracle(gpu_lib, oracle_lib):
    rng = np.random.default_rng(21)
    tri = mesh.icosphere(5, radius=1.3)
    nrm, src = mesh.triangle_normals(tri), mesh.centroids(tri)
    ssl, sdl = rng.uniform(-1, 1, (tri.shape[0], 3)), rng.uniform(-1, 1, (tri.shape[0], 3))
    tg = mesh.cube_targets(200, 2.0, 1.0, rng)
    a = gpu_lib.el3dtriadirecttarg(0.9, 1.4, tri, nrm, src, ssl, sdl, tg)
    b = oracle_lib.direct(0.9, 1.4, tri, nrm, src, ssl, sdl, tg)
    for x, y in zip(a, b):
        assert rel(x, y) < 1e-12


@pytest.mark.parametrize("iprec,tol", [(1, 0.5e-3), (2, 0.5e-6)])
def test_fmm_within_tolerance_of_direct(gpu_lib, iprec, tol):
    rng = np.random.default_rng(3)
    tri = mesh.icosphere(16)
    nrm, src = mesh.triangle_normals(tri), mesh.centroids(tri)
    ssl, sdl = rng.uniform(-1, 1, (tri.shape[0], 3)), rng.uniform(-1, 1, (tri.shape[0], 3))
    tg = mesh.cube_targets(2000, 2.0, 1.0, rng)
    ier, *fmm = gpu_lib.elfmm3dtriatarg(iprec, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
    assert ier == 0
    ref = gpu_lib.el3dtriadirecttarg(1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
    # displacement to the requested tolerance, strain one digit looser
    # (the reference's documented accuracy, src/elfmm3dtria.f:122-124)
    assert rel(fmm[0], ref[0]) < tol and rel(fmm[2], ref[2]) < tol
    assert rel(fmm[1], ref[1]) < 10 * tol and rel(fmm[3], ref[3]) < 10 * tol


def test_device_entry_matches_host_entry(gpu_lib):
    import torch
    g = load_golden("fmm_s6")
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    pt = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    st = torch.zeros((n, 3, 3), dtype=torch.float64, device=dev)
    ptt = torch.zeros((m, 3), dtype=torch.float64, device=dev)
    stt = torch.zeros((m, 3, 3), dtype=torch.float64, device=dev)
    ier = gpu_lib.elfmm3dtriatarg_dev(2, float(g["lam"]), float(g["mu"]), T(g["triangle"]), T(g["trinorm"]),
                                      T(g["source"]), T(g["sigma_sl"]), T(g["sigma_dl"]), T(g["target"]),
                                      pt, st, ptt, stt)
    torch.cuda.synchronize()
    assert ier == 0
    assert rel(pt.cpu().numpy(), g["ptfrc"]) < PARITY
    assert rel(stt.cpu().numpy(), g["straintarg"]) < PARITY


def test_repeated_calls_identical(gpu_lib):
    g = load_golden("fmm_s6")
    a = run_golden(gpu_lib, g)
    b = run_golden(gpu_lib, g)
    for x, y in zip(a[1:], b[1:]):
        assert rel(x, y) < 1e-14


@pytest.mark.parametrize("case", ["fmm_c1", "fmm_fault"])
def test_split_plan_matches_one_piece_plan(gpu_lib, case, monkeypatch):
    """A plain apply launches its near field before the far-field lists are
    built (two plan buffers); ELFMM_SPLIT_PLAN=0 plans in one piece first.
    Same lists and kernels: only the order of atomic sums may differ."""
    g = load_golden(case)
    monkeypatch.setenv("ELFMM_SPLIT_PLAN", "0")
    one = run_golden(gpu_lib, g)
    monkeypatch.setenv("ELFMM_SPLIT_PLAN", "1")
    split = run_golden(gpu_lib, g)
    assert one[0] == 0 and split[0] == 0
    for x, y in zip(one[1:], split[1:]):
        if x is not None and np.size(x):
            assert rel(y, x) < 1e-13
    f = [int(x) for x in g["flags"]]
    if f[0]:
        assert rel(split[1], g["ptfrc"]) < PARITY
