"""The full C2 workload (icosphere 64x64 per face: 81,920 triangles,
100,000 targets in [-2,2]^3 minus the unit ball, SL + DL, lam = mu = 1) on
the GPU against the reference FMM compiled from its sources
(tests/golden/c2_ref.npz, make_golden.py c2_ref) at iprec 2 and 3: outputs
at 3,000 sampled sources and targets within 1e-10 relative L2, the whole
output vectors through their norms and their dot products with seeded
Gaussian vectors (a projection error above 1e-9 of the vector's norm would
need an output difference well above 1e-10 relative), and the FMM against
the GPU direct sum at sampled targets within the requested tolerance (at
iprec 3 within 2e-9: the reference FMM's own accuracy at that order on this
workload is 7e-10 against a tolerance of 5e-10)."""
import numpy as np
import pytest

from conftest import load_golden, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    from elfmmlib3d_amd import mesh
    g = load_golden("c2_ref")
    inp = mesh.sphere_workload(int(g["nsplit"]), int(g["ntarg"]), int(g["seed"]))
    return g, inp, {}


def apply(gpu_lib, c2, iprec):
    g, (tri, nrm, src, ssl, sdl, tg), cache = c2
    if iprec not in cache:
        print(f"C2 iprec {iprec}: {tri.shape[0]} triangles, {tg.shape[0]} targets", flush=True)
        cache[iprec] = gpu_lib.elfmm3dtriatarg(iprec, 1.0, 1.0, tri, nrm, src, ssl, sdl, tg)
    return cache[iprec]


@pytest.mark.parametrize("iprec", [2, 3])
def test_c2_matches_reference_fmm(gpu_lib, c2, iprec):
    g = c2[0]
    ier, pt, st, ptt, stt = apply(gpu_lib, c2, iprec)
    assert ier == 0
    prng = np.random.default_rng(int(g["proj_seed"]))
    for k, a in (("ptfrc", pt), ("strain", st), ("ptfrctarg", ptt), ("straintarg", stt)):
        v = a.reshape(a.shape[0], -1)
        idx = g["src_idx"] if k in ("ptfrc", "strain") else g["tgt_idx"]
        e = rel(v[idx], g[f"p{iprec}_{k}_sample"])
        assert e < 1e-10, (k, e)
        nrm_ref = float(g[f"p{iprec}_{k}_norm"])
        assert abs(np.linalg.norm(v) - nrm_ref) < 1e-10 * nrm_ref, k
        R = prng.standard_normal((int(g["nproj"]), v.size))
        dp = np.abs(R @ v.ravel() - g[f"p{iprec}_{k}_proj"]).max()
        assert dp < 1e-9 * nrm_ref, (k, dp / nrm_ref)


@pytest.mark.parametrize("iprec,tol", [(2, 0.5e-6), (3, 2e-9)])
def test_c2_within_tolerance_of_direct_sum(gpu_lib, c2, iprec, tol):
    g, (tri, nrm, src, ssl, sdl, tg), _ = c2
    ier, _, _, ptt, stt = apply(gpu_lib, c2, iprec)
    assert ier == 0
    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(tg.shape[0], 400, replace=False))
    _, _, dpt, dst = gpu_lib.el3dtriadirecttarg(1.0, 1.0, tri, nrm, src, ssl, sdl, tg[idx], 0, 0, 1, 1)
    assert rel(ptt[idx], dpt) < tol and rel(stt[idx], dst) < 10 * tol, (rel(ptt[idx], dpt), rel(stt[idx], dst))
