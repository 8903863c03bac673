"""Half-space (Mindlin) path on the GPU against the reference compiled from
its sources (fixtures hs_fault_p2 / hs_fault_p1, tests/golden/make_golden.py):
each part alone (image near field of lists 1, 3, 4 and self with the
adaptive image integrals; anti-image far field; Mindlin harmonic far field),
the whole elhfmm3dtriatargiter (icomp 1, the 2 -> 3 store/replay cycle with
new densities, icomp 4), and the direct sum elh3dtriadirecttarg; then the
FMM against the direct sum within the requested tolerance."""
import numpy as np
import pytest

from conftest import load_golden, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


def inputs(g, second=False):
    return (float(g["lam"]), float(g["mu"]), g["triangle"], g["trinorm"], g["source"],
            g["sigma_sl2" if second else "sigma_sl"], g["sigma_dl2" if second else "sigma_dl"], g["target"])


@pytest.fixture(params=["hs_fault_p2", "hs_fault_p1"])
def case(request, gpu_lib):
    g = load_golden(request.param)
    yield gpu_lib, g
    gpu_lib.half_parts(7)


def run(E, g, mask, icomp=1, wsave=None, second=False):
    E.half_parts(mask)
    lam, mu, tri, nrm, src, ssl, sdl, tg = inputs(g, second)
    ier, lused, pt, st = E.elhfmm3dtriatargiter(int(g["iprec"]), lam, mu, tri, nrm, src, ssl, sdl, tg,
                                                icomp_type=icomp, wsave=wsave)
    return ier, lused, pt, st


def test_image_near_field_icomp1(case):
    E, g = case
    ier, _, pt, st = run(E, g, 1)
    assert ier == 0
    assert rel(pt, g["near1_pt"]) < TOL and rel(st, g["near1_st"]) < TOL, (rel(pt, g["near1_pt"]),
                                                                           rel(st, g["near1_st"]))


def test_image_near_field_store_and_replay(case):
    """icomp 5 keeps the unit-density integrals (the reference's icomp 2
    inside), icomp 3 replays them with a second density set; icomp 4 reports
    the reference's store size."""
    E, g = case
    ws = np.zeros(4)
    ier, _, pt, st = run(E, g, 1, icomp=5, wsave=ws)
    assert ier == 0
    assert rel(pt, g["near2_pt"]) < TOL and rel(st, g["near2_st"]) < TOL
    ier, _, pt2, st2 = run(E, g, 1, icomp=3, wsave=ws, second=True)
    assert ier == 0
    assert rel(pt2, g["near2b_pt"]) < TOL and rel(st2, g["near2b_st"]) < TOL
    assert E.iter_release(ws) == 1
    ier, lused, _, _ = run(E, g, 7, icomp=4)
    assert ier == 0 and int(E.last_stats()["iter_ref_store_words"]) == int(g["lused"])


def test_anti_image_far_field(case):
    E, g = case
    ier, _, pt, st = run(E, g, 2)
    assert ier == 0
    assert rel(pt, g["anti_pt"]) < TOL and rel(st, g["anti_st"]) < TOL, (rel(pt, g["anti_pt"]),
                                                                         rel(st, g["anti_st"]))


def test_direct_sum(gpu_lib):
    g = load_golden("hs_fault_p2")
    lam, mu, tri, nrm, src, ssl, sdl, tg = inputs(g)
    sub = g["direct_idx"]
    _, _, pt, st = gpu_lib.elh3dtriadirecttarg(lam, mu, tri, nrm, src, ssl, sdl, tg[sub])
    assert rel(pt, g["direct_pt"]) < 1e-11 and rel(st, g["direct_st"]) < 1e-11
