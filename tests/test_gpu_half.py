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


def test_mindlin_far_field(case):
    E, g = case
    ier, _, pt, st = run(E, g, 4)
    assert ier == 0
    assert rel(pt, g["mind_pt"]) < TOL and rel(st, g["mind_st"]) < TOL, (rel(pt, g["mind_pt"]),
                                                                         rel(st, g["mind_st"]))


def test_half_space_fmm(case):
    """The whole image part: icomp 1, then the store (icomp 2) and replay
    (icomp 3) cycle, against the reference's own parts (its icomp 2 and 3
    near field integrates the unit densities)."""
    E, g = case
    ier, _, pt, st = run(E, g, 7)
    assert ier == 0
    assert rel(pt, g["full_pt"]) < TOL and rel(st, g["full_st"]) < TOL, (rel(pt, g["full_pt"]),
                                                                         rel(st, g["full_st"]))
    want_pt = g["anti_pt"] + g["mind_pt"] + g["near2_pt"]
    want_st = g["anti_st"] + g["mind_st"] + g["near2_st"]
    ws = np.zeros(4)
    for icomp in (2, 3):
        ier, _, pt, st = run(E, g, 7, icomp=icomp, wsave=ws)
        assert ier == 0
        assert rel(pt, want_pt) < TOL and rel(st, want_st) < TOL, (icomp, rel(pt, want_pt), rel(st, want_st))
    assert E.iter_release(ws) == 1


@pytest.mark.parametrize("tag,iprec", [("p0", 0), ("m1", -1)])
def test_half_space_low_precision(gpu_lib, tag, iprec):
    """iprec 0 and -1: the Mindlin P2M on the 6- and 3-node triangle rules,
    smaller trees (nbox 60 and 30) and the lowest plane-wave table."""
    E = gpu_lib
    g = load_golden("hs_fault_low")
    args = (float(g["lam"]), float(g["mu"]), g["triangle"], g["trinorm"], g["source"], g["sigma_sl"],
            g["sigma_dl"], g["target"])
    try:
        E.half_parts(4)
        ier, _, pt, st = E.elhfmm3dtriatargiter(iprec, *args)
        assert ier == 0
        assert rel(pt, g[f"mind_{tag}_pt"]) < TOL and rel(st, g[f"mind_{tag}_st"]) < TOL, (
            rel(pt, g[f"mind_{tag}_pt"]), rel(st, g[f"mind_{tag}_st"]))
    finally:
        E.half_parts(7)
    ier, _, pt, st = E.elhfmm3dtriatargiter(iprec, *args)
    assert ier == 0
    assert rel(pt, g[f"full_{tag}_pt"]) < TOL and rel(st, g[f"full_{tag}_st"]) < TOL, (
        rel(pt, g[f"full_{tag}_pt"]), rel(st, g[f"full_{tag}_st"]))


def test_half_space_fmm_vs_direct(gpu_lib):
    """Direct arrival (elfmm3dtriatargiter) + image part (elhfmm3dtriatargiter)
    against the GPU half-space direct sum, within the requested tolerance
    (iprec 2: 0.5e-6 for displacement, one digit more for strain)."""
    E = gpu_lib
    g = load_golden("hs_fault_p2")
    lam, mu, tri, nrm, src, ssl, sdl, tg = inputs(g)
    sub = g["direct_idx"]
    ier, _, _, _, pa, sa = E.elfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl, sdl, tg[sub], ifptfrc=0,
                                                 ifstrain=0)
    assert ier == 0
    ier, _, ph, sh = E.elhfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl, sdl, tg[sub])
    assert ier == 0
    _, _, pd, sd = E.elh3dtriadirecttarg(lam, mu, tri, nrm, src, ssl, sdl, tg[sub])
    assert rel(pa + ph, pd) < 0.5e-6 and rel(sa + sh, sd) < 0.5e-5, (rel(pa + ph, pd), rel(sa + sh, sd))


def test_half_space_replay_reuses_the_stored_plans(gpu_lib):
    """A replay (icomp 3) with new densities on a store kept by icomp 2 (its
    octrees and plans, including the Mindlin far field's, reused) equals an
    icomp 2 apply with those densities on a fresh store, and a second
    replay repeats it."""
    E = gpu_lib
    g = load_golden("hs_fault_p2")
    lam, mu, tri, nrm, src, ssl, sdl, tg = inputs(g)
    ssl2, sdl2 = inputs(g, second=True)[5:7]
    ws, ws2 = np.zeros(4), np.zeros(4)
    ier, _, _, _ = E.elhfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl, sdl, tg, icomp_type=2, wsave=ws)
    assert ier == 0
    ier, _, pt3, st3 = E.elhfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl2, sdl2, tg, icomp_type=3, wsave=ws)
    assert ier == 0
    assert E.last_timings()["mindlin"] > 0
    ier, _, pt3b, st3b = E.elhfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl2, sdl2, tg, icomp_type=3, wsave=ws)
    assert ier == 0
    ier, _, pt2, st2 = E.elhfmm3dtriatargiter(2, lam, mu, tri, nrm, src, ssl2, sdl2, tg, icomp_type=2, wsave=ws2)
    assert ier == 0
    assert rel(pt3, pt2) < 1e-13 and rel(st3, st2) < 1e-13
    assert rel(pt3b, pt3) < 1e-14 and rel(st3b, st3) < 1e-14
    assert E.iter_release(ws) == 1 and E.iter_release(ws2) == 1
