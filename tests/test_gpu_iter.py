"""Iter engine (elfmm3dtriatargiter, src/elfmm3dtria_iter_1.f:160) against
the reference built from its half-space link set: outputs of icomp 1, 2 and
3 (replay with new densities) within 1e-10 of the reference's, the near-list
structure through the reference's own store size (icomp 4 lused), and the
handle protocol of the device-resident store."""
import numpy as np
import pytest

from conftest import load_golden, rel

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def run(gpu_lib, g, icomp, wsave=None, second=False, **kw):
    ssl = g["sigma_sl2"] if second else g["sigma_sl"]
    sdl = g["sigma_dl2"] if second else g["sigma_dl"]
    return gpu_lib.elfmm3dtriatargiter(int(g["iprec"]), float(g["lam"]), float(g["mu"]), g["triangle"],
                                       g["trinorm"], g["source"], ssl, sdl, g["target"], icomp_type=icomp,
                                       wsave=wsave, **kw)


def check(out, g, tag=""):
    ier, lused, pt, st, ptt, stt = out
    for a, k in zip((pt, st, ptt, stt), ("ptfrc", "strain", "ptfrctarg", "straintarg")):
        e = rel(a, g[k + tag])
        assert e < PARITY, (k + tag, e)


@pytest.mark.parametrize("case", ["iter_c1_p2", "iter_c1_p1"])
def test_iter_compute_matches_reference(gpu_lib, case):
    g = load_golden(case)
    print(f"case {case} icomp 1", flush=True)
    out = run(gpu_lib, g, 1)
    assert out[0] == 0
    check(out, g)


@pytest.mark.parametrize("case", ["iter_c1_p2", "iter_c1_p1"])
def test_iter_plan_matches_reference_store_size(gpu_lib, case):
    g = load_golden(case)
    print(f"case {case} icomp 4", flush=True)
    ier, lused, pt, *_ = run(gpu_lib, g, 4)
    assert ier == 0 and lused == 4
    assert not pt.any()                       # icomp 4 leaves the outputs alone
    assert gpu_lib.last_stats()["iter_ref_store_words"] == int(g["lused"])


def test_iter_store_and_replay(gpu_lib):
    g = load_golden("iter_c1_p2")
    ier, lused, *_ = run(gpu_lib, g, 4)
    wsave = np.zeros(lused)
    print("icomp 2", flush=True)
    out2 = run(gpu_lib, g, 2, wsave)
    assert out2[0] == 0 and wsave[0] != 0
    check(out2, g)
    print("icomp 3 with new densities", flush=True)
    out3 = run(gpu_lib, g, 3, wsave, second=True)
    assert out3[0] == 0
    check(out3, g, "2")
    out1 = run(gpu_lib, g, 1, second=True)
    for a, b in zip(out3[2:], out1[2:]):
        assert rel(a, b) < 1e-13
    # a second replay gives the same numbers (up to the order of the far
    # field's atomic additions)
    out3b = run(gpu_lib, g, 3, wsave, second=True)
    for a, b in zip(out3[2:], out3b[2:]):
        assert rel(a, b) < 1e-14
    assert gpu_lib.iter_release(wsave) == 1
    assert gpu_lib.iter_release(wsave) == 0


def test_iter_replay_without_store_computes_directly(gpu_lib):
    g = load_golden("iter_c1_p1")
    wsave = np.zeros(4)
    print("icomp 3 without a store", flush=True)
    out = run(gpu_lib, g, 3, wsave)
    assert out[0] == 64
    check(out, g)


def test_iter_near_only_adds_to_outputs(gpu_lib):
    g = load_golden("iter_c1_p1")
    wsave = np.zeros(4)
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    rng = np.random.default_rng(3)
    pre = [rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3, 3)), rng.uniform(-1, 1, (m, 3)),
           rng.uniform(-1, 1, (m, 3, 3))]
    print("icomp 5 (near field only, added)", flush=True)
    a = run(gpu_lib, g, 5, wsave, ptfrc=pre[0].copy(), strain=pre[1].copy(), ptfrctarg=pre[2],
            straintarg=pre[3])
    b = run(gpu_lib, g, 5, np.zeros(4))
    assert a[0] == 0 and b[0] == 0
    for x, y, p in zip(a[2:], b[2:], pre):
        assert np.abs((x - p) - y).max() <= 1e-13 * np.abs(y).max()
    # near field + far field = the full result
    full = run(gpu_lib, g, 1)
    farfield_plus = run(gpu_lib, g, 3, wsave)    # replay of the icomp-5 store plus the far field
    assert farfield_plus[0] == 0
    for x, y in zip(farfield_plus[2:], full[2:]):
        assert rel(x, y) < 1e-13
    gpu_lib.iter_release(wsave)


def test_iter_on_surface_entry(gpu_lib):
    """elfmm3dtriaiter = elfmm3dtriatargiter without targets."""
    g = load_golden("iter_c1_p1")
    import ctypes as C
    lib = gpu_lib.library()
    n = g["triangle"].shape[0]
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    tri, nrm, src, ssl, sdl = (f(g[k]) for k in ("triangle", "trinorm", "source", "sigma_sl", "sigma_dl"))
    pt, st = np.zeros((n, 3)), np.zeros((n, 3, 3))
    ier, lused = C.c_int(0), C.c_int(0)
    wsave = np.zeros(4)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    I = lambda v: C.byref(C.c_int(v))
    D = lambda v: C.byref(C.c_double(v))
    lib.elfmm3dtriaiter_(C.byref(ier), I(int(g["iprec"])), D(float(g["lam"])), D(float(g["mu"])), P(tri),
                         P(nrm), I(n), P(src), I(1), P(ssl), I(1), P(sdl), I(1), P(pt), I(1), P(st), I(1),
                         P(wsave), I(4), C.byref(lused))
    assert ier.value == 0
    ref = gpu_lib.elfmm3dtriatargiter(int(g["iprec"]), float(g["lam"]), float(g["mu"]), tri, nrm, src, ssl,
                                      sdl, None, icomp_type=1)
    assert rel(pt, ref[2]) < 1e-13 and rel(st, ref[3]) < 1e-13


@pytest.mark.parametrize("layers", ["sl", "dl", "both"])
@pytest.mark.parametrize("fields", ["disp", "strain", "both"])
def test_iter_replay_every_store_layout(gpu_lib, layers, fields):
    """Replay (icomp 3) with new densities equals a direct compute (icomp 1,
    pinned to the reference above) for each stored layout: one or both
    layers, displacement and/or strain fields."""
    g = load_golden("iter_c1_p1")
    ssl2 = g["sigma_sl2"] if layers in ("sl", "both") else None
    sdl2 = g["sigma_dl2"] if layers in ("dl", "both") else None
    ssl1 = g["sigma_sl"] if ssl2 is not None else None
    sdl1 = g["sigma_dl"] if sdl2 is not None else None
    fd, fs = int(fields != "strain"), int(fields != "disp")
    flags = dict(ifptfrc=fd, ifstrain=fs, ifptfrctarg=fd, ifstraintarg=fs)

    def call(icomp, ssl, sdl, wsave=None):
        return gpu_lib.elfmm3dtriatargiter(int(g["iprec"]), float(g["lam"]), float(g["mu"]), g["triangle"],
                                           g["trinorm"], g["source"], ssl, sdl, g["target"], icomp_type=icomp,
                                           wsave=wsave, **flags)

    print(f"layers {layers} fields {fields}", flush=True)
    ier, lused, *_ = call(4, ssl1, sdl1)
    wsave = np.zeros(max(lused, 4))
    assert call(2, ssl1, sdl1, wsave)[0] == 0
    out3 = call(3, ssl2, sdl2, wsave)
    out1 = call(1, ssl2, sdl2)
    assert out3[0] == 0 and out1[0] == 0
    for a, b, on in zip(out3[2:], out1[2:], (fd, fs, fd, fs)):
        if on:
            assert rel(a, b) < 1e-12
    assert gpu_lib.iter_release(wsave) == 1


def test_iter_replay_on_changed_coordinates_computes_directly(gpu_lib):
    """A replay whose sources differ from those the store was built on (one
    centroid moved) computes directly (ier 64) instead of reusing the
    store's octree and interactions."""
    g = load_golden("iter_c1_p1")
    wsave = np.zeros(4)
    out2 = run(gpu_lib, g, 2, wsave)
    assert out2[0] == 0
    src = g["source"].copy()
    src[7, 0] += 1e-9
    moved = dict(g)
    moved["source"] = src
    out3 = run(gpu_lib, moved, 3, wsave)
    assert out3[0] == 64
    out1 = run(gpu_lib, moved, 1)
    for a, b in zip(out3[2:], out1[2:]):
        assert rel(a, b) < 1e-13
    assert gpu_lib.iter_release(wsave) == 1


def test_iter_concurrent_replays_of_one_store(gpu_lib):
    """Replays of one store from several host threads at once equal the
    same replays made one after another."""
    import threading
    g = load_golden("iter_c1_p1")
    wsave = np.zeros(4)
    assert run(gpu_lib, g, 2, wsave)[0] == 0
    n = g["triangle"].shape[0]
    rng = np.random.default_rng(4)
    dens = [(rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))) for _ in range(4)]

    def replay(k, out):
        d = dict(g)
        d["sigma_sl"], d["sigma_dl"] = dens[k]
        out[k] = run(gpu_lib, d, 3, wsave)

    alone = [None] * 4
    for k in range(4):
        replay(k, alone)
    together = [None] * 4
    th = [threading.Thread(target=replay, args=(k, together)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(alone, together):
        assert a[0] == 0 and b[0] == 0
        for x, y in zip(a[2:], b[2:]):
            assert rel(x, y) < 1e-13
    assert gpu_lib.iter_release(wsave) == 1
