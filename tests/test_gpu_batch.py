"""Batched replays: one call replays a stored geometry for many density
sets (elfmm_gfx950_triatargiter_multi_dev / _htriatargiter_multi_dev) and
must equal the same replays made one set at a time, for the free-space
iter engine (both layers, displacement and strain at sources and targets)
and the half-space image part; 11 sets cross the 4- and 8-set passes of the
near field."""
import numpy as np
import pytest

from conftest import load_golden, rel

pytestmark = pytest.mark.gpu

NRHS = 11


def test_free_space_multi_replay(gpu_lib):
    import torch
    E = gpu_lib
    g = load_golden("iter_c1_p1")
    lam, mu = float(g["lam"]), float(g["mu"])
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    geo = [T(g[k]) for k in ("triangle", "trinorm", "source")]
    tg = T(g["target"])
    rng = np.random.default_rng(12)
    ssl = rng.uniform(-1, 1, (NRHS, n, 3))
    sdl = rng.uniform(-1, 1, (NRHS, n, 3))
    z = lambda *sh: torch.zeros(sh, dtype=torch.float64, device="cuda")
    ws = np.zeros(4)
    ier, _ = E.elfmm3dtriatargiter_dev(1, lam, mu, *geo, T(ssl[0]), T(sdl[0]), tg, z(n, 3), z(n, 3, 3), z(m, 3),
                                       z(m, 3, 3), 2, ws)
    assert ier == 0
    one = []
    for r in range(NRHS):
        out = [z(n, 3), z(n, 3, 3), z(m, 3), z(m, 3, 3)]
        ier, _ = E.elfmm3dtriatargiter_dev(1, lam, mu, *geo, T(ssl[r]), T(sdl[r]), tg, *out, 3, ws)
        assert ier == 0
        one.append([o.cpu().numpy() for o in out])
    outs = [z(NRHS, n, 3), z(NRHS, n, 3, 3), z(NRHS, m, 3), z(NRHS, m, 3, 3)]
    ier = E.triatargiter_multi_dev(1, lam, mu, *geo, T(ssl), T(sdl), tg, *outs, ws)
    assert ier == 0
    for r in range(NRHS):
        for k in range(4):
            assert rel(outs[k][r].cpu().numpy(), one[r][k]) < 1e-13, (r, k)
    assert E.iter_release(ws) == 1
    # without a store: ier 64, nothing computed
    assert E.triatargiter_multi_dev(1, lam, mu, *geo, T(ssl), T(sdl), tg, *outs, np.zeros(4)) == 64


def test_half_space_multi_replay(gpu_lib):
    import torch
    E = gpu_lib
    g = load_golden("hs_fault_p1")
    lam, mu = float(g["lam"]), float(g["mu"])
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    geo = [T(g[k]) for k in ("triangle", "trinorm", "source")]
    tg = T(g["target"])
    rng = np.random.default_rng(13)
    sdl = rng.uniform(-1, 1, (NRHS, n, 3))
    z = lambda *sh: torch.zeros(sh, dtype=torch.float64, device="cuda")
    ws = np.zeros(4)
    ier, _ = E.elhfmm3dtriatargiter_dev(1, lam, mu, *geo, None, T(sdl[0]), tg, z(m, 3), z(m, 3, 3), 2, ws)
    assert ier == 0
    one = []
    for r in range(NRHS):
        out = [z(m, 3), z(m, 3, 3)]
        ier, _ = E.elhfmm3dtriatargiter_dev(1, lam, mu, *geo, None, T(sdl[r]), tg, *out, 3, ws)
        assert ier == 0
        one.append([o.cpu().numpy() for o in out])
    outs = [z(NRHS, m, 3), z(NRHS, m, 3, 3)]
    ier = E.triatargiter_multi_dev(1, lam, mu, *geo, None, T(sdl), tg, None, None, *outs, ws, half=True)
    assert ier == 0
    for r in range(NRHS):
        for k in range(2):
            assert rel(outs[k][r].cpu().numpy(), one[r][k]) < 1e-13, (r, k)
    assert E.iter_release(ws) == 1
