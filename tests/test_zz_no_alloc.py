"""The allocation-free direct sum (elfmm_gfx950_el3dtriadirecttarg_noalloc)
against the allocating entry point el3dtriadirecttarg_ on the same inputs:
two calls in a row, the second with fewer triangles and targets (the static
workspace keeps the first call's contents), and the size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def call_noalloc(E, lam, mu, tri, nrm, src, ssl, sdl, tg):
    import torch
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    n, m = tri.shape[0], tg.shape[0]
    d = [t(tri), t(nrm), t(src), t(ssl), t(sdl), t(tg)]
    out = [torch.zeros((n, 3), dtype=torch.float64, device=dev), torch.zeros((n, 3, 3), dtype=torch.float64, device=dev),
           torch.zeros((m, 3), dtype=torch.float64, device=dev), torch.zeros((m, 3, 3), dtype=torch.float64, device=dev)]
    P = lambda x: C.c_void_p(x.data_ptr())
    stream = torch.cuda.current_stream()
    lib = E.library()
    lib.elfmm_gfx950_el3dtriadirecttarg_noalloc.restype = C.c_int
    ier = lib.elfmm_gfx950_el3dtriadirecttarg_noalloc(
        C.c_double(lam), C.c_double(mu), C.c_int(n), P(d[0]), P(d[1]), P(d[2]), 1, P(d[3]), 1, P(d[4]), 1,
        P(out[0]), 1, P(out[1]), C.c_int(m), P(d[5]), 1, P(out[2]), 1, P(out[3]), C.c_void_p(stream.cuda_stream))
    torch.cuda.synchronize()
    return ier, [o.cpu().numpy() for o in out]


def test_noalloc_matches_allocating_entry(gpu_lib):
    g = load_golden("direct_s4")
    lam, mu = float(g["lam"]), float(g["mu"])
    full = (g["triangle"], g["trinorm"], g["source"], g["sigma_sl"], g["sigma_dl"], g["target"])
    small = tuple(a[:100] for a in full[:5]) + (full[5][:20],)
    for args in (full, small):
        ier, got = call_noalloc(gpu_lib, lam, mu, *args)
        assert ier == 0
        want = gpu_lib.el3dtriadirecttarg(lam, mu, *args)
        for a, b in zip(got, want):
            assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max(), np.abs(a - b).max()


def test_noalloc_rejects_sizes_above_the_workspace(gpu_lib):
    lib = gpu_lib.library()
    lib.elfmm_gfx950_el3dtriadirecttarg_noalloc.restype = C.c_int
    z = C.c_void_p(0)
    ier = lib.elfmm_gfx950_el3dtriadirecttarg_noalloc(C.c_double(1.0), C.c_double(1.0), C.c_int(131073), z, z, z, 1,
                                                      z, 0, z, 1, z, 0, z, C.c_int(0), z, 0, z, 0, z, z)
    assert ier == 128
