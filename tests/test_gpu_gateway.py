"""The MATLAB-gateway calling convention (elfmmlib3d_amd.gateway, after
matlab/elfmm3d_r2013a.mw) on the GPU: column-major arguments and U outputs
equal to the reference's fixtures (free space, direct sums) and, for the
half space, the image part plus the free-space part within the requested
tolerance of the half-space direct sum."""
import numpy as np
import pytest

from conftest import load_golden, rel

pytestmark = pytest.mark.gpu


def matlab_args(g, n):
    tri = np.asarray(g["triangle"]).transpose(2, 1, 0)          # (3, 3, n): coordinate, vertex, triangle
    return (n, tri, g["trinorm"].T, g["source"].T)


def test_gateway_free_space_matches_reference(gpu_lib):
    from elfmmlib3d_amd import gateway
    g = load_golden("fmm_s6")
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    U = gateway.elfmm3dtria(int(g["iprec"]), float(g["lam"]), float(g["mu"]), *matlab_args(g, n), 1,
                            g["sigma_sl"].T, 1, g["sigma_dl"].T, 1, 1, m, g["target"].T)
    assert U.ier == 0
    assert rel(U.pot.T, g["ptfrc"]) < 1e-10 and rel(U.strain.transpose(2, 1, 0), g["strain"]) < 1e-10
    assert rel(U.pottarg.T, g["ptfrctarg"]) < 1e-10
    assert rel(U.straintarg.transpose(2, 1, 0), g["straintarg"]) < 1e-10
    # 11 arguments: source outputs only (a tree without the targets: equal
    # within the FMM tolerance)
    U2 = gateway.elfmm3dtria(int(g["iprec"]), float(g["lam"]), float(g["mu"]), *matlab_args(g, n), 1,
                             g["sigma_sl"].T, 1, g["sigma_dl"].T)
    assert not hasattr(U2, "pottarg") and rel(U2.pot, U.pot) < 1e-6


def test_gateway_direct_sum_matches_reference(gpu_lib):
    from elfmmlib3d_amd import gateway
    g = load_golden("direct_s4")
    n, m = g["triangle"].shape[0], g["target"].shape[0]
    U = gateway.el3dtriadirect(float(g["lam"]), float(g["mu"]), *matlab_args(g, n), 1, g["sigma_sl"].T, 1,
                               g["sigma_dl"].T, 1, 1, m, g["target"].T)
    assert rel(U.pot.T, g["ptfrc"]) < 1e-11 and rel(U.straintarg.transpose(2, 1, 0), g["straintarg"]) < 1e-11


def test_gateway_half_space(gpu_lib):
    from elfmmlib3d_amd import gateway
    g = load_golden("hs_fault_p2")
    sub = g["direct_idx"]
    n = g["triangle"].shape[0]
    tg = g["target"][sub]
    args = (float(g["lam"]), float(g["mu"]), *matlab_args(g, n), 1, g["sigma_sl"].T, 1, g["sigma_dl"].T, 0, 0,
            len(sub), tg.T, 1, 1)
    U = gateway.elhfmm3dtria(2, *args)
    assert U.ier == [0, 0]
    D = gateway.elh3dtriadirect(*args)
    assert rel(D.pottarg.T, g["direct_pt"]) < 1e-11
    assert rel(U.pottarg, D.pottarg) < 0.5e-6 and rel(U.straintarg, D.straintarg) < 0.5e-5
