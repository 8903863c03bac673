"""gmres_simple (matlab/gmres_simple.m restated) on dense systems, and the
iter-engine operator it drives on the GPU."""
import numpy as np
import pytest

from conftest import load_golden, rel
from elfmmlib3d_amd.solver import IterOperator, gmres_simple


def test_gmres_solves_well_conditioned_system():
    rng = np.random.default_rng(0)
    n = 60
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) / np.sqrt(n)
    x_true = rng.standard_normal(n)
    b = A @ x_true
    x, ier, hist = gmres_simple(A, b, tol=1e-12, maxit=60)
    assert ier == 0
    assert np.linalg.norm(x - x_true) / np.linalg.norm(x_true) < 1e-10
    assert all(h2 < h1 for h1, h2 in zip(hist, hist[1:]))


def test_gmres_reports_maxit():
    rng = np.random.default_rng(1)
    n = 40
    A = np.eye(n) + 0.5 * rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    x, ier, hist = gmres_simple(A, b, tol=1e-14, maxit=3)
    assert ier in (4, 8)


def test_gmres_torch_tensors():
    torch = pytest.importorskip("torch")
    g = torch.Generator().manual_seed(2)
    n = 30
    A = torch.eye(n, dtype=torch.float64) * 3 + torch.randn(n, n, generator=g, dtype=torch.float64) / 6
    b = torch.randn(n, generator=g, dtype=torch.float64)
    x, ier, _ = gmres_simple(lambda v: A @ v, b, tol=1e-12, maxit=30)
    assert ier == 0 and float(torch.linalg.norm(A @ x - b) / torch.linalg.norm(b)) < 1e-11


@pytest.mark.gpu
def test_iter_operator_solve(gpu_lib):
    """Double-layer operator at the centroids through the iter engine: the
    stored-then-replayed products equal direct iter applies, and GMRES
    solves the interior-limit equation (K - 2 pi) sigma = b (second kind:
    the reference's double layer of a constant density is -4 pi sigma
    inside, 0 outside, so its principal value is the mean, -2 pi sigma)."""
    g = load_golden("iter_c1_p1")
    tri, nrm, src = g["triangle"], g["trinorm"], g["source"]
    op = IterOperator(1, float(g["lam"]), float(g["mu"]), tri, nrm, src, layer="double")
    rng = np.random.default_rng(7)
    s1 = rng.uniform(-1, 1, tri.shape[0] * 3)
    y1 = op(s1)                      # icomp 2
    y2 = op(s1)                      # icomp 3
    assert rel(y2, y1) < 1e-13
    ref = gpu_lib.elfmm3dtriatargiter(1, float(g["lam"]), float(g["mu"]), tri, nrm, src, None,
                                      s1.reshape(-1, 3), None, ifstrain=0, icomp_type=1)
    assert rel(y1, ref[2].reshape(-1)) < 1e-13
    A = lambda v: op(v) - 2 * np.pi * v
    s_true = rng.uniform(-1, 1, tri.shape[0] * 3)
    b = A(s_true)
    x, ier, hist = gmres_simple(A, b, tol=1e-6, maxit=60)
    print("gmres", ier, len(hist), hist[-1] if hist else None)
    assert ier == 0
    assert rel(A(x), b) < 1e-5
    assert op.release() == 1


@pytest.mark.gpu
def test_iter_operator_on_device(gpu_lib):
    """The operator on torch tensors in HBM equals the host operator, and
    GMRES runs on device vectors; the half-space operator (direct arrival +
    image part at the centroids) equals the sum of the two library calls."""
    import torch
    g = load_golden("iter_c1_p1")
    tri, nrm, src = g["triangle"], g["trinorm"], g["source"]
    lam, mu = float(g["lam"]), float(g["mu"])
    host = IterOperator(1, lam, mu, tri, nrm, src, layer="double")
    dev = IterOperator(1, lam, mu, tri, nrm, src, layer="double", device="cuda")
    rng = np.random.default_rng(9)
    s = rng.uniform(-1, 1, tri.shape[0] * 3)
    st = torch.as_tensor(s, device="cuda")
    for _ in range(2):                                   # store, then replay
        yh = host(s)
        yd = dev(st)
        assert isinstance(yd, torch.Tensor) and yd.is_cuda
        assert rel(yd.cpu().numpy(), yh) < 1e-13
    A = lambda v: dev(v) - 2 * np.pi * v
    b = A(torch.as_tensor(rng.uniform(-1, 1, tri.shape[0] * 3), device="cuda"))
    x, ier, hist = gmres_simple(A, b, tol=1e-6, maxit=60)
    assert ier == 0 and x.is_cuda
    assert float(torch.linalg.norm(A(x) - b) / torch.linalg.norm(b)) < 1e-5
    assert host.release() == 1 and dev.release() == 1
    # half space
    h = load_golden("hs_fault_p1")
    tri, nrm, src = h["triangle"], h["trinorm"], h["source"]
    lam, mu = float(h["lam"]), float(h["mu"])
    op = IterOperator(1, lam, mu, tri, nrm, src, layer="double", half=True, device="cuda")
    s = rng.uniform(-1, 1, (tri.shape[0], 3))
    y = op(torch.as_tensor(s.reshape(-1), device="cuda")).cpu().numpy().reshape(-1, 3)
    _, _, pa, *_ = gpu_lib.elfmm3dtriatargiter(1, lam, mu, tri, nrm, src, None, s, None, ifstrain=0)
    wh = np.zeros(4)                # the operator's first product stores (icomp 2): compare with the same
    _, _, ph, _ = gpu_lib.elhfmm3dtriatargiter(1, lam, mu, tri, nrm, src, None, s, src, 1, 0, icomp_type=2,
                                               wsave=wh)
    assert gpu_lib.iter_release(wh) == 1
    assert rel(y, pa + ph) < 1e-12
    y2 = op(torch.as_tensor(s.reshape(-1), device="cuda")).cpu().numpy().reshape(-1, 3)
    # the replay's image near field is the stored unit-density integral (the
    # reference's icomp 2/3 form), equal to the store call's result
    assert rel(y2, y) < 1e-10
    assert op.release() == 2


def test_bicgstab_solves_nonsymmetric_system():
    """Converges to the dense solution; one step on A = c I is exact
    (alpha = 1/c, the residual vanishes)."""
    from elfmmlib3d_amd.solver import bicgstab_simple
    rng = np.random.default_rng(3)
    n = 50
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) / np.sqrt(n)
    x_true = rng.standard_normal(n)
    b = A @ x_true
    x, ier, hist = bicgstab_simple(A, b, tol=1e-12, maxit=60)
    assert ier == 0
    assert np.linalg.norm(x - x_true) / np.linalg.norm(x_true) < 1e-10
    assert np.linalg.norm(A @ x - b) < 1e-12 * np.linalg.norm(b)
    x, ier, hist = bicgstab_simple(2.5 * np.eye(7), b[:7], tol=1e-12, maxit=5)
    assert ier == 0 and len(hist) == 1 and np.allclose(x, b[:7] / 2.5, rtol=0, atol=1e-15)
    x, ier, hist = bicgstab_simple(A, b, tol=1e-14, maxit=3)
    assert ier == 4 and len(hist) == 2
    x, ier, _ = bicgstab_simple(np.array([[4.0]]), np.array([2.0]))
    assert ier == 0 and x[0] == 0.5


def test_bicgstab_torch_tensors():
    torch = pytest.importorskip("torch")
    from elfmmlib3d_amd.solver import bicgstab_simple
    g = torch.Generator().manual_seed(4)
    n = 30
    # complex, as the listing's header describes: its beta (a ratio of
    # square roots of <r0, r>) turns a real system complex when <r0, r> < 0
    A = torch.eye(n, dtype=torch.complex128) * 3 + torch.randn(n, n, generator=g, dtype=torch.complex128) / 6
    b = torch.randn(n, generator=g, dtype=torch.complex128)
    x, ier, _ = bicgstab_simple(lambda v: A @ v, b, tol=1e-12, maxit=30)
    assert ier == 0 and float(torch.linalg.norm(A @ x - b) / torch.linalg.norm(b)) < 1e-11
