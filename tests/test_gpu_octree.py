"""The device octree (elfm_gfx950_octree_dev) against the host octree
(elfmm_tree.cpp, itself pinned to the reference's d3tstrcr by
tests/test_tree.py): the same boxes (level, cube coordinates, parent,
child slots, point ranges), centres, point permutations and lists 1-5 in
the same order, bit for bit, on inputs chosen to stress the rules (ties on
box centres, deep and uneven trees, the half-space root, boxes with only
sources or only targets, no targets) and on the sizes of the workloads."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel
from elfmmlib3d_amd import mesh

pytestmark = pytest.mark.gpu

CSRC = os.path.join(ROOT, "elfmmlib3d_amd", "csrc")


@pytest.fixture(scope="module")
def host_tree_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("tree") / "libtree_check.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", out,
                    os.path.join(ROOT, "tests", "csrc", "tree_check.cpp"),
                    os.path.join(CSRC, "elfmm_tree.cpp"), os.path.join(CSRC, "elfmm_plan.cpp"), os.path.join(CSRC, "elfmm_ops.cpp"),
                    os.path.join(CSRC, "elfmm_pool.cpp")], check=True)
    return C.CDLL(out)


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def arrays(nb, sz, n, m):
    return (np.zeros((nb, 18), dtype=np.int32), np.zeros((nb, 3)), np.zeros(n, dtype=np.int32),
            np.zeros(max(m, 1), dtype=np.int32), np.zeros(5 * (nb + 1), dtype=np.int32),
            np.zeros(max(int(sz[2:].sum()), 1), dtype=np.int32))


def trim(a, m, sz):
    return a[0], a[1], a[2], a[3][:m], a[4], a[5][:int(sz[2:].sum())]


def host(lib, src, tgt, nbox, imindlin):
    n, m = len(src), len(tgt)
    s = np.ascontiguousarray(src, dtype=np.float64)
    t = np.ascontiguousarray(tgt if m else np.zeros((1, 3)), dtype=np.float64)
    nb = lib.tree_build(p(s), n, p(t), m, nbox, imindlin)
    assert nb > 0
    sz = np.zeros(7, dtype=np.int32)
    lib.tree_sizes(p(sz))
    a = arrays(nb, sz, n, m)
    lib.tree_get(p(np.zeros((nb, 7), dtype=np.int32)), p(a[1]), p(a[2]), p(a[3]), p(a[4]), p(a[5]))
    lib.tree_boxes_full(p(a[0]))
    return sz, trim(a, m, sz)


def device_tree(E, src, tgt, nbox, imindlin):
    """(number of boxes or error code, sizes, arrays) of the device build."""
    import torch
    n, m = len(src), len(tgt)
    ds = torch.as_tensor(np.ascontiguousarray(src, dtype=np.float64), device="cuda")
    dt = torch.as_tensor(np.ascontiguousarray(tgt, dtype=np.float64), device="cuda") if m else None
    sz = np.zeros(7, dtype=np.int32)
    lib = E.library()
    nb = lib.elfmm_gfx950_octree_dev(C.c_void_p(ds.data_ptr()), n, C.c_void_p(dt.data_ptr() if m else 0), m, nbox,
                                     imindlin, p(sz), C.c_void_p(0))
    if nb <= 0:
        return nb, sz, None
    a = arrays(nb, sz, n, m)
    lib.elfmm_gfx950_octree_get(*[p(x) for x in a])
    return nb, sz, trim(a, m, sz)


def cases():
    rng = np.random.default_rng(3)
    yield "uniform", rng.uniform(-1, 1, (3000, 3)), rng.uniform(-1, 1, (500, 3)), 20, 0
    yield "ties", rng.integers(-4, 5, (2000, 3)).astype(float), rng.integers(-4, 5, (300, 3)).astype(float), 10, 0
    yield "deep", mesh.centroids(mesh.icosphere(6)), np.zeros((0, 3)), 2, 0
    tri = mesh.fault_mesh(3, 2000, rng, box=6.0)
    yield "halfspace", mesh.centroids(tri), rng.uniform(-3, 3, (400, 3)) - [0, 0, 3.5], 40, 1
    # sources and targets in separate regions: boxes of one kind only
    yield "apart", rng.uniform(-1, 0, (4000, 3)), rng.uniform(0.2, 1, (4000, 3)), 30, 0
    # clusters a million times smaller than the root: deep, uneven levels
    c = rng.uniform(-1, 1, (6, 3))
    pts = np.concatenate([c[k] + 1e-4 * rng.standard_normal((500, 3)) for k in range(6)])
    yield "clusters", pts, rng.uniform(-1, 1, (300, 3)), 8, 0
    yield "c2", mesh.centroids(mesh.icosphere(64)), mesh.cube_targets(100000, 2.0, 1.0, rng), 40, 0
    tri = mesh.fault_mesh(8, 200000, np.random.default_rng(33))
    cen = mesh.centroids(tri)
    yield "c3_iter", cen, cen * [1, 1, -1], 240, 1


@pytest.mark.parametrize("case", list(range(8)))
def test_device_octree_matches_host(gpu_lib, host_tree_lib, case):
    name, src, tgt, nbox, imindlin = list(cases())[case]
    print(f"octree case {name}: {len(src)} sources, {len(tgt)} targets, nbox {nbox}", flush=True)
    hs, h = host(host_tree_lib, src, tgt, nbox, imindlin)
    nb, ds, d = device_tree(gpu_lib, src, tgt, nbox, imindlin)
    assert nb > 0 and hs[0] == ds[0], (nb, hs, ds)
    bad = np.argwhere(h[0] != d[0])
    assert len(bad) == 0, (name, "boxes differ", bad[:5].tolist(), h[0][bad[0][0]].tolist(), d[0][bad[0][0]].tolist())
    assert np.array_equal(hs, ds), (nb, hs, ds)
    for k, (a, b) in enumerate(zip(h, d)):
        assert np.array_equal(a, b), (name, ["boxes", "centers", "iz", "iztarg", "lstart", "litems"][k])


def test_too_deep_for_the_device_build(gpu_lib):
    """Clusters 1e-9 of the root wide need more than 21 levels: the device
    build reports it (-4), and an apply on them builds its tree on the host
    and gives the same result as with the host tree forced."""
    rng = np.random.default_rng(8)
    c = rng.uniform(-1, 1, (3, 3))
    pts = np.concatenate([c[k] + 1e-9 * rng.standard_normal((40, 3)) for k in range(3)])
    nb, _, _ = device_tree(gpu_lib, pts, np.zeros((0, 3)), 4, 0)
    assert nb == -4
    off = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=float) * 1e-10
    tri = pts[:, None, :] + off[None]
    nrm = mesh.triangle_normals(tri)
    src = mesh.centroids(tri)
    dens = rng.uniform(-1, 1, (len(tri), 3))
    tg = rng.uniform(-1, 1, (50, 3))
    a = gpu_lib.elfmm3dtriatarg(1, 1.0, 1.0, tri, nrm, src, dens, dens, tg)
    os.environ["ELFMM_HOST_TREE"] = "1"
    try:
        b = gpu_lib.elfmm3dtriatarg(1, 1.0, 1.0, tri, nrm, src, dens, dens, tg)
    finally:
        del os.environ["ELFMM_HOST_TREE"]
    assert a[0] == 0 and b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert rel(x, y) < 1e-13


def test_apply_same_with_device_and_host_trees(gpu_lib):
    """The C1 sphere with its targets: an apply on the device tree equals
    the apply on the host tree (the same tree, so the same arithmetic up to
    the order of atomic sums)."""
    from conftest import load_golden
    g = load_golden("fmm_c1")
    args = (2, float(g["lam"]), float(g["mu"]), g["triangle"], g["trinorm"], g["source"], g["sigma_sl"],
            g["sigma_dl"], g["target"])
    a = gpu_lib.elfmm3dtriatarg(*args)
    os.environ["ELFMM_HOST_TREE"] = "1"
    try:
        b = gpu_lib.elfmm3dtriatarg(*args)
    finally:
        del os.environ["ELFMM_HOST_TREE"]
    for x, y in zip(a[1:], b[1:]):
        assert rel(x, y) < 1e-14
