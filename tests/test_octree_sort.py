"""The device octree's construction (elfmm_octree.hip) restated with numpy
and checked against the host octree (elfmm_tree.cpp, pinned to the
reference's d3tstrcr by test_tree.py) on the CPU: octant paths by the
reference's floating centre formula, one stable sort of all points by
path, boxes opened level by level as runs of equal path prefixes among the
points of split boxes, point order within a leaf by input index. The GPU
test (test_gpu_octree.py) checks the device code itself against the same
host octree, lists included."""
import numpy as np
import pytest

from test_tree import lib_tree, tree_lib  # noqa: F401  (fixture)

D = 21


def paths(pts, centre, size):
    """Octant path of every point down to D levels (3 bits per level, z
    most significant), each centre (c - size/2) + (i-1)*side + side/2
    rounded operation by operation as the device does."""
    lo = centre - size * 0.5
    ijk = np.ones((len(pts), 3), dtype=np.int64)
    key = np.zeros(len(pts), dtype=np.uint64)
    for lev in range(D):
        side, side2 = np.ldexp(size, -lev), np.ldexp(size, -lev - 1)
        c = (lo[None, :] + (ijk - 1).astype(np.float64) * side) + side2
        bit = (pts > c).astype(np.int64)
        o = bit[:, 0] | (bit[:, 1] << 1) | (bit[:, 2] << 2)
        key = (key << np.uint64(3)) | o.astype(np.uint64)
        ijk = 2 * (ijk - 1) + 1 + bit
    return key


def sort_tree(src, tgt, nbox, imindlin=0):
    n, m = len(src), len(tgt)
    allp = np.concatenate([src, tgt]) if m else src
    mn, mx = allp.min(axis=0), allp.max(axis=0)
    size = max(mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2])
    centre = (mn + mx) / 2
    if imindlin:
        centre[2] = 0.0
        size = max(size, 2 * max(abs(mn[2]), abs(mx[2])))
    key = paths(allp, centre, size)
    order = np.argsort(key, kind="stable")
    K, P = key[order], order
    is_src = P < n
    sexcl = np.concatenate([[0], np.cumsum(is_src)[:-1]])
    boxes = [(0, 1, 1, 1, -1, 0, n, 0, m)]          # level, i, j, k, parent, src start/count, tgt start/count
    active = np.full(len(P), n > nbox or m > nbox)
    cur = np.zeros(len(P), dtype=np.int64)
    lev = 0
    while active.any():
        lev += 1
        pre = K >> np.uint64(3 * (D - lev))
        head = active & np.concatenate([[True], (~active[:-1]) | (pre[1:] != pre[:-1])])
        start = len(boxes)
        bid = np.cumsum(head) - 1 + start
        newcur = np.where(active, bid, cur)
        for i in np.flatnonzero(head):
            # run of the box: consecutive active points with this prefix
            e = i + 1
            while e < len(P) and active[e] and pre[e] == pre[i]:
                e += 1
            c = [0, 0, 0]
            for t in range(lev):
                o = int(pre[i] >> np.uint64(3 * (lev - 1 - t))) & 7
                c = [2 * c[d] + ((o >> d) & 1) for d in range(3)]
            ns = int(is_src[i:e].sum())
            boxes.append((lev, c[0] + 1, c[1] + 1, c[2] + 1, int(cur[i]), int(sexcl[i]), ns, i - int(sexcl[i]),
                          (e - i) - ns))
        cur = newcur
        split = np.array([b[6] > nbox or b[8] > nbox for b in boxes])
        active = active & split[cur]
    leafpos = np.array([b[5] + b[7] for b in boxes])[cur]     # first sorted point of each point's leaf
    pos = np.empty(len(P), dtype=np.int64)
    pos[P] = leafpos
    iz = np.argsort(pos[:n], kind="stable")
    izt = np.argsort(pos[n:], kind="stable") if m else np.zeros(0, dtype=np.int64)
    return np.array(boxes), iz, izt


def cases():
    from elfmmlib3d_amd import mesh
    rng = np.random.default_rng(3)
    yield rng.uniform(-1, 1, (1500, 3)), rng.uniform(-1, 1, (300, 3)), 20, 0
    yield rng.integers(-4, 5, (1000, 3)).astype(float), rng.integers(-4, 5, (200, 3)).astype(float), 10, 0
    tri = mesh.fault_mesh(3, 1200, rng, box=6.0)
    yield mesh.centroids(tri), rng.uniform(-3, 3, (300, 3)) - [0, 0, 3.5], 40, 1
    yield rng.uniform(-1, 0, (800, 3)), rng.uniform(0.2, 1, (800, 3)), 30, 0


@pytest.mark.parametrize("case", range(4))
def test_sorted_paths_give_the_reference_tree(tree_lib, case):  # noqa: F811
    src, tgt, nbox, imindlin = list(cases())[case]
    t = lib_tree(tree_lib, src, tgt, nbox, imindlin)
    boxes, iz, izt = sort_tree(src, tgt, nbox, imindlin)
    assert len(boxes) == t["boxes"].shape[0]
    assert np.array_equal(boxes[:, 0:4], t["boxes"][:, 0:4])          # level, i, j, k
    assert np.array_equal(boxes[:, 4], t["boxes"][:, 4])              # parent
    assert np.array_equal(boxes[:, 6], t["boxes"][:, 6])              # source count
    assert np.array_equal(iz, t["iz"]) and np.array_equal(izt, t["iztarg"])
