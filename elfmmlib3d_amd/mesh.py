"""Mesh and target generators for the layer-potential workloads.

Triangles use the reference's "standard format": ``triangle[k, v, c]`` is
coordinate ``c`` of vertex ``v`` of triangle ``k`` (the transpose of the
Fortran array ``triangle(3,3,n)``, so the bytes are identical), normals are
``trinorm[k, c]`` and centroids ``source[k, c]``.

Mesh helpers follow the reference's trilib routines:
``triangle_norm`` (src/trilib.f:243-279), ``triangle_area``
(src/trilib.f:286-318) and the Cart3D ``.tri`` reader ``atrireadchk``
(src/trilib.f:59-89).
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "icosahedron", "icosphere", "triangle_normals", "triangle_areas",
    "centroids", "read_tri", "write_tri", "shell_targets", "cube_targets",
    "fault_mesh",
]


def icosahedron():
    """Unit icosahedron with poles on the z axis and rings at z=+-1/sqrt(5).

    This is the orientation of the reference's sphere fixtures
    (examples/sphere*.a.tri: the first vertex is (cos 72deg, sin 72deg, 1/2)
    scaled to the unit sphere). Faces are oriented with outward normals.
    """
    z = 1.0 / np.sqrt(5.0)
    rho = 2.0 / np.sqrt(5.0)
    verts = [(0.0, 0.0, 1.0)]
    for i in range(5):
        a = 2.0 * np.pi * i / 5.0
        verts.append((rho * np.cos(a), rho * np.sin(a), z))
    for i in range(5):
        a = 2.0 * np.pi * (i + 0.5) / 5.0
        verts.append((rho * np.cos(a), rho * np.sin(a), -z))
    verts.append((0.0, 0.0, -1.0))
    v = np.array(verts)
    faces = []
    for i in range(5):
        u0, u1 = 1 + i, 1 + (i + 1) % 5
        l0, l1 = 6 + i, 6 + (i + 1) % 5
        faces.append((0, u0, u1))
        faces.append((u0, l0, u1))
        faces.append((u1, l0, l1))
        faces.append((11, l1, l0))
    f = np.array(faces, dtype=np.int64)
    # orient outward
    c = v[f].mean(axis=1)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = np.einsum("ij,ij->i", n, c) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def icosphere(nsplit, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Flat-triangle sphere: each icosahedron face split into nsplit x nsplit
    triangles on a barycentric grid, every vertex projected to the sphere.

    Returns ``triangle`` of shape (20*nsplit**2, 3, 3) (unshared vertices,
    counter-clockwise seen from outside).
    """
    v, f = icosahedron()
    n = int(nsplit)
    out = []
    for a, b, c in f:
        A, B, C = v[a], v[b], v[c]

        def P(i, j):
            # grid point i steps along A->B, j steps along A->C
            p = A + (B - A) * (i / n) + (C - A) * (j / n)
            return p / np.linalg.norm(p)

        for i in range(n):
            for j in range(n - i):
                out.append((P(i, j), P(i + 1, j), P(i, j + 1)))
                if i + j < n - 1:
                    out.append((P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)))
    tri = np.array(out) * float(radius) + np.asarray(center, dtype=np.float64)
    return np.ascontiguousarray(tri)


def triangle_normals(triangle):
    """Unit normals: (v2-v1) x (v3-v2), normalised (src/trilib.f:243-279)."""
    x = triangle[:, 1] - triangle[:, 0]
    y = triangle[:, 2] - triangle[:, 1]
    n = np.cross(x, y)
    return np.ascontiguousarray(n / np.linalg.norm(n, axis=1)[:, None])


def triangle_areas(triangle):
    """Areas: |(v2-v1) x (v3-v1)|/2 (src/trilib.f:286-318)."""
    x = triangle[:, 1] - triangle[:, 0]
    y = triangle[:, 2] - triangle[:, 0]
    return 0.5 * np.linalg.norm(np.cross(x, y), axis=1)


def centroids(triangle):
    return np.ascontiguousarray(triangle.mean(axis=1))


def read_tri(path):
    """Cart3D ``.tri``: "nverts nfaces", vertex rows, 1-based face rows."""
    with open(path) as fh:
        toks = fh.read().split()
    nv, nf = int(toks[0]), int(toks[1])
    vals = toks[2:]
    verts = np.array(vals[: 3 * nv], dtype=np.float64).reshape(nv, 3)
    faces = np.array(vals[3 * nv: 3 * nv + 3 * nf], dtype=np.int64).reshape(nf, 3) - 1
    return np.ascontiguousarray(verts[faces])


def write_tri(path, triangle):
    n = triangle.shape[0]
    with open(path, "w") as fh:
        fh.write(f"{3 * n} {n}\n")
        for p in triangle.reshape(-1, 3):
            fh.write(f"{p[0]:.16E} {p[1]:.16E} {p[2]:.16E}\n")
        for k in range(n):
            fh.write(f"{3 * k + 1} {3 * k + 2} {3 * k + 3}\n")


def shell_targets(m, rmin, rmax, rng):
    """m points uniform in volume in the shell rmin < r < rmax."""
    u = rng.random(m)
    r = (rmin ** 3 + u * (rmax ** 3 - rmin ** 3)) ** (1.0 / 3.0)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.ascontiguousarray(d * r[:, None])


def cube_targets(m, half, rexclude, rng):
    """m points uniform in [-half, half]^3 with |x| > rexclude (rejection)."""
    out = np.empty((0, 3))
    while out.shape[0] < m:
        p = rng.uniform(-half, half, size=(2 * m, 3))
        p = p[np.linalg.norm(p, axis=1) > rexclude]
        out = np.vstack([out, p])
    return np.ascontiguousarray(out[:m])


def fault_mesh(nfaults, ntri_total, rng, box=20.0, top=(0.0, 3.0)):
    """Planar rectangular faults in z<0 meshed into right triangles.

    Length U[2,10], width U[1,5], dip U[15,90] deg, top depth U[top] (default
    U[0,3]; every fault starts 1e-3 below it), random strike and centre in a
    box x box square; a uniform cell size h is chosen so that the total
    triangle count is close to ntri_total.
    """
    L = rng.uniform(2.0, 10.0, nfaults)
    W = rng.uniform(1.0, 5.0, nfaults)
    dip = np.deg2rad(rng.uniform(15.0, 90.0, nfaults))
    top = rng.uniform(top[0], top[1], nfaults)
    strike = rng.uniform(0.0, 2.0 * np.pi, nfaults)
    cx = rng.uniform(-box / 2, box / 2, nfaults)
    cy = rng.uniform(-box / 2, box / 2, nfaults)
    area = float(np.sum(L * W))
    h = np.sqrt(2.0 * area / ntri_total)
    tris = []
    for f in range(nfaults):
        nl = max(1, int(round(L[f] / h)))
        nw = max(1, int(round(W[f] / h)))
        s = np.array([np.cos(strike[f]), np.sin(strike[f]), 0.0])
        d = np.array([-np.sin(strike[f]) * np.cos(dip[f]),
                      np.cos(strike[f]) * np.cos(dip[f]),
                      -np.sin(dip[f])])
        o = np.array([cx[f], cy[f], -top[f] - 1e-3]) - s * L[f] / 2
        for i in range(nl):
            for j in range(nw):
                p00 = o + s * (L[f] * i / nl) + d * (W[f] * j / nw)
                p10 = o + s * (L[f] * (i + 1) / nl) + d * (W[f] * j / nw)
                p01 = o + s * (L[f] * i / nl) + d * (W[f] * (j + 1) / nw)
                p11 = o + s * (L[f] * (i + 1) / nl) + d * (W[f] * (j + 1) / nw)
                tris.append((p00, p10, p11))
                tris.append((p00, p11, p01))
    return np.ascontiguousarray(np.array(tris))


def sphere_workload(nsplit=64, ntarg=100000, seed=2024):
    """The sphere workload of the benchmark (C2 at the defaults): icosphere
    nsplit x nsplit per face, densities uniform in [-1,1]^3, ntarg targets
    uniform in [-2,2]^3 minus the unit ball. Returns (triangle, trinorm,
    source, sigma_sl, sigma_dl, target)."""
    rng = np.random.default_rng(seed)
    tri = icosphere(nsplit)
    nrm = triangle_normals(tri)
    src = centroids(tri)
    ssl = rng.uniform(-1, 1, (tri.shape[0], 3))
    sdl = rng.uniform(-1, 1, (tri.shape[0], 3))
    tg = cube_targets(ntarg, 2.0, 1.0, rng)
    return tri, nrm, src, ssl, sdl, tg


def sphere_batch(nmesh=256, tmin=500, tmax=20000, seed=2025, box=20.0):
    """The many-mesh workload: nmesh independent icospheres whose triangle
    counts 20*nsplit**2 follow a log-uniform draw in [tmin, tmax] (nsplit
    rounded, at least 1), radii uniform in [0.5, 3], centres uniform in
    [-box, box]^3, densities uniform in [-1,1]^3. Returns a list of
    (triangle, trinorm, source, sigma_sl, sigma_dl)."""
    rng = np.random.default_rng(seed)
    unit = {}
    out = []
    for _ in range(int(nmesh)):
        cnt = np.exp(rng.uniform(np.log(tmin), np.log(tmax)))
        ns = max(1, int(round(np.sqrt(cnt / 20.0))))
        r = rng.uniform(0.5, 3.0)
        c = rng.uniform(-box, box, 3)
        if ns not in unit:
            unit[ns] = icosphere(ns)
        tri = np.ascontiguousarray(unit[ns] * r + c)
        n = tri.shape[0]
        out.append((tri, triangle_normals(tri), centroids(tri), rng.uniform(-1, 1, (n, 3)),
                    rng.uniform(-1, 1, (n, 3))))
    return out


def strain_to_stress(strain, lam, mu):
    """Isotropic Hooke's law, sigma = lam tr(eps) I + 2 mu eps, for strains of
    shape (..., 3, 3) (el3dstrain2stress of the reference's MATLAB tools,
    matlab/Contents.m:27)."""
    strain = np.asarray(strain, dtype=np.float64)
    tr = np.trace(strain, axis1=-2, axis2=-1)
    return lam * tr[..., None, None] * np.eye(3) + 2.0 * mu * strain


def stress_to_traction(stress, normal):
    """Traction (stress vector) t_i = sigma_ij n_j for stresses (n, 3, 3) and
    unit normals (n, 3) (matlab/el3dstress2traction.m)."""
    return np.einsum("nij,nj->ni", np.asarray(stress, dtype=np.float64), np.asarray(normal, dtype=np.float64))
