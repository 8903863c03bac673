"""Mesh front end: the Cart3D .tri reader against the reference's own mesh
file (tests/golden/sphere2880.a.tri, a copy of examples/sphere2880.a.tri),
which is the generator's icosphere(12) (the C1 mesh); normals and areas
against the reference's trilib routines (triangle_norm, triangle_area,
src/trilib.f:243, :286) when the compiled reference is present; write/read
round trip; strain -> stress -> traction against Hooke's law and the
traction of a hydrostatic and a uniaxial state."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT
from elfmmlib3d_amd import mesh

TRI = os.path.join(ROOT, "tests", "golden", "sphere2880.a.tri")


def test_reference_sphere_is_the_generator_icosphere():
    t = mesh.read_tri(TRI)
    ic = mesh.icosphere(12)
    assert t.shape == ic.shape == (2880, 3, 3)
    # same triangles (order and vertex rotation may differ): match centroids
    ca, cb = mesh.centroids(t), mesh.centroids(ic)
    key = lambda c: np.lexsort(np.round(c, 12).T[::-1])
    ia, ib = key(ca), key(cb)
    assert np.abs(ca[ia] - cb[ib]).max() < 1e-14
    assert np.abs(mesh.triangle_normals(t)[ia] - mesh.triangle_normals(ic)[ib]).max() < 1e-13
    assert np.abs(mesh.triangle_areas(t)[ia] - mesh.triangle_areas(ic)[ib]).max() < 1e-15
    # all vertices on the unit sphere
    assert np.abs(np.linalg.norm(t.reshape(-1, 3), axis=1) - 1).max() < 1e-14


def test_normals_and_areas_match_reference_trilib():
    from oracle import ref
    if not ref.available():
        pytest.skip("reference library not built")
    lib = ref.lib()
    t = mesh.read_tri(TRI)[:200]
    nr = mesh.triangle_normals(t)
    ar = mesh.triangle_areas(t)
    for j in range(t.shape[0]):
        tri = np.ascontiguousarray(t[j])                 # (3 vertices, 3 coords) = Fortran (3, 3)
        nv = np.zeros(3)
        a = C.c_double(0.0)
        lib.triangle_norm_(tri.ctypes.data_as(C.c_void_p), nv.ctypes.data_as(C.c_void_p))
        lib.triangle_area_(tri.ctypes.data_as(C.c_void_p), C.byref(a))
        assert np.abs(nv - nr[j]).max() < 1e-15 and abs(a.value - ar[j]) < 1e-17


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    tri = mesh.fault_mesh(2, 300, rng)
    p = tmp_path / "f.tri"
    mesh.write_tri(str(p), tri)
    back = mesh.read_tri(str(p))
    assert np.abs(back - tri).max() < 1e-14 * np.abs(tri).max()


def test_strain_stress_traction():
    rng = np.random.default_rng(4)
    lam, mu = 1.3, 0.7
    e = rng.normal(size=(50, 3, 3))
    e = (e + e.transpose(0, 2, 1)) / 2
    s = mesh.strain_to_stress(e, lam, mu)
    for k in range(50):
        assert np.allclose(s[k], lam * np.trace(e[k]) * np.eye(3) + 2 * mu * e[k], rtol=0, atol=1e-14)
    nv = rng.normal(size=(50, 3))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    t = mesh.stress_to_traction(s, nv)
    assert np.allclose(t, np.array([s[k] @ nv[k] for k in range(50)]), rtol=0, atol=1e-14)
    # hydrostatic strain e I: traction (3 lam + 2 mu) e n
    h = mesh.stress_to_traction(mesh.strain_to_stress(np.tile(0.2 * np.eye(3), (50, 1, 1)), lam, mu), nv)
    assert np.allclose(h, (3 * lam + 2 * mu) * 0.2 * nv, rtol=0, atol=1e-14)
    # uniaxial strain along x, normal e_x: sigma_xx = lam + 2 mu
    u = np.zeros((1, 3, 3))
    u[0, 0, 0] = 1.0
    assert np.allclose(mesh.stress_to_traction(mesh.strain_to_stress(u, lam, mu), np.array([[1.0, 0, 0]])),
                       [[lam + 2 * mu, 0, 0]], rtol=0, atol=1e-15)
