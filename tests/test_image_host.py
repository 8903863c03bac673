"""The half-space image kernels and adaptive quadrature of the library
(elfmm_image.h, the code the image near-field kernels run), built for the
host, against the reference's own routines compiled from its sources
(oracle/_ref/libelfmm_refh.so): green3eluh_image_eval / green3elth_image_eval
(src/el3greenh.f:643, :118) at random point pairs, and eluh3triaadap /
elth3triaadap (src/elhfmm3dtria_iter.f:2653, :2777, adaptive tria3adam) on
random triangles and targets, including the centroid (the self-evaluation
point of the half-space path). Both outputs are requested, the setting in
which the reference's integrand fills every entry its subdivision test
reads."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "tests", "csrc")


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    from oracle import ref
    if not ref.available_h():
        pytest.skip("reference half-space library not built (make -C oracle refh)")
    out = str(tmp_path_factory.mktemp("img") / "image_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, os.path.join(CSRC, "image_host.cpp")],
                   check=True)
    return C.CDLL(out), ref.libh()


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)


def I(v):
    return C.byref(C.c_int(v))


def D(v):
    return C.byref(C.c_double(v))


def test_point_kernels(libs):
    h, ref = libs
    rng = np.random.default_rng(1)
    for t in range(400):
        lam, mu = rng.uniform(0.5, 2), rng.uniform(0.5, 2)
        x = np.array([*rng.uniform(-2, 2, 2), rng.uniform(-3, -0.01)])
        y = np.array([*rng.uniform(-2, 2, 2), rng.uniform(-3, -0.01)])
        s = rng.uniform(-1, 1, 3)
        nv = rng.normal(size=3)
        nv /= np.linalg.norm(nv)
        for layer in (0, 1):
            fv, st = np.zeros(3), np.zeros((3, 3))
            if layer == 0:
                ref.green3eluh_image_eval_(D(lam), D(mu), P(x), P(s), P(y), fv.ctypes.data_as(C.c_void_p), I(1),
                                           st.ctypes.data_as(C.c_void_p))
            else:
                ref.green3elth_image_eval_(D(lam), D(mu), P(x), P(s), P(nv), P(y), fv.ctypes.data_as(C.c_void_p),
                                           I(1), st.ctypes.data_as(C.c_void_p))
            f = np.zeros(12)
            h.image_host_point(layer, C.c_double(lam), C.c_double(mu), P(x), P(s), P(nv), P(y), 1, 1,
                               f.ctypes.data_as(C.c_void_p))
            r = np.concatenate([fv, st.T.ravel()])
            assert np.abs(f - r).max() <= 1e-12 * np.abs(r).max(), (t, layer)


def test_adaptive_integrals(libs):
    h, ref = libs
    rng = np.random.default_rng(2)
    for t in range(120):
        lam, mu = 1.0, 0.8
        c = np.array([*rng.uniform(-1, 1, 2), -rng.uniform(0.0, 1.0)])
        tri = c + rng.uniform(-0.5, 0.5, (3, 3))
        tri[:, 2] = np.minimum(tri[:, 2], -1e-3)
        y = tri.mean(0) if t % 4 == 0 else c + rng.uniform(-0.7, 0.7, 3)
        y[2] = min(y[2], -1e-3)
        s = rng.uniform(-1, 1, 3)
        nv = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        nv /= np.linalg.norm(nv)
        for layer in (0, 1):
            fv, st, nf = np.zeros(3), np.zeros((3, 3)), C.c_int(0)
            fn = ref.eluh3triaadap_ if layer == 0 else ref.elth3triaadap_
            fn(D(lam), D(mu), I(1), P(tri), P(s), P(nv), P(y), I(1), fv.ctypes.data_as(C.c_void_p), I(1),
               st.ctypes.data_as(C.c_void_p), C.byref(nf))
            f, ns = np.zeros(12), C.c_int(0)
            h.image_host_adap(layer, C.c_double(lam), C.c_double(mu), P(tri), P(s), P(nv), P(y), 1, 1,
                              f.ctypes.data_as(C.c_void_p), C.byref(ns))
            r = np.concatenate([fv, st.T.ravel()])
            assert np.abs(f - r).max() <= 1e-12 * np.abs(r).max(), (t, layer)
            # same subdivision: 12 evaluations for the triangle, 48 per step
            assert nf.value == 12 + 48 * ns.value, (t, layer, nf.value, ns.value)
