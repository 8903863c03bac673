"""Launch orders of the host planner (elfmm_plan.cpp), checked on the CPU
against a recomputation from the octree: the M2L, M2M and L2L tiles,
reordered so that workgroup b runs on XCD b mod 8 with the work balanced
over the eight XCDs, still cover every (pair group, row block) of every
operator exactly once (M2L pairs under the operator of their offset,
src/lfmm3drouts.f:297-496); near-field leaves (self + list 1,
src/elfmm3dtria.f:621-1144) and P2M leaves keep their task contents and run
heaviest first."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "elfmmlib3d_amd", "csrc")


@pytest.fixture(scope="module")
def plan_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("po") / "libplan_order.so")
    srcs = [os.path.join(ROOT, "tests", "csrc", "plan_order.cpp")] + [
        os.path.join(CSRC, f) for f in ("elfmm_tree.cpp", "elfmm_plan.cpp", "elfmm_ops.cpp", "elfmm_pool.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", out, *srcs], check=True)
    return C.CDLL(out)


def geometry(n, nt, seed):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((n, 3))
    s /= np.linalg.norm(s, axis=1)[:, None]
    s[:, 2] *= 0.6                                   # an ellipsoid: uneven boxes
    t = rng.uniform(-2.0, 2.0, (nt, 3))
    return np.ascontiguousarray(s), np.ascontiguousarray(t)


@pytest.mark.parametrize("n,nt,iprec,spread", [(4000, 2000, 1, 1.25), (60000, 30000, 2, 1.02)])
def test_launch_orders(plan_lib, n, nt, iprec, spread):
    s, t = geometry(n, nt, 7 + n)
    out = np.zeros(6)
    rc = plan_lib.plan_order_check(s.ctypes.data_as(C.c_void_p), n, t.ctypes.data_as(C.c_void_p), nt,
                                   iprec, out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    assert out[5] > 0
    assert out[0] == 1.0, "M2L tiles do not cover the pairs once"
    assert out[2] == 1.0, "M2M/L2L tiles do not cover the pairs once"
    assert out[3] == 1.0, "near-field boxes changed or not heaviest first"
    assert out[4] == 1.0, "P2M boxes changed or not most sources first"
    # the largest XCD share of M2L work against the mean share
    assert out[1] <= spread, out[1]
