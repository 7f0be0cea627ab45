"""GPU tests of the set-up on the GPU (include/moris_gpu_setup.h) against the host set-up
(mgpu_problem_build + mgpu_set_create), the reference's pipeline of
cl_XTK_Integration_Mesh_Generator.cpp:53-173 (cut), cl_MSI_Dof_Manager.cpp:703-745 (adof ids),
cl_HMR_T_Matrix_Base.cpp:306-457 (T rows) and cl_Sparse_Matrix_EpetraFECrs.cpp:195-240 (pattern).

* Planes and spheres: every integer array (elements, sub-cell offsets, T rows, topology, CSR pattern)
  and every coordinate array (sub-cell vertices, boxes, vertex level set, T weights) bit-identical;
  material within 2 ulp (exp, log and cos of the per-element conductivity come from the device
  math library); p = 1 and p = 2, 2D and 3D, phase 0 and 1.
* Gyroid (sin / cos from the device library): integer arrays identical, coordinates within 1e-13.
* A Set made from the GPU set-up's descriptor assembles K and R bit-identical to the host one.
* The headline batch's set-up (32 sphere-array variants, 64^3 elasticity): variants 0, 13 and 31
  bit-identical, device time per variant printed beside the host set-up's wall time.
"""
import time

import numpy as np
import pytest

import moris_amd as M
from helpers import SEED
from moris_amd import gpu_setup as G
from moris_amd import workloads as W

pytestmark = pytest.mark.gpu

INT_KEYS = ("elem_cell", "ig_ptr", "t_ptr", "t_adof")
FLOAT_KEYS = ("ig_xi", "ip_box", "vertex_phi", "t_weight")

CASES = {
    "sphere3_diff_p1": lambda: M.make_spec(3, (16, 16, 16), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5, seed=SEED + 1),
    "sphere3_diff_p1_phase0": lambda: M.make_spec(3, (12, 12, 12), spheres=[[0.45, 0.5, 0.55, 0.31]], k_sigma=0.5,
                                                  seed=SEED + 2, material_phase=0),
    "spheres3_elast_p1": lambda: M.make_spec(3, (20, 20, 20), iwg=M.IWG_STRUC_LINEAR_BULK,
                                             spheres=[[0.3, 0.3, 0.5, 0.2], [0.7, 0.6, 0.5, 0.22]]),
    "plane3_elast_p1": lambda: M.make_spec(3, (10, 10, 10), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_PLANE,
                                           plane=[0.3, 0.5, 0.81, -0.6]),
    "plane3_through_nodes": lambda: M.make_spec(3, (8, 8, 8), ls_kind=M.LS_PLANE, plane=[0.0, 0.0, 1.0, -0.5]),
    "sphere3_diff_p2": lambda: M.make_spec(3, (12, 12, 12), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.3, seed=SEED,
                                           bspline_order=2),
    "spheres3_elast_p2": lambda: M.make_spec(3, (10, 10, 10), iwg=M.IWG_STRUC_LINEAR_BULK,
                                             spheres=[[0.3, 0.3, 0.5, 0.2], [0.7, 0.6, 0.5, 0.22]], bspline_order=2),
    "circle2_elast_p1": lambda: M.make_spec(2, (32, 32), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.0, 0.25]]),
    "circle2_diff_p2_phase0": lambda: M.make_spec(2, (24, 24), spheres=[[0.5, 0.5, 0.0, 0.3]], k_sigma=0.4,
                                                  seed=SEED, material_phase=0, bspline_order=2),
    "circle2_elast_p2": lambda: M.make_spec(2, (16, 16), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 1.0, 0.0, 0.3]],
                                            bspline_order=2),
}


def _ulps(a, b):
    a = np.asarray(a, np.float64).view(np.int64)
    b = np.asarray(b, np.float64).view(np.int64)
    return int(np.abs(a - b).max()) if a.size else 0


def _compare(spec, keep, float_tol=0.0):
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Gs = G.GpuSetup(spec, keep)
    assert (Gs.n_elements, Gs.n_cut, Gs.n_ig, Gs.n_adofs_per_type) == (P.n_elements, P.n_cut, P.n_ig,
                                                                       P.n_adofs_per_type)
    assert (Gs.n_rows, Gs.nnz) == (S.n_rows, S.nnz)
    a, b = P.arrays(), Gs.arrays()
    for k in INT_KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in FLOAT_KEYS:
        if float_tol == 0.0:
            assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), (k, np.abs(a[k] - b[k]).max())
        else:
            assert np.abs(a[k] - b[k]).max() <= float_tol, k
    assert _ulps(a["material"], b["material"]) <= 2
    rp, ci = S.pattern()
    grp, gci = Gs.pattern()
    assert np.array_equal(rp, grp) and np.array_equal(ci, gci)
    assert np.array_equal(S.topology(), Gs.topology())
    return P, S, Gs


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_setup_equals_host_setup(case):
    spec, keep = CASES[case]()
    P, S, Gs = _compare(spec, keep)
    print(case, "elements", Gs.n_elements, "cut", Gs.n_cut, "sub-cells", Gs.n_ig, "rows", Gs.n_rows, "nnz", Gs.nnz,
          "device ms %.3f" % Gs.ms)


def test_gpu_setup_gyroid_to_tolerance():
    n = 24
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0)
    P, S, Gs = _compare(spec, keep, float_tol=1e-13)
    assert Gs.n_cut > 0.3 * Gs.n_elements


def test_set_from_gpu_setup_assembles_identically():
    spec, keep = CASES["spheres3_elast_p1"]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Gs = G.GpuSetup(spec, keep)
    S2 = M.Set(Gs.desc, keepalive=Gs)
    u = np.random.default_rng(SEED + 5).uniform(-1.0, 1.0, S.n_rows)
    v1, r1 = S.compute_jacobian_and_residual(u)
    v2, r2 = S2.compute_jacobian_and_residual(u)
    assert np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
    assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64))


def test_rejects_unsupported_specs():
    for kw in (dict(lagrange_order=2, bspline_order=2), dict(refine_levels=1, bspline_order=2), dict(bspline_order=3)):
        spec, keep = M.make_spec(3, (4, 4, 4), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.5, 0.3]], **kw)
        with pytest.raises(M.MgpuError):
            G.GpuSetup(spec, keep)


@pytest.mark.timeout(600)
def test_headline_batch_setup_variants():
    """The 64^3 sphere-array elasticity variants of the headline batch: GPU set-up == host set-up."""
    for k in (0, 13, 31):
        spec, keep = W.elasticity_spec(64, W.sphere_array(k))
        t0 = time.perf_counter()
        P = M.Problem(spec, keep)
        S = M.Set(P.desc, keepalive=P)
        host_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        Gs = G.GpuSetup(spec, keep)
        gpu_s = time.perf_counter() - t0
        a, b = P.arrays(), Gs.arrays()
        for key in INT_KEYS:
            assert np.array_equal(a[key], b[key]), key
        for key in FLOAT_KEYS:
            assert np.array_equal(a[key].view(np.uint64), b[key].view(np.uint64)), key
        rp, ci = S.pattern()
        grp, gci = Gs.pattern()
        assert np.array_equal(rp, grp) and np.array_equal(ci, gci)
        print("variant %d: elements %d cut %d sub-tets %d nnz %d | host set-up %.3f s, GPU set-up %.3f s wall, "
              "%.3f ms device" % (k, Gs.n_elements, Gs.n_cut, Gs.n_ig, Gs.nnz, host_s, gpu_s, Gs.ms), flush=True)
        del Gs, S, P
