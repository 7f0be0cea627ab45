"""GPU tests of the set-up on the GPU (include/moris_gpu_setup.h) against the host set-up
(mgpu_problem_build + mgpu_set_create), the reference's pipeline of
cl_XTK_Integration_Mesh_Generator.cpp:53-173 (cut), cl_MSI_Dof_Manager.cpp:703-745 (adof ids),
cl_HMR_T_Matrix_Base.cpp:306-457 (T rows) and cl_Sparse_Matrix_EpetraFECrs.cpp:195-240 (pattern).

* Planes and spheres: every integer array (elements, sub-cell offsets, T rows, topology, CSR pattern)
  and every coordinate array (sub-cell vertices, boxes, vertex level set, T weights) bit-identical;
  material within 2 ulp (exp, log and cos of the per-element conductivity come from the device
  math library); p = 1 and p = 2, 2D and 3D, phase 0 and 1, and HEX27 cells with p = 2 or nodal rows.
* Gyroid (sin / cos from the device library): integer arrays identical, coordinates within 1e-13.
* A Set made from the GPU set-up's descriptor assembles K and R bit-identical to the host one (HEX8
  p = 1; HEX27 p = 2 with the subdivision and the moment-fitted rule).
* A Set made directly from the set-up's device arrays (mgpu_set_create_from_setup) equals the set
  mgpu_set_create makes from host copies of the same arrays: K and R bit-identical (FP64, mixed,
  mixed + learned), pattern, topology and statistics equal, a side set on it equal.
* The headline batch's set-up (32 sphere-array variants, 64^3 elasticity): variants 0, 13 and 31
  bit-identical, device time per variant printed beside the host set-up's wall time, and the
  set made from the device arrays.
"""
import time

import numpy as np
import pytest

import moris_amd as M
from helpers import SEED
from moris_amd import gpu_setup as G
from moris_amd import learned as L
from moris_amd import sideset as SS
from moris_amd import train_lq
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
    # HEX27 IP cells with p = 2 rows at the 27 nodes (the quadratic configuration without refinement)
    "sphere3_elast_hex27_p2": lambda: M.make_spec(3, (10, 9, 8), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                  spheres=[[0.45, 0.5, 0.55, 0.3]], bspline_order=2, lagrange_order=2),
    "plane3_diff_hex27_p2": lambda: M.make_spec(3, (7, 7, 7), ls_kind=M.LS_PLANE, plane=[0.3, 0.5, 0.8, -0.6],
                                                k_sigma=0.4, seed=SEED + 3, bspline_order=2, lagrange_order=2),
    # HEX27 IP cells with nodal rows (adofs on the half-step node grid)
    "sphere3_elast_hex27_nodal": lambda: M.make_spec(3, (7, 6, 5), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                     spheres=[[0.45, 0.5, 0.55, 0.32]], lagrange_order=2),
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


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("case", ["spheres3_elast_p1", "sphere3_diff_p1", "circle2_elast_p1", "plane3_through_nodes"])
def test_set_from_device_arrays_equals_host_set(case):
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    Gs = G.GpuSetup(spec, keep)
    # the reference: a set made through mgpu_set_create from host copies of the same set-up's arrays
    # (the host set-up's own set differs only where the device math library rounds the material)
    S = M.Set(Gs.desc, keepalive=Gs)
    D = Gs.make_set()
    assert (D.n_rows, D.nnz, D.n_cut) == (S.n_rows, S.nnz, S.n_cut)
    u = np.random.default_rng(SEED + 9).uniform(-1.0, 1.0, S.n_rows)
    modes = [(M.PREC_FP64, M.QUAD_SUBDIVISION), (M.PREC_MIXED, M.QUAD_SUBDIVISION)]
    if spec.dim == 3 and spec.iwg == M.IWG_STRUC_LINEAR_BULK:
        net = train_lq.load() or L.MLP(seed=3)
        L.attach(S, net)
        L.attach(D, net)
        modes.append((M.PREC_MIXED, M.QUAD_LEARNED))
    for prec, quad in modes:
        S.options(prec, quad)
        D.options(prec, quad)
        v1, r1 = S.compute_jacobian_and_residual(u)
        v2, r2 = D.compute_jacobian_and_residual(u)
        assert _same(v1, v2) and _same(r1, r2), (prec, quad)
    assert np.array_equal(S.statistics(), D.statistics())
    rp, ci = S.pattern()
    rp2, ci2 = D.pattern()
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    assert np.array_equal(S.topology(), D.topology())
    # a side set on each: the same values
    geom = SS.GEOM_TRI3 if spec.dim == 3 else SS.GEOM_LINE2
    iwg = SS.IWG_SPATIALDIFF_NEUMANN if spec.iwg == M.IWG_SPATIALDIFF_BULK else SS.IWG_STRUC_LINEAR_NEUMANN
    e, x, nrm = SS.facets(P, SS.SIDE_INTERFACE, geom)
    val = [0.3, -0.2, 0.1][: (1 if spec.iwg == M.IWG_SPATIALDIFF_BULK else spec.dim)]
    S.options(M.PREC_FP64, M.QUAD_SUBDIVISION)
    D.options(M.PREC_FP64, M.QUAD_SUBDIVISION)
    a1 = SS.assemble(S, [SS.SideSet(S, iwg, geom, e, x, nrm, value=val)], u)
    a2 = SS.assemble(D, [SS.SideSet(D, iwg, geom, e, x, nrm, value=val)], u)
    assert _same(a1[0], a2[0]) and _same(a1[1], a2[1])


def test_hex27_set_from_gpu_setup_assembles_identically():
    """A HEX27 set (generic path) made from the GPU set-up's arrays equals the host set-up's set bit
    for bit, with the subdivision rule and with the moment-fitted rule."""
    spec, keep = CASES["sphere3_elast_hex27_p2"]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Gs = G.GpuSetup(spec, keep)
    S2 = M.Set(Gs.desc, keepalive=Gs)
    u = np.random.default_rng(SEED + 6).uniform(-1.0, 1.0, S.n_rows)
    for quad in (M.QUAD_SUBDIVISION, M.QUAD_MOMENT_FIT):
        S.options(M.PREC_FP64, quad)
        S2.options(M.PREC_FP64, quad)
        v1, r1 = S.compute_jacobian_and_residual(u)
        v2, r2 = S2.compute_jacobian_and_residual(u)
        assert _same(v1, v2) and _same(r1, r2), quad


def test_set_from_device_arrays_needs_nodal_rows():
    spec, keep = CASES["sphere3_diff_p2"]()
    Gs = G.GpuSetup(spec, keep)
    with pytest.raises(M.MgpuError):
        Gs.make_set()


def test_rejects_unsupported_specs():
    for kw in (dict(lagrange_order=3), dict(refine_levels=1, bspline_order=2), dict(bspline_order=3),
               dict(enrich=1)):
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
        # the same variant end to end on the device: set-up, set from its arrays, one assembly
        M.synchronize()
        t0 = time.perf_counter()
        G2 = G.GpuSetup(spec, keep)
        D = G2.make_set()
        ub = M.DeviceBuffer.from_array(np.zeros(D.n_rows))
        vb, rb = M.DeviceBuffer(8 * D.nnz), M.DeviceBuffer(8 * D.n_rows)
        D.assemble_device(ub, vb, rb)
        M.synchronize()
        dev_s = time.perf_counter() - t0
        v1, _ = S.compute_jacobian_and_residual(np.zeros(S.n_rows))
        assert _same(v1, vb.to_array(np.float64, D.nnz))
        print("variant %d: elements %d cut %d sub-tets %d nnz %d | host set-up %.3f s, GPU set-up %.3f s wall, "
              "%.3f ms device; GPU set-up + set from its arrays + first assembly %.3f s" % (
                  k, Gs.n_elements, Gs.n_cut, Gs.n_ig, Gs.nnz, host_s, gpu_s, Gs.ms, dev_s), flush=True)
        del D, G2, Gs, S, P
