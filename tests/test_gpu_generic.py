"""GPU tests of the generic element path (T rows with several weighted entries, sets without a
lattice): element K_p on the GPU, K_e = T^T K_p T in FP64, deterministic gather into the CSR
values, R = K u.

* quadratic B-spline T rows (bspline_order 2: every Lagrange node carries 2^dim splines of weight
  2^-dim) against the oracle's dense T^T K T: CSR pattern and element topology bit-exact, K and R
  within 1e-12 (FP64) and 1e-6 (mixed precision);
* HEX27 IP cells (lagrange_order 2: HEX_3x3x3 on whole cells, TET_35 on XTK sub-tets, p = 2
  B-spline rows of 8 to 27 entries or nodal rows) to the same bars;
* one THB refinement level (refine_levels 1) on HEX8 and HEX27 cells, up to 38 adofs per type and
  element, to the same bars;
* a nodal set with its lattice hint removed runs the generic path and reproduces the lattice
  stencil path to 1e-12;
* graph replays of generic sets are bit-identical to eager calls; learned quadrature is refused.
"""
import ctypes

import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, rel, two_bodies
from moris_amd import learned as L
from oracle import py as O

pytestmark = pytest.mark.gpu

CASES = {
    "p2_sphere_8_elasticity": lambda: M.make_spec(3, (8, 8, 8), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                  spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2),
    "p2_sphere_10_diffusion": lambda: M.make_spec(3, (10, 10, 10), spheres=[[0.45, 0.5, 0.55, 0.33]],
                                                  k_sigma=0.5, bspline_order=2),
    "p2_circle_16x12_plane_stress": lambda: M.make_spec(2, (16, 12), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                        spheres=[[0.5, 0.5, 0.0, 0.3]], bspline_order=2),
    "p2_plane_2d_diffusion_phase0": lambda: M.make_spec(2, (9, 7), ls_kind=M.LS_PLANE, plane=[0.6, 0.8, 0.0, -0.55],
                                                        material_phase=0, bspline_order=2),
    "p2_uncut_elasticity": lambda: M.make_spec(3, (3, 4, 2), iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2),
    # HEX27 IP cells (HEX_3x3x3 uncut, TET_35 on the sub-tets): p = 2 B-splines and nodal Lagrange rows
    "hex27_p2_gyroid_6_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                        ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi * 1.5,
                                                        bspline_order=2, lagrange_order=2),
    "hex27_p2_sphere_6_diffusion": lambda: M.make_spec(3, (6, 5, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5,
                                                       bspline_order=2, lagrange_order=2),
    "hex27_lagrange_sphere_5_elasticity": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                              spheres=[[0.4, 0.5, 0.6, 0.35]], lagrange_order=2),
    "hex27_p2_uncut_elasticity": lambda: M.make_spec(3, (3, 4, 2), iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2,
                                                     lagrange_order=2),
    # one THB level (truncated hierarchical quadratic B-splines, up to 38 adofs per type and element)
    "thb_sphere_6_hex8_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                        spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                                        refine_levels=1),
    "thb_plane_6_hex8_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_PLANE,
                                                       plane=[0.3, 0.5, 0.8, -0.6], bspline_order=2, refine_levels=1),
    "thb_plane_5_hex27_elasticity": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_PLANE,
                                                        plane=[0.3, 0.5, 0.8, -0.6], bspline_order=2,
                                                        lagrange_order=2, refine_levels=1),
    "thb_sphere_4_hex27_diffusion": lambda: M.make_spec(3, (4, 4, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.4,
                                                        bspline_order=2, lagrange_order=2, refine_levels=1),
    # XTK enrichment: a cell split into two subphase elements, bases replicated per level; nodal
    # p = 1 rows that leave the lattice stencil for the generic path
    "enriched_two_spheres_p1_elasticity": lambda: M.make_spec(3, (10, 10, 10), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                              spheres=two_bodies(3, 10), material_phase=0, enrich=1),
    "enriched_two_circles_p1_plane_stress": lambda: M.make_spec(2, (16, 16), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                                spheres=two_bodies(2, 16), material_phase=0,
                                                                enrich=1),
    "enriched_gyroid_7_p1_diffusion": lambda: M.make_spec(3, (7, 7, 7), ls_kind=M.LS_GYROID, gyroid_scale=9.0,
                                                          k_sigma=0.5, enrich=1),
    "enriched_gyroid_6_p2_hex27_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                                 ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi * 1.6,
                                                                 bspline_order=2, lagrange_order=2, enrich=1),
}


def _oracle_topology(s, nb, nf, N):
    rows = []
    for e in range(len(s["elem_cell"])):
        q0, q1 = s["t_ptr"][e * nb], s["t_ptr"][(e + 1) * nb]
        u = np.unique(s["t_adof"][q0:q1])
        rows.append(np.concatenate([u + f * N for f in range(nf)]))
    return rows


@pytest.mark.parametrize("case", sorted(CASES))
def test_generic_path_matches_oracle(case):
    print("case", case, flush=True)
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Q = O.OracleProblem(spec)
    s = Q.setup()
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    nb = P.desc.n_bases
    nf = spec.dim if spec.iwg == M.IWG_STRUC_LINEAR_BULK else 1
    topo, want = S.topology_lists(), _oracle_topology(s, nb, nf, Q.n_adofs_per_type)
    assert len(topo) == len(want) and all(np.array_equal(a, b) for a, b in zip(topo, want))
    u = np.random.default_rng(SEED + 11).uniform(-1.0, 1.0, S.n_rows)
    vo, Ro = Q.assemble(u)
    vals, R = S.compute_jacobian_and_residual(u)
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12, (rel(vals, vo), rel(R, Ro))
    S.options(M.PREC_MIXED, M.QUAD_SUBDIVISION)
    vals, R = S.compute_jacobian_and_residual(u)
    assert rel(vals, vo) <= 1e-6 and rel(R, Ro) <= 1e-6, (rel(vals, vo), rel(R, Ro))


def _without_lattice(desc):
    d = M.SetDesc()
    ctypes.memmove(ctypes.byref(d), ctypes.byref(desc), ctypes.sizeof(M.SetDesc))
    d.lattice.n[0] = 0
    d.elem_cell = None
    return d


@pytest.mark.parametrize("iwg", [M.IWG_SPATIALDIFF_BULK, M.IWG_STRUC_LINEAR_BULK])
def test_generic_path_reproduces_lattice_stencil(iwg):
    spec, keep = M.make_spec(3, (9, 9, 9), iwg=iwg, spheres=[[0.5, 0.5, 0.5, 0.33]], k_sigma=0.4)
    P = M.Problem(spec, keep)
    S_lat = M.Set(P.desc, keepalive=P)
    S_gen = M.Set(_without_lattice(P.desc), keepalive=P)
    assert np.array_equal(S_lat.pattern()[0], S_gen.pattern()[0])
    assert np.array_equal(S_lat.pattern()[1], S_gen.pattern()[1])
    u = np.random.default_rng(3).uniform(-1.0, 1.0, S_lat.n_rows)
    v1, R1 = S_lat.compute_jacobian_and_residual(u)
    v2, R2 = S_gen.compute_jacobian_and_residual(u)
    assert rel(v2, v1) <= 1e-12 and rel(R2, R1) <= 1e-12


@pytest.mark.parametrize("lagrange_order", [1, 2])
def test_generic_element_order_and_autotune(lagrange_order):
    """Generic sets in gather mode have two launch shapes: element workgroups in element order, or
    most sub-cells first.  Both give bit-identical K and R (each K_e has its own pool slot and the
    gather sums in a fixed order); the autotuner times both, keeps the faster, and the result still
    matches the oracle."""
    spec, keep = M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * 1.5, bspline_order=2, lagrange_order=lagrange_order)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    u = np.random.default_rng(SEED + 5).uniform(-1.0, 1.0, S.n_rows)
    ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
    out = []
    for order in (0, 1):
        S.tuning(cut_order=order)
        S.assemble_device(ub, vb, rb)
        M.synchronize()
        out.append((vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows)))
    assert np.array_equal(out[0][0].view(np.uint64), out[1][0].view(np.uint64))
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))
    ms = S.autotune(ub, vb, rb, iters=2, ignore_cache=True)
    assert set(ms) == {(0, 0), (0, 1)} and all(t > 0 for t in ms.values()), ms
    _, order = S.tuning()
    assert ms[(0, order)] == min(ms.values())
    S.assemble_device(ub, vb, rb)
    M.synchronize()
    assert np.array_equal(vb.to_array(np.float64, S.nnz).view(np.uint64), out[0][0].view(np.uint64))
    Q = O.OracleProblem(spec)
    s = Q.setup()
    vo, Ro, _ = O.assemble_desc(P.desc, s["row_ptr"], s["col_idx"], u)
    assert rel(out[0][0], vo) <= 1e-12 and rel(out[0][1], Ro) <= 1e-12


def test_generic_batch_replays_and_learned_refused():
    sets, bufs, eager = [], [], []
    for k, r in enumerate((0.28, 0.31)):
        spec, keep = M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.5, r]],
                                 bspline_order=2)
        P = M.Problem(spec, keep)
        S = M.Set(P.desc, keepalive=P)
        u = np.random.default_rng(k).uniform(-1.0, 1.0, S.n_rows)
        ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
        S.assemble_device(ub, vb, rb)
        M.synchronize()
        eager.append((vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows)))
        sets.append(S)
        bufs.append((ub, vb, rb))
    B = M.Batch(sets, [b[0] for b in bufs], [b[1] for b in bufs], [b[2] for b in bufs], n_branches=2)
    for _ in range(3):
        for S, (_, vb, rb) in zip(sets, bufs):
            z = np.zeros(S.nnz)
            M.load_library().mgpu_memcpy_htod(vb.ptr, z.ctypes.data, z.nbytes)
        M.synchronize()
        B.launch()
        B.synchronize()
        for S, (_, vb, rb), (ve, re) in zip(sets, bufs, eager):
            assert np.array_equal(vb.to_array(np.float64, S.nnz), ve)
            assert np.array_equal(rb.to_array(np.float64, S.n_rows), re)
    B.close()
    with pytest.raises(M.MgpuError, match="learned quadrature"):
        L.attach(sets[0], L.MLP(seed=1))


def test_quadratic_thb_gyroid_size_independent_properties():
    """The quadratic configuration at a size the oracle does not reach in seconds (12^3 base, one THB
    level on the intersected cells, HEX27 cells, gyroid of period 4 cells, elasticity): K is
    symmetric, rigid translations (partition of unity) are in its null space, u^T K u >= 0, R = K u,
    and the mixed path stays within 1e-6 of FP64."""
    n = 12
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2, refine_levels=1)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    print("elements", P.n_elements, "cut", P.n_cut, "rows", S.n_rows, "nnz", S.nnz, flush=True)
    rp, ci = S.pattern()
    N = S.n_rows // 3
    u = np.random.default_rng(SEED + 41).uniform(-1.0, 1.0, S.n_rows)
    vals, R = S.compute_jacobian_and_residual(u)
    import scipy.sparse as sp
    K = sp.csr_matrix((vals, ci, rp), shape=(S.n_rows, S.n_rows))
    kmax = np.abs(vals).max()
    assert abs(K - K.T).max() <= 1e-13 * kmax
    for f in range(3):
        t = np.zeros(S.n_rows)
        t[f * N:(f + 1) * N] = 1.0
        assert np.abs(K @ t).max() <= 1e-12 * kmax * np.sqrt(S.n_rows)
    assert u @ (K @ u) >= -1e-12 * kmax * S.n_rows
    assert rel(R, K @ u) <= 1e-12
    S.options(M.PREC_MIXED, M.QUAD_SUBDIVISION)
    vm, Rm = S.compute_jacobian_and_residual(u)
    assert rel(vm, vals) <= 1e-6 and rel(Rm, R) <= 1e-6


HEX27_CASES = sorted(c for c in CASES if "hex27" in c)


@pytest.mark.parametrize("case", HEX27_CASES)
def test_hex27_scatter_mode_matches_oracle_and_gather(case, monkeypatch):
    """Scatter mode of the HEX27 kernel (colour classes adding K_e into the CSR values, the mode of
    sets too large for a K_e pool), forced on small sets: the same pattern, K and R within 1e-12 of
    the oracle and of the gather mode (only the order of summation differs), 1e-6 on the mixed path,
    and bit-identical repeated calls."""
    print("case", case, flush=True)
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    S_gather = M.Set(P.desc, keepalive=P)
    monkeypatch.setenv("MGPU_GENERIC_SCATTER", "1")
    S = M.Set(P.desc, keepalive=P)
    monkeypatch.delenv("MGPU_GENERIC_SCATTER")
    Q = O.OracleProblem(spec)
    s = Q.setup()
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    u = np.random.default_rng(SEED + 12).uniform(-1.0, 1.0, S.n_rows)
    vo, Ro = Q.assemble(u)
    vals, R = S.compute_jacobian_and_residual(u)
    vg, Rg = S_gather.compute_jacobian_and_residual(u)
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12, (rel(vals, vo), rel(R, Ro))
    assert rel(vals, vg) <= 1e-12 and rel(R, Rg) <= 1e-12
    v2, R2 = S.compute_jacobian_and_residual(u)
    assert np.array_equal(v2, vals) and np.array_equal(R2, R)
    S.options(M.PREC_MIXED, M.QUAD_SUBDIVISION)
    vm, Rm = S.compute_jacobian_and_residual(u)
    assert rel(vm, vo) <= 1e-6 and rel(Rm, Ro) <= 1e-6, (rel(vm, vo), rel(Rm, Ro))


@pytest.mark.parametrize("quad", [M.QUAD_SUBDIVISION, M.QUAD_MOMENT_FIT])
def test_hex27_scatter_mode_graph_replays(monkeypatch, quad):
    """Graph replays of scatter-mode sets zero the values themselves (a fill kernel, no memset
    node) and are bit-identical to eager calls, whatever the buffers held before; with the
    subdivision and with the moment-fitted rule."""
    monkeypatch.setenv("MGPU_GENERIC_SCATTER", "1")
    sets, bufs, eager = [], [], []
    for k, n in enumerate((5, 6)):
        spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                                 gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2,
                                 refine_levels=k)
        P = M.Problem(spec, keep)
        S = M.Set(P.desc, keepalive=P)
        S.options(M.PREC_FP64, quad)
        u = np.random.default_rng(k).uniform(-1.0, 1.0, S.n_rows)
        ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
        S.assemble_device(ub, vb, rb)
        M.synchronize()
        eager.append((vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows)))
        sets.append(S)
        bufs.append((ub, vb, rb))
    monkeypatch.delenv("MGPU_GENERIC_SCATTER")
    B = M.Batch(sets, [b[0] for b in bufs], [b[1] for b in bufs], [b[2] for b in bufs], n_branches=2)
    for _ in range(3):
        for S, (_, vb, rb) in zip(sets, bufs):
            z = np.full(S.nnz, 7.0)
            M.load_library().mgpu_memcpy_htod(vb.ptr, z.ctypes.data, z.nbytes)
        M.synchronize()
        B.launch()
        B.synchronize()
        for S, (_, vb, rb), (ve, re) in zip(sets, bufs, eager):
            assert np.array_equal(vb.to_array(np.float64, S.nnz), ve)
            assert np.array_equal(rb.to_array(np.float64, S.n_rows), re)
    B.close()


def test_thb_scatter_mode_chosen_by_size_properties():
    """A THB set large enough (20^3 base, one level, about 3.3e8 element-matrix entries) to take the
    scatter mode without being forced: symmetric K, rigid translations in its null space,
    u^T K u >= 0, R = K u, mixed within 1e-6 of FP64."""
    n = 20
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2, refine_levels=1)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    entries = int(np.sum(np.diff(S.topology_offsets()).astype(np.int64) ** 2))
    print("elements", P.n_elements, "rows", S.n_rows, "nnz", S.nnz, "K_e entries", entries, flush=True)
    assert entries > 2 ** 28
    rp, ci = S.pattern()
    N = S.n_rows // 3
    u = np.random.default_rng(SEED + 43).uniform(-1.0, 1.0, S.n_rows)
    vals, R = S.compute_jacobian_and_residual(u)
    import scipy.sparse as sp
    K = sp.csr_matrix((vals, ci, rp), shape=(S.n_rows, S.n_rows))
    kmax = np.abs(vals).max()
    assert abs(K - K.T).max() <= 1e-13 * kmax
    for f in range(3):
        t = np.zeros(S.n_rows)
        t[f * N:(f + 1) * N] = 1.0
        assert np.abs(K @ t).max() <= 1e-12 * kmax * np.sqrt(S.n_rows)
    assert u @ (K @ u) >= -1e-12 * kmax * S.n_rows
    assert rel(R, K @ u) <= 1e-12
    S.options(M.PREC_MIXED, M.QUAD_SUBDIVISION)
    vm, Rm = S.compute_jacobian_and_residual(u)
    assert rel(vm, vals) <= 1e-6 and rel(Rm, R) <= 1e-6
