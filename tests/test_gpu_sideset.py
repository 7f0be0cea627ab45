"""GPU tests of the side-set IWGs (include/moris_gpu_sideset.h).

* Bulk set plus side sets assembled on the GPU against the oracle's bulk assembly plus its own
  facets and side integrands, at 1e-12 for K and R: Nitsche Dirichlet (SYMMETRIC and UNSYMMETRIC)
  and Neumann, on the interface and on domain boundary faces (TRI3 sub-cell faces, QUAD4 hex faces,
  2D LINE2 sides), diffusion and elasticity, random u; a second call is bit-identical.  The same
  on generic-path sets: p = 2 B-splines on HEX8 and QUAD4 cells, HEX27 cells, one THB level.
* The reference's known answer (UT_MDL_XTK_DiffusionElement.cpp:91-340) assembled on the GPU:
  T = 5 + 20 z at every node of the 1 x 1 x 4 lattice cut at z = 3.1, solved by a direct solver
  and by GPU GMRES (the system is not symmetric).
* C1 with the reference's weak boundary conditions: 64 x 64 QUAD4 plane stress around the circular
  hole, left edge clamped by SYMMETRIC Nitsche (alpha 100), unit traction (1, 0) on the right edge
  by Neumann; K u = f solved on the GPU (CG) for K of the oracle, of the FP64 GPU path and of the
  mixed path: displacement errors in the Euclidean norm.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import moris_amd as M
from helpers import NV, SEED, oracle_side_system, rel, two_bodies
from moris_amd import sideset as SS
from moris_amd import solve as SV

pytestmark = pytest.mark.gpu

CASES = {
    "diff3_sphere": (lambda: M.make_spec(3, (8, 8, 8), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5, seed=SEED + 1),
                     [(SS.IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 3.0, [0.4]),
                      (SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_XMAX, SS.GEOM_QUAD4, 0.0, [1.5])]),
    "diff3_face_sphere_unsym": (lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.0, 0.5, 0.5, 0.3]], k_sigma=0.5,
                                                    seed=SEED + 2),
                                [(SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_TRI3, 2.0,
                                  [1.0]),
                                 (SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_QUAD4, 2.0,
                                  [1.0]),
                                 (SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 0.0, [-2.0])]),
    "elast3_sphere_array": (lambda: M.make_spec(3, (10, 10, 10), iwg=1, spheres=[[0.3, 0.3, 0.5, 0.2],
                                                                                 [0.7, 0.6, 0.5, 0.22]]),
                            [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_QUAD4, 10.0,
                              [0.01, -0.02, 0.005]),
                             (SS.IWG_STRUC_LINEAR_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 5.0,
                              [0.0, 0.1, 0.0]),
                             (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_XMAX, SS.GEOM_QUAD4, 0.0, [1.0, 0.5, -0.25])]),
    "elast2_circle": (lambda: M.make_spec(2, (24, 24), iwg=1, spheres=[[0.5, 0.5, 0.0, 0.25]]),
                      [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_LINE2, 100.0, [0, 0]),
                       (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_XMAX, SS.GEOM_LINE2, 0.0, [1.0, 0.0]),
                       (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_LINE2, 0.0, [0.1, -0.3])]),
    "diff2_circle_ymax": (lambda: M.make_spec(2, (16, 16), spheres=[[0.5, 1.0, 0.0, 0.3]], k_sigma=0.3, seed=SEED),
                          [(SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_YMAX, SS.GEOM_LINE2, 4.0, [2.0]),
                           (SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_LINE2, 0.0, [3.0])]),
}


# side sets on generic-path sets: p = 2 B-splines on HEX8 / QUAD4, HEX27 cells, THB rows
GENERIC_CASES = {
    "p2_hex8_diff_sphere": (lambda: M.make_spec(3, (8, 8, 8), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5,
                                                seed=SEED + 3, bspline_order=2),
                            [(SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 0.0, [1.5]),
                             (SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_QUAD4, 3.0, [0.4])]),
    "hex27_elast_sphere": (lambda: M.make_spec(3, (6, 6, 6), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                               lagrange_order=2),
                           [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 10.0,
                             [0.01, -0.02, 0.005]),
                            (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_XMAX, SS.GEOM_QUAD4, 0.0, [1.0, 0.5, -0.25])]),
    "hex27_diff_face_sphere": (lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.0, 0.5, 0.5, 0.3]], k_sigma=0.3,
                                                   seed=SEED, bspline_order=2, lagrange_order=2),
                               [(SS.IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_TRI3, 2.0, [1.0]),
                                (SS.IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_QUAD4, 2.0,
                                 [1.0])]),
    "thb_elast_sphere": (lambda: M.make_spec(3, (4, 4, 4), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                             lagrange_order=2, refine_levels=1),
                         [(SS.IWG_STRUC_LINEAR_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 5.0,
                           [0.0, 0.1, 0.0])]),
    "p2_quad4_elast_circle": (lambda: M.make_spec(2, (16, 16), iwg=1, spheres=[[0.5, 0.5, 0.0, 0.25]],
                                                  bspline_order=2),
                              [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_LINE2, 100.0,
                                [0.0, 0.0]),
                               (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_LINE2, 0.0, [0.1, -0.3])]),
    # XTK enrichment: interface facets of two bodies that share a cell, on subphase elements
    "enriched_two_spheres_elast": (lambda: M.make_spec(3, (10, 10, 10), iwg=1, spheres=two_bodies(3, 10),
                                                       material_phase=0, enrich=1),
                                   [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_INTERFACE, SS.GEOM_TRI3,
                                     10.0, [0.01, -0.02, 0.005])]),
}


def _gpu_sets(spec, keep, sides):
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    side_sets = []
    for iwg, where, geom, alpha, value in sides:
        e, x, n = SS.facets(P, where, geom)
        side_sets.append(SS.SideSet(S, iwg, geom, e, x, n, alpha=alpha, value=value))
    return P, S, side_sets


@pytest.mark.parametrize("case", sorted(CASES))
def test_sidesets_match_oracle(case):
    mk, sides = CASES[case]
    spec, keep = mk()
    P, S, side_sets = _gpu_sets(spec, keep, sides)
    assert all(ss.n_facets > 0 for ss in side_sets), [ss.n_facets for ss in side_sets]
    u = np.random.default_rng(SEED + 90).uniform(-1.0, 1.0, S.n_rows)
    vals, R = SS.assemble(S, side_sets, u)
    Q, s, vo, Ro = oracle_side_system(spec, sides, u)
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    print(case, "facets", [ss.n_facets for ss in side_sets], "|dK|/|K| %.2e |dR|/|R| %.2e" % (rel(vals, vo), rel(R, Ro)))
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12
    v2, R2 = SS.assemble(S, side_sets, u)
    assert np.array_equal(v2.view(np.uint64), vals.view(np.uint64)) and np.array_equal(R2.view(np.uint64), R.view(np.uint64))


@pytest.mark.parametrize("case", sorted(GENERIC_CASES))
def test_generic_sidesets_match_oracle(case):
    mk, sides = GENERIC_CASES[case]
    spec, keep = mk()
    P, S, side_sets = _gpu_sets(spec, keep, sides)
    assert all(ss.n_facets > 0 for ss in side_sets), [ss.n_facets for ss in side_sets]
    u = np.random.default_rng(SEED + 91).uniform(-1.0, 1.0, S.n_rows)
    vals, R = SS.assemble(S, side_sets, u)
    Q, s, vo, Ro = oracle_side_system(spec, sides, u)
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    print(case, "facets", [ss.n_facets for ss in side_sets], "|dK|/|K| %.2e |dR|/|R| %.2e" % (rel(vals, vo), rel(R, Ro)))
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12
    v2, R2 = SS.assemble(S, side_sets, u)
    assert np.array_equal(v2.view(np.uint64), vals.view(np.uint64)) and np.array_equal(R2.view(np.uint64), R.view(np.uint64))


def test_known_answer_xtk_cut_diffusion_5_plus_20z_on_gpu():
    spec, keep = M.make_spec(3, (1, 1, 4), hi=(1.0, 1.0, 4.0), ls_kind=M.LS_PLANE, plane=[0.0, 0.0, 1.0, -3.1],
                             material_phase=0, k_sigma=0.0)
    sides = [(SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_ZMIN, SS.GEOM_QUAD4, 1.0, [5.0]),
             (SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 0.0, [20.0])]
    P, S, side_sets = _gpu_sets(spec, keep, sides)
    assert S.n_rows == 20
    vals, R0 = SS.assemble(S, side_sets, np.zeros(S.n_rows))
    rp, ci = S.pattern()
    K = sp.csr_matrix((vals, ci, rp), shape=(S.n_rows, S.n_rows))
    T = spla.spsolve(K.tocsc(), -R0)
    z = np.repeat(np.arange(5.0), 4)
    assert np.abs(T - (5.0 + 20.0 * z)).max() < 1e-8, T
    # the same system (non-symmetric: unsymmetric Nitsche) solved on the GPU by GMRES
    Tg, info = SV.gmres(SV.DeviceCSR(rp, ci, vals), -R0, rtol=1e-14, restart=30)
    assert info.converged == 1
    assert np.abs(Tg - (5.0 + 20.0 * z)).max() < 1e-8, Tg


def test_c1_weak_boundary_conditions_displacement():
    """C1 as configured: 64 x 64 plane stress, circle hole r = 0.25, left edge clamped weakly,
    unit traction on the right edge."""
    spec, keep = M.make_spec(2, (64, 64), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.0, 0.25]], seed=SEED)
    sides = [(SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, SS.SIDE_XMIN, SS.GEOM_LINE2, 100.0, [0.0, 0.0]),
             (SS.IWG_STRUC_LINEAR_NEUMANN, SS.SIDE_XMAX, SS.GEOM_LINE2, 0.0, [1.0, 0.0])]
    P, S, side_sets = _gpu_sets(spec, keep, sides)
    u0 = np.zeros(S.n_rows)
    Q, s, vo, Ro = oracle_side_system(spec, sides, u0)
    rp, ci = S.pattern()

    def solve(vals, R0):
        A = SV.DeviceCSR(rp, ci, vals)
        x, info = SV.cg(A, -R0, rtol=1e-13, max_iter=100000)
        assert info.converged == 1
        return x, info

    u_ref, info = solve(vo, Ro)
    out = {}
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        S.options(prec, M.QUAD_SUBDIVISION)
        vals, R0 = SS.assemble(S, side_sets, u0)
        u, inf = solve(vals, R0)
        out[prec] = (rel(vals, vo), rel(R0, Ro), rel(u, u_ref), inf.iterations)
    print("C1 weak BCs: oracle cg iterations", info.iterations, "fp64 / mixed (|dK|, |dR|, |du|, its)", out)
    assert out[M.PREC_FP64][0] <= 1e-12 and out[M.PREC_FP64][1] <= 1e-12 and out[M.PREC_FP64][2] <= 1e-9
    assert out[M.PREC_MIXED][0] <= 1e-6 and out[M.PREC_MIXED][2] <= 1e-5
