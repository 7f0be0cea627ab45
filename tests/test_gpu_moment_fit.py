"""GPU tests of the moment-fitted quadrature of cut HEX27 cells (MGPU_QUAD_MOMENT_FIT): each cut
cell's TET_35 points replaced by the 5 x 5 x 5 Gauss-Legendre nodes weighted by the cell's degree-4
Lagrange moments over those points, which gives the subdivision's sum for every integrand of degree
<= 4 per axis, the box HEX27 stiffness among them.

* against the oracle (TET_35 on every sub-tet) at 1e-12 on the FP64 path and 1e-6 on the mixed
  path: p = 2 B-spline and nodal rows, diffusion and elasticity, one THB level, an enriched set;
  pattern unchanged; repeated calls bit-identical;
* scatter mode (colour classes) to the same bars;
* size-independent on a 16^3 gyroid HEX27 set and a 12^3-base THB set: within 1e-12 of the
  subdivision path on the GPU;
* an element given as 750 sub-tets tiling its box (several staging batches of the weight fit): fit,
  subdivision and oracle within 1e-12, and within 1e-8 of the uncut HEX_3x3x3 element (the TET_35
  sum over the tiling differs from it by 9e-10 on M, computed on the CPU);
* 3D lattice (hex8) sets: the 27 HEX_3x3x3 weights from the degree-2 Lagrange moments; K and R
  against the oracle at 1e-6 (FP32 MFMA product), every cut element accepted, the weights summing
  to each element's material volume; refused in 2D;
* generic-path HEX8 / QUAD4 cells (p = 2 and THB rows): 3 nodes per axis, TET_11 / TRI_7 points,
  against the oracle at 1e-12 (FP64) and 1e-6 (mixed).
"""
import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, rel, two_bodies
from moris_amd import learned as L
from oracle import py as O

pytestmark = pytest.mark.gpu

CASES = {
    "hex27_p2_gyroid_6_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                        ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi * 1.5,
                                                        bspline_order=2, lagrange_order=2),
    "hex27_p2_sphere_6_diffusion": lambda: M.make_spec(3, (6, 5, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5,
                                                       bspline_order=2, lagrange_order=2),
    "hex27_lagrange_sphere_5_elasticity": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                              spheres=[[0.4, 0.5, 0.6, 0.35]], lagrange_order=2),
    "thb_plane_5_hex27_elasticity": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_PLANE,
                                                        plane=[0.3, 0.5, 0.8, -0.6], bspline_order=2,
                                                        lagrange_order=2, refine_levels=1),
    "thb_sphere_4_hex27_diffusion": lambda: M.make_spec(3, (4, 4, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.4,
                                                        bspline_order=2, lagrange_order=2, refine_levels=1),
    "enriched_two_spheres_6_hex27_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                                   spheres=two_bodies(3, 6), material_phase=0,
                                                                   bspline_order=2, lagrange_order=2, enrich=1),
}


def _check_against_oracle(S, spec):
    Q = O.OracleProblem(spec)
    s = Q.setup()
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    u = np.random.default_rng(SEED + 21).uniform(-1.0, 1.0, S.n_rows)
    vo, Ro = Q.assemble(u)
    S.options(M.PREC_FP64, M.QUAD_MOMENT_FIT)
    vals, R = S.compute_jacobian_and_residual(u)
    print("FP64 moment fit vs oracle |dK|/|K| %.2e |dR|/|R| %.2e" % (rel(vals, vo), rel(R, Ro)), flush=True)
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12
    v2, R2 = S.compute_jacobian_and_residual(u)
    assert np.array_equal(v2.view(np.uint64), vals.view(np.uint64)) and np.array_equal(R2.view(np.uint64), R.view(np.uint64))
    S.options(M.PREC_MIXED, M.QUAD_MOMENT_FIT)
    vm, Rm = S.compute_jacobian_and_residual(u)
    print("mixed moment fit vs oracle |dK|/|K| %.2e" % rel(vm, vo), flush=True)
    assert rel(vm, vo) <= 1e-6 and rel(Rm, Ro) <= 1e-6


@pytest.mark.parametrize("case", sorted(CASES))
def test_moment_fit_matches_oracle(case):
    print("case", case, flush=True)
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    _check_against_oracle(S, spec)


@pytest.mark.parametrize("case", ["hex27_p2_gyroid_6_elasticity", "thb_sphere_4_hex27_diffusion"])
def test_moment_fit_scatter_mode_matches_oracle(case, monkeypatch):
    print("case", case, flush=True)
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    monkeypatch.setenv("MGPU_GENERIC_SCATTER", "1")
    S = M.Set(P.desc, keepalive=P)
    monkeypatch.delenv("MGPU_GENERIC_SCATTER")
    _check_against_oracle(S, spec)


@pytest.mark.parametrize("refine", [0, 1])
def test_moment_fit_equals_subdivision_on_larger_sets(refine):
    n = 16 if refine == 0 else 12
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2,
                             refine_levels=refine)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    u = np.random.default_rng(SEED + 22).uniform(-1.0, 1.0, S.n_rows)
    v_sub, R_sub = S.compute_jacobian_and_residual(u)
    S.options(M.PREC_FP64, M.QUAD_MOMENT_FIT)
    v_fit, R_fit = S.compute_jacobian_and_residual(u)
    S.options(M.PREC_MIXED, M.QUAD_MOMENT_FIT)
    v_mix, _ = S.compute_jacobian_and_residual(u)
    print("elements", P.n_elements, "fit vs subdivision %.2e, mixed fit %.2e" % (rel(v_fit, v_sub), rel(v_mix, v_sub)),
          flush=True)
    assert rel(v_fit, v_sub) <= 1e-12 and rel(R_fit, R_sub) <= 1e-12
    assert rel(v_mix, v_sub) <= 1e-6


GENERIC_LINEAR = {
    "p2_sphere_8_hex8_elasticity": lambda: M.make_spec(3, (8, 8, 8), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                       spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2),
    "p2_sphere_10_hex8_diffusion": lambda: M.make_spec(3, (10, 10, 10), spheres=[[0.45, 0.5, 0.55, 0.33]],
                                                       k_sigma=0.5, bspline_order=2),
    "p2_circle_16x12_quad4_plane_stress": lambda: M.make_spec(2, (16, 12), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                              spheres=[[0.5, 0.5, 0.0, 0.3]], bspline_order=2),
    "thb_sphere_6_hex8_elasticity": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                        spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                                        refine_levels=1),
}


@pytest.mark.parametrize("case", sorted(GENERIC_LINEAR))
def test_generic_hex8_quad4_moment_fit_matches_oracle(case):
    """Generic-path HEX8 / QUAD4 cells (p = 2 and THB rows): the 3^dim tensor Gauss points with the
    degree-2 Lagrange moments of the TET_11 / TRI_7 points as weights; K and R against the oracle at
    1e-12 (FP64) and 1e-6 (mixed)."""
    print("case", case, flush=True)
    spec, keep = GENERIC_LINEAR[case]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    _check_against_oracle(S, spec)


def _box_tets(m):
    """The box [-1, 1]^3 as m^3 sub-cubes of 6 tets each (Kuhn split), every tet positively oriented
    in the kernels' convention (rows X_0 - X_2, X_1 - X_2, X_3 - X_2): [m^3 6][4][3]."""
    kuhn = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]
    corner = np.array([[(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1] for v in range(8)], float)
    out = []
    h = 2.0 / m
    for k in range(m):
        for j in range(m):
            for i in range(m):
                base = np.array([-1.0 + i * h, -1.0 + j * h, -1.0 + k * h])
                for t in kuhn:
                    X = base + h * corner[list(t)]
                    J = np.stack([X[0] - X[2], X[1] - X[2], X[3] - X[2]])
                    if np.linalg.det(J) < 0:
                        X = X[[1, 0, 2, 3]]
                    out.append(X)
    return np.array(out)


def test_moment_fit_element_with_many_sub_tets():
    """An element whose box is given as 750 sub-tets (more than one staging batch of the weight fit)
    next to uncut elements: moment fit, subdivision and the oracle agree."""
    spec, keep = M.make_spec(3, (3, 2, 2), iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2, lagrange_order=2)
    P = M.Problem(spec, keep)
    d = M.SetDesc()
    M.ctypes.memmove(M.ctypes.byref(d), M.ctypes.byref(P.desc), M.ctypes.sizeof(M.SetDesc))
    tets = _box_tets(5)
    ig_ptr = np.zeros(P.n_elements + 1, np.int32)
    ig_ptr[1:] = len(tets)                       # element 0 holds every sub-tet, the others none
    ig_xi = np.ascontiguousarray(tets.reshape(-1))
    d.ig_ptr = ig_ptr.ctypes.data_as(M.ctypes.POINTER(M.ctypes.c_int32))
    d.ig_xi = ig_xi.ctypes.data_as(M.ctypes.POINTER(M.ctypes.c_double))
    keepalive = (P, ig_ptr, ig_xi)
    S = M.Set(d, keepalive=keepalive)
    rp, ci = S.pattern()
    u = np.random.default_rng(SEED + 23).uniform(-1.0, 1.0, S.n_rows)
    vo, Ro, _ = O.assemble_desc(d, rp, ci, u)
    v_sub, R_sub = S.compute_jacobian_and_residual(u)
    S.options(M.PREC_FP64, M.QUAD_MOMENT_FIT)
    v_fit, R_fit = S.compute_jacobian_and_residual(u)
    print("750 sub-tets: subdivision vs oracle %.2e, fit vs oracle %.2e" % (rel(v_sub, vo), rel(v_fit, vo)), flush=True)
    assert rel(v_sub, vo) <= 1e-12 and rel(R_sub, Ro) <= 1e-12
    assert rel(v_fit, vo) <= 1e-12 and rel(R_fit, Ro) <= 1e-12
    # the sub-tets tile the box: the cut element's K equals the uncut rule's up to the TET_35
    # error on this degree-10 integrand, and exactly where the box is integrated by HEX_3x3x3
    Su = M.Set(P.desc, keepalive=P)
    v_box, _ = Su.compute_jacobian_and_residual(u)
    assert rel(v_fit, v_box) <= 1e-8


LATTICE = {
    "sphere_16_elasticity": lambda: M.make_spec(3, (16, 16, 16), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                spheres=[[0.45, 0.5, 0.55, 0.31]]),
    "spheres_12_diffusion": lambda: M.make_spec(3, (12, 12, 12), spheres=[[0.3, 0.3, 0.5, 0.2], [0.7, 0.6, 0.5, 0.22]],
                                                k_sigma=0.5),
}


@pytest.mark.parametrize("case", sorted(LATTICE))
def test_lattice_moment_fit_matches_oracle(case):
    """3D lattice sets: the 27 weights from the degree-2 Lagrange moments reproduce the subdivision
    for the trilinear stiffness; K and R against the oracle within the FP32 MFMA product's 1e-6,
    every cut element on the fitted rule, the weights summing to the material's parametric volume."""
    print("case", case, flush=True)
    spec, keep = LATTICE[case]()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Q = O.OracleProblem(spec)
    s = Q.setup()
    u = np.random.default_rng(SEED + 24).uniform(-1.0, 1.0, S.n_rows)
    vo, Ro = Q.assemble(u)
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        S.options(prec, M.QUAD_MOMENT_FIT)
        vals, R = S.compute_jacobian_and_residual(u)
        print(case, "precision", prec, "|dK|/|K| %.2e |dR|/|R| %.2e" % (rel(vals, vo), rel(R, Ro)), flush=True)
        assert rel(vals, vo) <= 1e-6 and rel(R, Ro) <= 1e-6
    st = S.statistics()
    assert st[2] == P.n_cut and st[1] == 0          # every cut element accepted, none on the subdivision
    _, w_fit, flags = L.result(S)
    assert np.all(flags == 1)
    a = P.arrays()
    ig = a["ig_ptr"]
    X = a["ig_xi"].reshape(-1, 4, 3)
    J = np.stack([X[:, 0] - X[:, 2], X[:, 1] - X[:, 2], X[:, 3] - X[:, 2]], axis=1)
    vol = np.abs(np.linalg.det(J)) / 6.0
    csum = np.concatenate([[0.0], np.cumsum(vol)])
    cut = np.flatnonzero(np.diff(ig) > 0)
    v_el = csum[ig[cut + 1]] - csum[ig[cut]]
    assert np.abs(w_fit.sum(axis=1) - v_el).max() <= 1e-5 * 8.0


def test_lattice_moment_fit_refused_in_2d():
    spec, keep = M.make_spec(2, (16, 16), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.0, 0.3]])
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    with pytest.raises(M.MgpuError, match="3D"):
        S.options(M.PREC_MIXED, M.QUAD_MOMENT_FIT)
