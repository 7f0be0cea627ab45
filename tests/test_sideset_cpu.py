"""Side sets on the CPU: the product's facet extraction against the oracle's own, closed-form facet
measures and normals, properties of the oracle's side-set IWGs, and the reference's known answer.

* Facets (domain boundary faces of material cells, interface facets of material sub-cells): the
  product's list equals the oracle's element for element and vertex for vertex (bit-exact), normals
  within 1e-14.
* Interface of a plane z = c: measure 1 and normal -e_z out of the material z > c; a tilted plane
  through all vertical edges of the unit cube: measure 1 / |n_z|; a sphere: measure -> 4 pi r^2 at
  O(h^2); the material part of a boundary face cut by a sphere centred on it: 1 - pi r^2 at O(h^2).
* Oracle side IWGs (cl_FEM_IWG_Isotropic_Struc_Linear_Dirichlet.cpp:52-240, ..._Neumann.cpp:38-85):
  a Neumann load sums to -t_N times the facet measure per field; the SYMMETRIC Nitsche matrix is
  symmetric, the UNSYMMETRIC one is not and its antisymmetric part is the consistency term; a field
  equal to u_D on the facets has zero Nitsche residual.
* The known answer of UT_MDL_XTK_DiffusionElement.cpp:91-340: 1 x 1 x 4 hex8 lattice cut by the
  plane z = 3.1, material below (phase 0), k = 1, unsymmetric Nitsche T = 5 on z = 0 (alpha 1),
  Neumann flux 20 on the interface: the solved nodal temperatures are 5 + 20 z at every node
  (the reference's gold values 25, 45, 65, 85 and 5).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import moris_amd as M
from helpers import NV, SEED, csr_matvec, oracle_desc as _desc_of, oracle_side_system as _oracle_system
from moris_amd import sideset as SS
from oracle import py as O


CASES = {
    "sphere3_interface": (lambda: M.make_spec(3, (8, 8, 8), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]]),
                          SS.SIDE_INTERFACE, SS.GEOM_TRI3),
    "plane3_interface": (lambda: M.make_spec(3, (5, 5, 5), ls_kind=M.LS_PLANE, plane=[0.3, 0.5, 0.81, -0.6]),
                         SS.SIDE_INTERFACE, SS.GEOM_TRI3),
    "sphere_on_face_xmin_tri": (lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.0, 0.5, 0.5, 0.3]]), SS.SIDE_XMIN,
                                SS.GEOM_TRI3),
    "sphere_on_face_xmin_quad": (lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.0, 0.5, 0.5, 0.3]]), SS.SIDE_XMIN,
                                 SS.GEOM_QUAD4),
    "circle2_interface": (lambda: M.make_spec(2, (16, 16), iwg=1, spheres=[[0.5, 0.5, 0.0, 0.25]]),
                          SS.SIDE_INTERFACE, SS.GEOM_LINE2),
    "circle2_ymax": (lambda: M.make_spec(2, (12, 12), iwg=1, spheres=[[0.5, 1.0, 0.0, 0.3]]), SS.SIDE_YMAX,
                     SS.GEOM_LINE2),
}


def _measure(P, elem, xi):
    """Physical measure of every facet (triangle fan / segment) from the element boxes."""
    a = P.arrays()
    dim = P.dim
    box = a["ip_box"].reshape(-1, 2 * dim)[elem]
    X = box[:, None, :dim] + 0.5 * (xi + 1.0) * box[:, None, dim:]
    if xi.shape[1] == 2:
        return np.linalg.norm(X[:, 1] - X[:, 0], axis=1)
    m = np.zeros(len(elem))
    for t in range(xi.shape[1] - 2):
        m += 0.5 * np.linalg.norm(np.cross(X[:, t + 1] - X[:, 0], X[:, t + 2] - X[:, 0]), axis=1)
    return m


@pytest.mark.parametrize("case", sorted(CASES))
def test_facets_match_oracle(case):
    mk, where, geom = CASES[case]
    spec, keep = mk()
    P = M.Problem(spec, keep)
    e, x, n = SS.facets(P, where, geom)
    Q = O.OracleProblem(spec)
    eo, xo, no = O.facets(Q, where, NV[geom])
    assert len(e) > 0, case
    assert np.array_equal(e, eo)
    assert np.array_equal(x.view(np.uint64), xo.view(np.uint64))
    assert np.abs(n - no).max() <= 1e-14
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)


def test_plane_interface_measure_and_normal():
    for c in (0.37, 0.6):
        spec, keep = M.make_spec(3, (4, 4, 4), ls_kind=M.LS_PLANE, plane=[0.0, 0.0, 1.0, -c])
        P = M.Problem(spec, keep)
        e, x, n = SS.facets(P, SS.SIDE_INTERFACE, SS.GEOM_TRI3)
        assert abs(_measure(P, e, x).sum() - 1.0) < 1e-13
        assert np.abs(n - [0.0, 0.0, -1.0]).max() < 1e-14      # out of the material z > c
    nrm = np.array([0.3, 0.2, 0.9])
    nrm /= np.linalg.norm(nrm)
    spec, keep = M.make_spec(3, (5, 5, 5), ls_kind=M.LS_PLANE, plane=list(nrm) + [-0.5 * nrm.sum()])
    P = M.Problem(spec, keep)
    e, x, n = SS.facets(P, SS.SIDE_INTERFACE, SS.GEOM_TRI3)
    assert abs(_measure(P, e, x).sum() - 1.0 / nrm[2]) < 1e-13
    assert np.abs(n + nrm).max() < 1e-13


def test_sphere_and_boundary_measures_converge():
    err_s, err_b = [], []
    for n in (8, 16, 32):
        spec, keep = M.make_spec(3, (n, n, n), spheres=[[0.5, 0.5, 0.5, 0.3]])
        P = M.Problem(spec, keep)
        e, x, _ = SS.facets(P, SS.SIDE_INTERFACE, SS.GEOM_TRI3)
        err_s.append(abs(_measure(P, e, x).sum() - 4 * np.pi * 0.09))
        spec, keep = M.make_spec(3, (n, n, n), spheres=[[0.0, 0.5, 0.5, 0.3]])
        P = M.Problem(spec, keep)
        area = 0.0
        for geom in (SS.GEOM_TRI3, SS.GEOM_QUAD4):
            e, x, nr = SS.facets(P, SS.SIDE_XMIN, geom)
            assert np.all(nr == [-1.0, 0.0, 0.0])
            area += _measure(P, e, x).sum()
        err_b.append(abs(area - (1.0 - np.pi * 0.09)))
    for err in (err_s, err_b):
        assert err[1] < 0.4 * err[0] and err[2] < 0.4 * err[1], err


@pytest.mark.parametrize("iwg,field", [(SS.IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE, 0),
                                       (SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE, 1)])
def test_oracle_nitsche_symmetry_and_consistency(iwg, field):
    spec, keep = M.make_spec(3, (4, 4, 4), iwg=field, spheres=[[0.5, 0.5, 0.5, 0.3]])
    Q = O.OracleProblem(spec)
    s = Q.setup()
    d, k = _desc_of(Q, s)
    rp, ci = s["row_ptr"], s["col_idx"]
    nr = Q.n_rows
    e, x, n = O.facets(Q, SS.SIDE_INTERFACE, 3)
    value = [0.7, -0.2, 0.4][: (3 if field else 1)]
    for unsym in (False, True):
        code = iwg + (1 if unsym else 0)
        vals, R = np.zeros(Q.nnz), np.zeros(nr)
        O.side_assemble(d, rp, ci, code, 3, e, x, n, 2.0, value, np.zeros(nr), vals, R)
        K = sp.csr_matrix((vals, ci, rp), shape=(nr, nr))
        asym = abs(K - K.T).max() / abs(K).max()
        assert (asym > 1e-3) if unsym else (asym < 1e-14)
        # the field equal to u_D everywhere has zero Nitsche residual: R(u) = K u + R(0)
        N = nr // (3 if field else 1)
        u = np.concatenate([np.full(N, v) for v in value])
        v2, R2 = np.zeros(Q.nnz), np.zeros(nr)
        O.side_assemble(d, rp, ci, code, 3, e, x, n, 2.0, value, u, v2, R2)
        assert np.abs(R2).max() < 1e-13 * np.abs(R).max()
        assert np.abs(csr_matvec(rp, ci, vals, u) + R).max() < 1e-13 * np.abs(R).max()


@pytest.mark.parametrize("field", [0, 1])
def test_oracle_neumann_load_sums_to_flux_times_area(field):
    spec, keep = M.make_spec(3, (6, 6, 6), iwg=field, ls_kind=M.LS_PLANE, plane=[0.0, 0.0, 1.0, -0.43])
    Q = O.OracleProblem(spec)
    s = Q.setup()
    d, k = _desc_of(Q, s)
    e, x, n = O.facets(Q, SS.SIDE_INTERFACE, 3)
    t = [1.5, -0.5, 2.0][: (3 if field else 1)]
    nr = Q.n_rows
    vals, R = np.zeros(Q.nnz), np.zeros(nr)
    O.side_assemble(d, s["row_ptr"], s["col_idx"], SS.IWG_STRUC_LINEAR_NEUMANN if field else SS.IWG_SPATIALDIFF_NEUMANN,
                    3, e, x, n, 0.0, t, np.zeros(nr), vals, R)
    assert np.abs(vals).max() == 0.0
    N = nr // len(t)
    for f, tf in enumerate(t):
        assert abs(R[f * N:(f + 1) * N].sum() + tf * 1.0) < 1e-13


def test_known_answer_xtk_cut_diffusion_5_plus_20z():
    spec, keep = M.make_spec(3, (1, 1, 4), hi=(1.0, 1.0, 4.0), ls_kind=M.LS_PLANE, plane=[0.0, 0.0, 1.0, -3.1],
                             material_phase=0, k_sigma=0.0)
    sides = [(SS.IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE, SS.SIDE_ZMIN, SS.GEOM_QUAD4, 1.0, [5.0]),
             (SS.IWG_SPATIALDIFF_NEUMANN, SS.SIDE_INTERFACE, SS.GEOM_TRI3, 0.0, [20.0])]
    Q = O.OracleProblem(spec)
    nr = Q.n_rows
    Q, s, vals, R0 = _oracle_system(spec, sides, np.zeros(nr))
    K = sp.csr_matrix((vals, s["col_idx"], s["row_ptr"]), shape=(nr, nr))
    T = spla.spsolve(K.tocsc(), -R0)
    # adofs are the 20 lattice nodes in ascending order (x fastest): z of node j
    z = np.repeat(np.arange(5.0), 4)
    assert nr == 20
    assert np.abs(T - (5.0 + 20.0 * z)).max() < 1e-8, T


GENERIC = {
    "p2_hex8_diff": lambda: M.make_spec(3, (5, 5, 5), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.4, seed=SEED,
                                        bspline_order=2),
    "p2_hex27_elast": lambda: M.make_spec(3, (4, 4, 4), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                          lagrange_order=2),
    "thb_elast": lambda: M.make_spec(3, (4, 4, 4), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                     lagrange_order=2, refine_levels=1),
    "p2_quad4_elast": lambda: M.make_spec(2, (8, 8), iwg=1, spheres=[[0.5, 0.5, 0.0, 0.25]], bspline_order=2),
}


@pytest.mark.parametrize("case", sorted(GENERIC))
def test_oracle_side_sets_on_generic_rows(case):
    """Side IWGs over T rows with several entries (p = 2 B-splines, HEX27 cells, THB): a Neumann
    load sums to -t_N times the interface measure per field (partition of unity of the splines);
    the SYMMETRIC Nitsche matrix is symmetric, the UNSYMMETRIC one is not; a field equal to u_D
    (the same coefficient on every spline) has zero Nitsche residual."""
    spec, keep = GENERIC[case]()
    Q = O.OracleProblem(spec)
    s = Q.setup()
    d, k = _desc_of(Q, s)
    rp, ci = s["row_ptr"], s["col_idx"]
    nr = Q.n_rows
    dim = spec.dim
    nvs = 3 if dim == 3 else 2
    nf = dim if spec.iwg == 1 else 1
    e, x, n = O.facets(Q, SS.SIDE_INTERFACE, nvs)
    assert len(e) > 0
    # interface measure from the facets and their elements' boxes
    box = s["ip_box"].reshape(-1, 2 * dim)[e]
    X = box[:, None, :dim] + 0.5 * (x + 1.0) * box[:, None, dim:]
    if nvs == 2:
        area = np.linalg.norm(X[:, 1] - X[:, 0], axis=1).sum()
    else:
        area = 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1).sum()
    t = [1.5, -0.5, 2.0][:nf]
    neu = SS.IWG_STRUC_LINEAR_NEUMANN if nf > 1 else SS.IWG_SPATIALDIFF_NEUMANN
    vals, R = np.zeros(Q.nnz), np.zeros(nr)
    O.side_assemble(d, rp, ci, neu, nvs, e, x, n, 0.0, t, np.zeros(nr), vals, R)
    N = nr // nf
    for f, tf in enumerate(t):
        assert abs(R[f * N:(f + 1) * N].sum() + tf * area) < 1e-12 * max(1.0, area)
    sym = SS.IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE if nf > 1 else SS.IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE
    value = [0.7, -0.2, 0.4][:nf]
    u = np.concatenate([np.full(N, v) for v in value])
    for unsym in (False, True):
        code = sym + (1 if unsym else 0)
        vals, R = np.zeros(Q.nnz), np.zeros(nr)
        O.side_assemble(d, rp, ci, code, nvs, e, x, n, 2.0, value, np.zeros(nr), vals, R)
        K = sp.csr_matrix((vals, ci, rp), shape=(nr, nr))
        asym = abs(K - K.T).max() / abs(K).max()
        assert (asym > 1e-3) if unsym else (asym < 1e-13), (unsym, asym)
        v2, R2 = np.zeros(Q.nnz), np.zeros(nr)
        O.side_assemble(d, rp, ci, code, nvs, e, x, n, 2.0, value, u, v2, R2)
        assert np.abs(R2).max() < 1e-12 * np.abs(R).max()


def test_hex27_facets_equal_hex8_facets():
    """The facets of HEX27 problems are those of the same cut on HEX8 cells (the cut is that of the
    cell box)."""
    for where, geom in ((SS.SIDE_INTERFACE, SS.GEOM_TRI3), (SS.SIDE_XMIN, SS.GEOM_QUAD4), (SS.SIDE_XMIN, SS.GEOM_TRI3)):
        got = []
        for lag in (1, 2):
            spec, keep = M.make_spec(3, (5, 5, 5), iwg=1, spheres=[[0.0, 0.5, 0.5, 0.35]], bspline_order=2,
                                     lagrange_order=lag)
            P = M.Problem(spec, keep)
            got.append(SS.facets(P, where, geom))
        for a, b in zip(*got):
            assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))

