"""The oracle against what the mathematics fixes, so that a plausible mistake in it (a dropped term, a
wrong sign or index, a transposed operand, a wrong weight or Jacobian) fails a test:

* the unit-cube hex8 Laplace stiffness (tensor product of 1D stiffness/mass, closed form);
* strain energy of affine fields on cut meshes: u^T K u = (eps : D : eps) * V_material and
  k |g|^2 V_material, with V_material of a planar cut in closed form (half-space inside a box);
* patch tests on cut meshes (uniform flux / uniform stress with zero traction on the planar
  interface) solved with Dirichlet data on the box boundary -- the analogue of the reference's XTK-cut
  known-answer test T = 5 + 20 z (projects/FEM/MDL/test/UT_MDL_XTK_DiffusionElement.cpp:91-340);
* R(u) computed through the flux (IWG::compute_residual) equals K u (compute_jacobian) -- the
  linear-IWG form of IWG::check_jacobian (projects/FEM/INT/src/IWG/cl_FEM_IWG.cpp:2498-2610);
* rigid-body null space, symmetry, volume conservation of the XTK decomposition
  (projects/XTK/test/xtk/UT_XTK_Cut_Mesh_RegSub.cpp:172-175) and O(h^2) convergence of the sphere
  volume;
* the CSR graph equals the brute-force union of element topologies.
"""
from itertools import product
from math import factorial

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import moris_amd as M
from helpers import adof_coords, csr_matvec, csr_to_scipy, rel, spec_sphere3
from oracle import py as O


def _plane_spec(dim, n, normal, d, iwg, E=1.0, nu=0.3):
    nrm = np.asarray(normal, float)
    nrm = nrm / np.linalg.norm(nrm)
    plane = list(nrm) + [0.0] * (3 - dim) + [d]
    return M.make_spec(dim, (n,) * dim, iwg=iwg, ls_kind=M.LS_PLANE, plane=plane, E=E, nu=nu, k_sigma=0.0), nrm


def _halfspace_volume(nrm, d):
    """|{x in [0,1]^dim : n.x + d > 0}| for n with positive components (closed form)."""
    dim = len(nrm)
    below = 0.0
    for v in product((0, 1), repeat=dim):
        t = -d - np.dot(nrm, v)
        below += (-1) ** sum(v) * max(0.0, t) ** dim
    below /= factorial(dim) * np.prod(nrm)
    return 1.0 - below


def _sub_volumes(setup, dim):
    nv = dim + 1
    X = setup["ig_xi"].reshape(-1, nv, dim)
    J = X[:, :dim, :] - X[:, dim:dim + 1, :] if dim == 2 else X[:, [0, 1, 3], :] - X[:, 2:3, :]
    fac = 2.0 if dim == 2 else 6.0
    return np.linalg.det(J) / fac   # in parametric units of the IP cell


def _oracle(spec_keep):
    spec, keep = spec_keep
    return O.OracleProblem(spec), spec, keep


def test_hex8_unit_cube_laplace_stiffness_closed_form():
    n = 2
    (spec, keep) = M.make_spec(3, (n, n, n), iwg=M.IWG_SPATIALDIFF_BULK, k_sigma=0.0)
    P = O.OracleProblem(spec)
    u = np.zeros(P.n_rows)
    ids, Ke, Re = O.element(P, 0, u)
    h = 1.0 / n
    # 1D linear element on [0,h]: stiffness [[1,-1],[-1,1]]/h, mass [[2,1],[1,2]] h/6
    k1 = np.array([[1.0, -1.0], [-1.0, 1.0]]) / h
    m1 = np.array([[2.0, 1.0], [1.0, 2.0]]) * h / 6.0
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    Kc = np.zeros((8, 8))
    for a, oa in enumerate(offs):
        for b, ob in enumerate(offs):
            Kc[a, b] = (k1[oa[0], ob[0]] * m1[oa[1], ob[1]] * m1[oa[2], ob[2]] +
                        m1[oa[0], ob[0]] * k1[oa[1], ob[1]] * m1[oa[2], ob[2]] +
                        m1[oa[0], ob[0]] * m1[oa[1], ob[1]] * k1[oa[2], ob[2]])
    # oracle K_e is in sorted adof order: vertex order 0,1,3,2,4,5,7,6 on the lattice
    perm = [0, 1, 3, 2, 4, 5, 7, 6]
    assert np.max(np.abs(Ke - Kc[np.ix_(perm, perm)])) < 1e-14


@pytest.mark.parametrize("dim,iwg", [(3, 0), (3, 1), (2, 0), (2, 1)])
def test_affine_field_energy_on_planar_cut(dim, iwg):
    n = 6 if dim == 3 else 12
    (spec, keep), nrm = _plane_spec(dim, n, [0.3, 0.5, 0.8][:dim], -0.45 * sum([0.3, 0.5, 0.8][:dim]) /
                                    np.linalg.norm([0.3, 0.5, 0.8][:dim]), iwg)
    P = O.OracleProblem(spec)
    s = P.setup()
    V = _halfspace_volume(nrm, spec.plane[3])
    _, X = adof_coords(s, dim, (n,) * dim)
    rp, ci = s["row_ptr"], s["col_idx"]
    E, nu = 1.0, 0.3
    if iwg == 0:
        g = np.array([0.7, -0.4, 0.2][:dim])
        u = X @ g
        vals, _ = P.assemble(u)
        energy = u @ csr_matvec(rp, ci, vals, u)
        assert abs(energy - g @ g * V) < 1e-11 * max(1.0, g @ g * V)
    else:
        N = P.n_adofs_per_type
        # u_x = a x (normal strain), then u_x = gam * y (engineering shear)
        for case in ("normal", "shear"):
            u = np.zeros(dim * N)
            if case == "normal":
                u[:N] = 0.01 * X[:, 0]
                if dim == 3:
                    expect = 0.01 ** 2 * E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) * V
                else:
                    expect = 0.01 ** 2 * E / (1 - nu ** 2) * V
            else:
                u[:N] = 0.01 * X[:, 1]
                expect = 0.01 ** 2 * E / (2 * (1 + nu)) * V
            vals, _ = P.assemble(u)
            energy = u @ csr_matvec(rp, ci, vals, u)
            assert abs(energy - expect) < 1e-12 * max(expect, 1e-30) + 1e-16, (case, energy, expect)


@pytest.mark.parametrize("dim,iwg", [(3, 0), (3, 1), (2, 0), (2, 1)])
def test_patch_test_on_planar_cut(dim, iwg):
    n = 5 if dim == 3 else 10
    nrm0 = [0.3, 0.5, 0.8][:dim]
    (spec, keep), nrm = _plane_spec(dim, n, nrm0, -0.47 * sum(nrm0) / np.linalg.norm(nrm0), iwg)
    P = O.OracleProblem(spec)
    s = P.setup()
    _, X = adof_coords(s, dim, (n,) * dim)
    N = P.n_adofs_per_type
    on_box = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1.0) < 1e-12), axis=1)
    # a field that satisfies the natural (zero-flux / traction-free) condition on the plane
    t = np.cross(nrm, [1.0, 0.0, 0.0]) if dim == 3 else np.array([-nrm[1], nrm[0]])
    t = t / np.linalg.norm(t)
    if iwg == 0:
        u_ex = X @ (0.8 * t)
        fixed = on_box
    else:
        E, nu, sgm = 1.0, 0.3, 0.02
        I = np.eye(dim)
        eps = (sgm / E) * ((1 + nu) * np.outer(t, t) - nu * I)
        u_ex = np.concatenate([X @ eps[i] for i in range(dim)])
        fixed = np.tile(on_box, dim)
    vals, R = P.assemble(u_ex)
    K = csr_to_scipy(s["row_ptr"], s["col_idx"], vals, P.n_rows)
    free = ~fixed
    # the exact field is in equilibrium at every free row (residual at roundoff level) ...
    scale = abs(K).max() * np.max(np.abs(u_ex))
    assert np.max(np.abs(R[free])) < 1e-13 * scale
    # ... and the Dirichlet solve reproduces it (limited by the conditioning of sliver cut cells)
    rhs = -K[free][:, fixed] @ u_ex[fixed]
    u_free = spla.spsolve(K[free][:, free].tocsc(), rhs)
    assert rel(u_free, u_ex[free]) < 1e-6


@pytest.mark.parametrize("iwg", [0, 1])
def test_residual_equals_jacobian_times_u_and_symmetry(iwg):
    P = O.OracleProblem(spec_sphere3(6, iwg=iwg)[0])
    s = P.setup()
    u = np.random.default_rng(3).uniform(-1, 1, P.n_rows)
    vals, R = P.assemble(u)
    Ku = csr_matvec(s["row_ptr"], s["col_idx"], vals, u)
    assert rel(R, Ku) < 1e-13
    K = csr_to_scipy(s["row_ptr"], s["col_idx"], vals, P.n_rows)
    assert abs(K - K.T).max() <= 1e-14 * abs(K).max()


def test_rigid_body_null_space():
    # diffusion: constants; elasticity: three translations and three rotations
    P = O.OracleProblem(spec_sphere3(5, iwg=0)[0])
    vals, R = P.assemble(np.ones(P.n_rows))
    assert np.max(np.abs(R)) < 1e-13 * np.max(np.abs(vals))
    (spec, keep) = spec_sphere3(5, iwg=1)
    P = O.OracleProblem(spec)
    s = P.setup()
    _, X = adof_coords(s, 3, (5, 5, 5))
    N = P.n_adofs_per_type
    modes = []
    for d in range(3):
        u = np.zeros(3 * N)
        u[d * N:(d + 1) * N] = 1.0
        modes.append(u)
    for (a, b) in [(0, 1), (1, 2), (0, 2)]:
        u = np.zeros(3 * N)
        u[a * N:(a + 1) * N] = -X[:, b]
        u[b * N:(b + 1) * N] = X[:, a]
        modes.append(u)
    for u in modes:
        vals, R = P.assemble(u)
        assert np.max(np.abs(R)) < 1e-13 * np.max(np.abs(vals))


@pytest.mark.parametrize("dim", [2, 3])
def test_decomposition_conserves_volume_and_orientation(dim):
    n = 7
    if dim == 3:
        mk = lambda ph: M.make_spec(3, (n, n, n), spheres=[[0.43, 0.52, 0.47, 0.31]], material_phase=ph)
    else:
        mk = lambda ph: M.make_spec(2, (n, n), spheres=[[0.43, 0.52, 0.0, 0.31]], material_phase=ph)
    vols = {}
    for ph in (0, 1):
        spec, keep = mk(ph)
        P = O.OracleProblem(spec)
        s = P.setup()
        v = _sub_volumes(s, dim)
        assert np.all(v > 0.0)            # templates keep the positive orientation
        per = np.add.reduceat(v, s["ig_ptr"][:-1][np.diff(s["ig_ptr"]) > 0]) if v.size else v
        cells = s["elem_cell"][np.diff(s["ig_ptr"]) > 0]
        vols[ph] = dict(zip(cells.tolist(), per.tolist()))
    full = 2.0 ** dim
    common = set(vols[0]) & set(vols[1])
    assert len(common) > 10
    for c in common:
        assert abs(vols[0][c] + vols[1][c] - full) < 1e-13


def test_sphere_volume_converges_quadratically():
    errs = []
    exact = 1.0 - 4.0 / 3.0 * np.pi * 0.3 ** 3
    for n in (8, 16, 32):
        spec, keep = spec_sphere3(n, k_sigma=0.0)
        P = O.OracleProblem(spec)
        s = P.setup()
        h = 1.0 / n
        cut = np.diff(s["ig_ptr"]) > 0
        V = (np.count_nonzero(~cut) + np.sum(_sub_volumes(s, 3)) / 8.0) * h ** 3
        errs.append(abs(V - exact))
    assert errs[2] < errs[1] < errs[0]
    assert errs[1] / errs[2] > 3.0     # ~4 for O(h^2)


def test_graph_is_union_of_element_topologies():
    spec, keep = spec_sphere3(4, iwg=1)
    P = O.OracleProblem(spec)
    s = P.setup()
    N = P.n_adofs_per_type
    t = s["t_adof"].reshape(-1, 8)
    rows = [set() for _ in range(3 * N)]
    for e in range(t.shape[0]):
        ids = sorted(set(t[e].tolist()))
        full = [a + f * N for f in range(3) for a in ids]
        for r in full:
            rows[r].update(full)
    rp = [0]
    ci = []
    for r in rows:
        ci.extend(sorted(r))
        rp.append(len(ci))
    assert np.array_equal(np.array(rp), s["row_ptr"])
    assert np.array_equal(np.array(ci), s["col_idx"])


def _greville_coords(setup, dim, n):
    """Coordinates of the adofs of a quadratic B-spline set (bspline_order 2): the used splines in
    ascending spline index, each at its Greville abscissa lo + h (s - 1/2) per axis (spline s of an
    axis starts at knot s - 2, so its middle knots are s - 1 and s)."""
    ec = setup["elem_cell"].astype(np.int64)
    nx, ny = n[0], n[1]
    ijk = [ec % nx, (ec // nx) % ny, ec // (nx * ny) if dim == 3 else np.zeros_like(ec)]
    ns = [nn + 2 for nn in n] + [1] * (3 - dim)
    ids = []
    for off in product(range(3), repeat=dim):
        s = [ijk[d] + off[d] for d in range(dim)] + [np.zeros_like(ec)] * (3 - dim)
        ids.append(s[0] + ns[0] * (s[1] + ns[1] * s[2]))
    used = np.unique(np.concatenate(ids))
    sx, sy, sz = used % ns[0], (used // ns[0]) % ns[1], used // (ns[0] * ns[1])
    h = [1.0 / nn for nn in n]
    X = np.stack([(sx - 0.5) * h[0], (sy - 0.5) * h[1]] + ([(sz - 0.5) * h[2]] if dim == 3 else []), axis=1)
    return X


@pytest.mark.parametrize("dim,iwg", [(3, 0), (3, 1), (2, 0), (2, 1)])
def test_quadratic_bspline_affine_energy_on_planar_cut(dim, iwg):
    """p = 2 B-spline T rows (dense T^T K T in the oracle): Greville coefficients reproduce an affine
    field, whose energy on the cut box is the closed form over the half-space volume."""
    n = 5 if dim == 3 else 9
    nrm0 = [0.3, 0.5, 0.8][:dim]
    d = -0.45 * sum(nrm0) / np.linalg.norm(nrm0)
    nrm = np.asarray(nrm0) / np.linalg.norm(nrm0)
    plane = list(nrm) + [0.0] * (3 - dim) + [d]
    spec, keep = M.make_spec(dim, (n,) * dim, iwg=iwg, ls_kind=M.LS_PLANE, plane=plane, bspline_order=2)
    P = O.OracleProblem(spec)
    s = P.setup()
    assert np.allclose(np.add.reduceat(s["t_weight"], s["t_ptr"][:-1]), 1.0)   # partition of unity
    V = _halfspace_volume(nrm, d)
    X = _greville_coords(s, dim, (n,) * dim)
    N = P.n_adofs_per_type
    assert X.shape[0] == N
    rp, ci = s["row_ptr"], s["col_idx"]
    E, nu = 1.0, 0.3
    if iwg == 0:
        g = np.array([0.7, -0.4, 0.2][:dim])
        u = X @ g
        vals, _ = P.assemble(u)
        energy = u @ csr_matvec(rp, ci, vals, u)
        assert abs(energy - g @ g * V) < 1e-11 * g @ g * V
    else:
        u = np.zeros(dim * N)
        u[:N] = 0.01 * X[:, 1]          # engineering shear gamma_xy = 0.01
        expect = 0.01 ** 2 * E / (2 * (1 + nu)) * V
        vals, _ = P.assemble(u)
        energy = u @ csr_matvec(rp, ci, vals, u)
        assert abs(energy - expect) < 1e-11 * expect, (energy, expect)


def test_quadratic_bspline_rigid_rotation_null_space():
    n = 4
    spec, keep = M.make_spec(3, (n, n, n), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2)
    P = O.OracleProblem(spec)
    s = P.setup()
    X = _greville_coords(s, 3, (n, n, n))
    N = P.n_adofs_per_type
    rp, ci = s["row_ptr"], s["col_idx"]
    vals, _ = P.assemble(np.zeros(3 * N))
    w = np.array([0.3, -0.2, 0.5])
    rot = np.cross(w, X - 0.5)
    for u in (np.concatenate([rot[:, 0], rot[:, 1], rot[:, 2]]), np.concatenate([np.ones(N), np.zeros(2 * N)])):
        Ku = csr_matvec(rp, ci, vals, u)
        assert np.linalg.norm(Ku) < 1e-12 * np.abs(vals).max() * np.linalg.norm(u) * np.sqrt(len(u))


def _hex27_node_coords(setup, n):
    """Coordinates of the adofs of a nodal HEX27 set (lagrange_order 2, bspline_order 1): the used
    nodes of the half-step grid in ascending index (x fastest)."""
    ec = setup["elem_cell"].astype(np.int64)
    ijk = [ec % n[0], (ec // n[0]) % n[1], ec // (n[0] * n[1])]
    g = [2 * nn + 1 for nn in n]
    ids = []
    for off in product(range(3), repeat=3):
        q = [2 * ijk[d] + off[d] for d in range(3)]
        ids.append(q[0] + g[0] * (q[1] + g[1] * q[2]))
    used = np.unique(np.concatenate(ids))
    gx, gy, gz = used % g[0], (used // g[0]) % g[1], used // (g[0] * g[1])
    return np.stack([0.5 * gx / n[0], 0.5 * gy / n[1], 0.5 * gz / n[2]], axis=1)


@pytest.mark.parametrize("iwg", [0, 1])
def test_hex27_quadratic_field_energy_on_planar_cut(iwg):
    """HEX27 Lagrange cells reproduce a quadratic field; its energy over the material part of a
    planar cut z < 0.45 (through the cell layer 0.4 < z < 0.6: sub-tets integrated with TET_35,
    whole cells with HEX_3x3x3) is the closed-form integral of a degree-2 integrand."""
    n = 5
    zc = 0.45
    spec, keep = M.make_spec(3, (n, n, n), iwg=iwg, ls_kind=M.LS_PLANE, plane=[0.0, 0.0, -1.0, zc], lagrange_order=2)
    P = O.OracleProblem(spec)
    s = P.setup()
    X = _hex27_node_coords(s, (n, n, n))
    N = P.n_adofs_per_type
    assert X.shape[0] == N and P.n_cut == n * n
    rp, ci = s["row_ptr"], s["col_idx"]
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    if iwg == 0:
        u = x ** 2 + 0.5 * y * z                       # grad u = (2x, z/2, y/2)
        expect = 4.0 / 3.0 * zc + 0.25 * zc ** 3 / 3.0 + 0.25 / 3.0 * zc
    else:
        E, nu = 1.0, 0.3
        u = np.zeros(3 * N)
        u[:N] = 0.01 * y ** 2                          # gamma_xy = 0.02 y, every other strain 0
        expect = E / (2 * (1 + nu)) * 4e-4 / 3.0 * zc
    vals, _ = P.assemble(u)
    energy = u @ csr_matvec(rp, ci, vals, u)
    assert abs(energy - expect) < 1e-12 * expect, (energy, expect)


@pytest.mark.parametrize("iwg", [0, 1])
def test_hex27_quadratic_bspline_affine_energy_on_planar_cut(iwg):
    """HEX27 IP cells with p = 2 B-spline T rows: Greville coefficients reproduce an affine field
    (the B-spline field is triquadratic per cell, which HEX27 interpolates exactly)."""
    n = 4
    nrm0 = [0.3, 0.5, 0.8]
    d = -0.45 * sum(nrm0) / np.linalg.norm(nrm0)
    nrm = np.asarray(nrm0) / np.linalg.norm(nrm0)
    spec, keep = M.make_spec(3, (n, n, n), iwg=iwg, ls_kind=M.LS_PLANE, plane=list(nrm) + [d], bspline_order=2,
                             lagrange_order=2)
    P = O.OracleProblem(spec)
    s = P.setup()
    V = _halfspace_volume(nrm, d)
    X = _greville_coords(s, 3, (n, n, n))
    N = P.n_adofs_per_type
    assert X.shape[0] == N
    rp, ci = s["row_ptr"], s["col_idx"]
    if iwg == 0:
        g = np.array([0.7, -0.4, 0.2])
        u = X @ g
        expect = g @ g * V
    else:
        E, nu = 1.0, 0.3
        u = np.zeros(3 * N)
        u[:N] = 0.01 * X[:, 1]
        expect = 0.01 ** 2 * E / (2 * (1 + nu)) * V
    vals, _ = P.assemble(u)
    energy = u @ csr_matvec(rp, ci, vals, u)
    assert abs(energy - expect) < 1e-11 * expect, (energy, expect)


@pytest.mark.parametrize("bspline_order", [1, 2])
def test_hex27_rigid_motions_null_space(bspline_order):
    n = 3
    spec, keep = M.make_spec(3, (n, n, n), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=bspline_order,
                             lagrange_order=2)
    P = O.OracleProblem(spec)
    s = P.setup()
    X = _greville_coords(s, 3, (n, n, n)) if bspline_order == 2 else _hex27_node_coords(s, (n, n, n))
    N = P.n_adofs_per_type
    rp, ci = s["row_ptr"], s["col_idx"]
    vals, _ = P.assemble(np.zeros(3 * N))
    K = csr_to_scipy(rp, ci, vals, 3 * N)
    assert abs(K - K.T).max() < 1e-13 * np.abs(vals).max()
    w = np.array([0.3, -0.2, 0.5])
    rot = np.cross(w, X - 0.5)
    for u in (np.concatenate([rot[:, 0], rot[:, 1], rot[:, 2]]), np.concatenate([np.ones(N), np.zeros(2 * N)])):
        Ku = csr_matvec(rp, ci, vals, u)
        assert np.linalg.norm(Ku) < 1e-12 * np.abs(vals).max() * np.linalg.norm(u) * np.sqrt(len(u))


def _node_coords(setup, nb):
    """Physical coordinates of every (element, basis) Lagrange node, in T-row order."""
    box = setup["ip_box"].reshape(-1, 6)
    if nb == 27:
        off = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
                        [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1],
                        [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2], [1, 1, 1], [1, 1, 0], [1, 1, 2], [0, 1, 1],
                        [2, 1, 1], [1, 0, 1], [1, 2, 1]], dtype=float)
    else:
        off = 2.0 * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    return (box[:, None, :3] + 0.5 * off[None, :, :] * box[:, None, 3:]).reshape(-1, 3)


def _t_dense(setup, n_adofs):
    tp, ta, tw = setup["t_ptr"], setup["t_adof"], setup["t_weight"]
    T = np.zeros((len(tp) - 1, n_adofs))
    for r in range(len(tp) - 1):
        for q in range(tp[r], tp[r + 1]):
            T[r, ta[q]] += tw[q]
    return T


@pytest.mark.parametrize("iwg", [0, 1])
def test_thb_truncated_basis_affine_energy_on_planar_cut(iwg):
    """One THB level on the cells the plane intersects, HEX27 cells: the truncated hierarchical
    quadratic B-splines are a partition of unity and reproduce linear fields (their coefficients are
    each level's Greville abscissae, which truncation preserves); the energy of an affine field over
    the cut box is then the closed form over the half-space volume."""
    n = 4
    nrm0 = [0.3, 0.5, 0.8]
    d = -0.45 * sum(nrm0) / np.linalg.norm(nrm0)
    nrm = np.asarray(nrm0) / np.linalg.norm(nrm0)
    spec, keep = M.make_spec(3, (n, n, n), iwg=iwg, ls_kind=M.LS_PLANE, plane=list(nrm) + [d], bspline_order=2,
                             lagrange_order=2, refine_levels=1)
    P = O.OracleProblem(spec)
    s = P.setup()
    N = P.n_adofs_per_type
    T = _t_dense(s, N)
    X = _node_coords(s, 27)
    assert np.abs(T.sum(axis=1) - 1.0).max() < 1e-14                      # partition of unity
    # adof coordinates (the truncated functions are independent on the box, not necessarily on its
    # material cells: near the refinement boundary more than 27 are non-zero on one cell)
    G = np.linalg.lstsq(T, X, rcond=None)[0]
    assert np.abs(T @ G - X).max() < 1e-13                                 # linear reproduction
    h = s["ip_box"].reshape(-1, 6)[:, 3]
    assert (h < 0.75 / n).sum() > 0 and (h > 0.75 / n).sum() > 0           # both levels present
    V = _halfspace_volume(nrm, d)
    rp, ci = s["row_ptr"], s["col_idx"]
    if iwg == 0:
        g = np.array([0.7, -0.4, 0.2])
        u = G @ g
        expect = g @ g * V
    else:
        u = np.zeros(3 * N)
        u[:N] = 0.01 * G[:, 1]
        expect = 0.01 ** 2 * 1.0 / (2 * 1.3) * V
    vals, _ = P.assemble(u)
    energy = u @ csr_matvec(rp, ci, vals, u)
    assert abs(energy - expect) < 1e-11 * expect, (energy, expect)


def test_thb_rigid_motions_null_space():
    n = 4
    spec, keep = M.make_spec(3, (n, n, n), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                             lagrange_order=2, refine_levels=1)
    P = O.OracleProblem(spec)
    s = P.setup()
    N = P.n_adofs_per_type
    G = np.linalg.lstsq(_t_dense(s, N), _node_coords(s, 27), rcond=None)[0]
    rp, ci = s["row_ptr"], s["col_idx"]
    vals, _ = P.assemble(np.zeros(3 * N))
    rot = np.cross(np.array([0.3, -0.2, 0.5]), G - 0.5)
    for u in (np.concatenate([rot[:, 0], rot[:, 1], rot[:, 2]]), np.concatenate([np.ones(N), np.zeros(2 * N)])):
        Ku = csr_matvec(rp, ci, vals, u)
        assert np.linalg.norm(Ku) < 1e-12 * np.abs(vals).max() * np.linalg.norm(u) * np.sqrt(len(u))


@pytest.mark.parametrize("lagrange_order", [1, 2])
def test_quadratic_bspline_rows_reproduce_quadratics(lagrange_order):
    """p = 2 T rows at the Lagrange nodes reproduce quadratic fields.  The coefficient of x^2 on the
    uniform quadratic B-spline with knots t, t+h, t+2h, t+3h is the product of its two interior knots
    (the blossom of x^2), (t+h)(t+2h); that of xy the product of the two axes' Greville abscissae.
    Partition of unity and linear reproduction leave the weights at a cell-centre node as the family
    (a, 1-2a, a); only this check fixes a = 1/8 (and 1/2, 1/2 at a lattice node)."""
    n = 4
    spec, keep = M.make_spec(3, (n, n, n), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                             lagrange_order=lagrange_order)
    P = O.OracleProblem(spec)
    s = P.setup()
    nb = 27 if lagrange_order == 2 else 8
    N = P.n_adofs_per_type
    T = _t_dense(s, N)
    Xn = _node_coords(s, nb)                    # every (element, basis) node, T-row order
    G = _greville_coords(s, 3, (n, n, n))       # adof Greville abscissae (s - 1/2) h
    h = 1.0 / n
    knot_lo = G - 1.5 * h                       # first knot t of each spline
    c2 = (knot_lo + h) * (knot_lo + 2 * h)      # interior-knot products per axis
    for d in range(3):
        assert np.abs(T @ c2[:, d] - Xn[:, d] ** 2).max() < 1e-14
    assert np.abs(T @ (G[:, 0] * G[:, 1]) - Xn[:, 0] * Xn[:, 1]).max() < 1e-14
    if lagrange_order == 2:
        # the three non-zero weights of a cell-centre coordinate, per axis
        w = np.sort(np.unique(np.round(s["t_weight"], 15)))
        assert np.any(np.isclose(w, 0.125 ** 3)) and np.any(np.isclose(w, 0.75 ** 3))


def test_thb_truncated_basis_reproduces_quadratics():
    """The truncated hierarchical space contains the level-0 quadratic splines, hence every quadratic
    polynomial: least-squares coefficients reproduce x^2, y z and x y at every Lagrange node of both
    levels to round-off.  A wrong two-scale weight of the truncation breaks this while keeping the
    partition of unity."""
    n = 4
    spec, keep = M.make_spec(3, (n, n, n), iwg=1, spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                             lagrange_order=2, refine_levels=1)
    P = O.OracleProblem(spec)
    s = P.setup()
    T = _t_dense(s, P.n_adofs_per_type)
    X = _node_coords(s, 27)
    h = s["ip_box"].reshape(-1, 6)[:, 3]
    assert (h < 0.75 / n).sum() > 0 and (h > 0.75 / n).sum() > 0
    for f in (X[:, 0] ** 2, X[:, 1] * X[:, 2], X[:, 0] * X[:, 1], X[:, 2] ** 2):
        c = np.linalg.lstsq(T, f, rcond=None)[0]
        assert np.abs(T @ c - f).max() < 1e-13
