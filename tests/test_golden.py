"""Oracle building blocks against the reference's own fixtures and against closed forms.

Shape functions: projects/FEM/INT/test/cl_FEM_Interpolation_Function_Lagrange_Hex8.cpp:95-177 (same
fixtures, same 1e-12 tolerance).  Quadrature: projects/FEM/INT/test/cl_FEM_Integrator.cpp:28-46
(integrate_*.bin test polynomials, 1e-12), plus exactness of the simplex rules on monomials, whose
integrals over the unit simplex are a!b!c!/(a+b+c+d)!.
"""
from math import factorial

import numpy as np
import pytest

from helpers import load_bin
from oracle import py as O


@pytest.mark.parametrize("name,dim", [("quad4", 2), ("hex8", 3)])
def test_shape_functions_match_reference_fixtures(name, dim):
    xi = load_bin("points_%dd" % dim).T                  # [n][dim]
    phihat = load_bin("lagrange_%s_phihat" % name)[:, 0]  # nodal values
    phi = load_bin("lagrange_%s_phi" % name)[0]          # [n]
    dphi = load_bin("lagrange_%s_dphidxi" % name)        # [dim][n]
    N, dN = O.shape(dim, xi)
    assert np.all(np.abs(N.sum(axis=1) - 1.0) < 1e-12)   # partition of unity
    assert np.max(np.abs(N @ phihat - phi)) < 1e-12
    assert np.max(np.abs(np.einsum("kdb,b->dk", dN, phihat) - dphi)) < 1e-12


def _test_polynomial(name):
    """Integrator_Test_Polynomial file layout (cl_MTK_Integrator_Test_Polynomial.cpp:18-58)."""
    d = load_bin(name)[:, 0]
    ndim, ncoef, integral = int(d[0]), int(d[1]), d[2]
    coeffs = d[3:3 + ncoef]
    expo = d[3 + ncoef:3 + ncoef + ncoef * ndim].reshape(ncoef, ndim)
    return ndim, coeffs, expo, integral


@pytest.mark.parametrize("fixture,rule", [("integrate_quad_2x2", "QUAD_2x2"), ("integrate_hex_2x2x2", "HEX_2x2x2"),
                                          ("integrate_hex_3x3x3", "HEX_3x3x3")])
def test_gauss_rules_integrate_reference_polynomials(fixture, rule):
    ndim, coeffs, expo, integral = _test_polynomial(fixture)
    xi, w = O.rule(rule)
    assert xi.shape[1] == ndim
    vals = np.array([np.sum(coeffs * np.prod(np.power(x[None, :], expo), axis=1)) for x in xi])
    assert abs(np.dot(w, vals) - integral) < 1e-12


def _simplex_monomials(dim, max_deg):
    if dim == 2:
        return [(a, b) for a in range(max_deg + 1) for b in range(max_deg + 1 - a)]
    return [(a, b, c) for a in range(max_deg + 1) for b in range(max_deg + 1 - a) for c in range(max_deg + 1 - a - b)]


@pytest.mark.parametrize("rule,dim,degree", [("TRI_7", 2, 5), ("TET_11", 3, 4), ("TET_35", 3, 6)])
def test_simplex_rules_exact_to_their_degree(rule, dim, degree):
    xi, w = O.rule(rule)
    vol = 0.5 if dim == 2 else 1.0 / 6.0
    # weights are normalised to sum to 1 (detJ of the unit simplex carries 1/2 or 1/6)
    assert abs(w.sum() - 1.0) < 1e-14
    for m in _simplex_monomials(dim, degree):
        exact = np.prod([factorial(k) for k in m]) / factorial(sum(m) + dim)
        approx = vol * np.sum(w * np.prod(xi ** np.array(m)[None, :], axis=1))
        assert abs(approx - exact) < 2e-15, (rule, m, approx, exact)
    # and not exact one degree higher (the rule is what it claims, not a higher one)
    errs = []
    for m in _simplex_monomials(dim, degree + 1):
        if sum(m) == degree + 1:
            exact = np.prod([factorial(k) for k in m]) / factorial(sum(m) + dim)
            approx = vol * np.sum(w * np.prod(xi ** np.array(m)[None, :], axis=1))
            errs.append(abs(approx - exact))
    assert max(errs) > 1e-8


def test_hex27_shape_functions_match_reference_fixtures():
    """HEX27 against projects/FEM/INT/test/data/lagrange_hex27_{phihat,phi,dphidxi}.bin (the fixtures of
    cl_FEM_Interpolation_Function_Lagrange_Hex27.cpp, 1e-12), plus N_a(node_b) = delta_ab."""
    xi = load_bin("points_3d").T
    phihat = load_bin("lagrange_hex27_phihat")[:, 0]
    phi = load_bin("lagrange_hex27_phi")[0]
    dphi = load_bin("lagrange_hex27_dphidxi")
    N, dN = O.shape(3, xi, nb=27)
    assert np.all(np.abs(N.sum(axis=1) - 1.0) < 1e-12)
    assert np.all(np.abs(dN.sum(axis=2)) < 1e-12)
    assert np.max(np.abs(N @ phihat - phi)) < 1e-12
    assert np.max(np.abs(np.einsum("kdb,b->dk", dN, phihat) - dphi)) < 1e-12
    # node order of get_param_coords: corners, edge midpoints, centre, face centres
    nodes = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                      [-1, 1, 1], [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [-1, -1, 0], [1, -1, 0],
                      [1, 1, 0], [-1, 1, 0], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1], [0, 0, 0],
                      [0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]], dtype=float)
    Nn, _ = O.shape(3, nodes, nb=27)
    assert np.max(np.abs(Nn - np.eye(27))) < 1e-15
