"""The mathematics of the moment-fitted quadrature of cut HEX27 cells (MGPU_QUAD_MOMENT_FIT), on the
CPU: the 5 x 5 x 5 Gauss-Legendre nodes xi_g with weights W_g = sum_q w_q L_i(x_q) L_j(y_q) L_k(z_q)
(L the degree-4 Lagrange polynomials through the nodes) give sum_g W_g f(xi_g) = sum_q w_q f(xi_q)
exactly when f has degree <= 4 in each axis.

Pins:
  * the nodes are the roots of the Legendre polynomial P_5 and L_i(x_j) = delta_ij, sum_i L_i = 1;
  * on the oracle's TET_35 points of sliver-to-regular sub-tets of a cell: random polynomials of
    degree <= 4 per axis agree to 1e-14, a degree-5 monomial does not (the bound is sharp);
  * the HEX27 moment integrand dN_a/dxi_i dN_b/dxi_j (the oracle's shape functions) summed over the
    fitted nodes equals its TET_35 sum to 1e-13 for every (a, i, b, j);
  * the lattice (hex8) variant: 3 nodes per axis (the HEX_3x3x3 rule's literal nodes), TET_11
    points: degree <= 2 per axis reproduced, degree 3 not, and the trilinear moment matrix equal.
"""
import numpy as np
from numpy.polynomial import legendre

from oracle import py as O

NODES = np.array([-0.90617984593866399279762687829939, -0.53846931010568309103631442070021, 0.0,
                  0.53846931010568309103631442070021, 0.90617984593866399279762687829939])


def _lagrange5(x):
    x = np.asarray(x, float)
    L = np.ones((len(x), 5))
    for i in range(5):
        for m in range(5):
            if m != i:
                L[:, i] *= (x - NODES[m]) / (NODES[i] - NODES[m])
    return L


def _points():
    """TET_35 points and weights (w detJ) of random sub-tets of [-1, 1]^3, some of them slivers."""
    xi35, w35 = O.rule("TET_35")
    rng = np.random.default_rng(7)
    P, W = [], []
    for t in range(12):
        V = rng.uniform(-1, 1, (4, 3))
        if t % 3 == 0:
            V[3] = V[0] + 1e-3 * rng.uniform(-1, 1, 3) + 0.5 * (V[1] - V[0])   # a sliver
        J = np.stack([V[0] - V[2], V[1] - V[2], V[3] - V[2]])
        det = abs(np.linalg.det(J)) / 6.0
        bary = np.stack([xi35[:, 0], xi35[:, 1], 1.0 - xi35.sum(axis=1), xi35[:, 2]], axis=1)
        P.append(bary @ V)
        W.append(w35 * det)
    return np.concatenate(P), np.concatenate(W)


def _fit(P, w):
    Lx, Ly, Lz = _lagrange5(P[:, 0]), _lagrange5(P[:, 1]), _lagrange5(P[:, 2])
    Wg = np.einsum("q,qi,qj,qk->ijk", w, Lx, Ly, Lz)
    G = np.stack(np.meshgrid(NODES, NODES, NODES, indexing="ij"), axis=-1).reshape(-1, 3)
    return G, Wg.reshape(-1)


def test_nodes_and_lagrange_basis():
    a, b = np.sqrt(5.0 - 2.0 * np.sqrt(10.0 / 7.0)) / 3.0, np.sqrt(5.0 + 2.0 * np.sqrt(10.0 / 7.0)) / 3.0
    assert np.abs(NODES - np.array([-b, -a, 0.0, a, b])).max() <= 2e-16
    assert np.abs(legendre.legval(NODES, [0, 0, 0, 0, 0, 1])).max() <= 1e-15   # roots of P_5
    assert np.abs(_lagrange5(NODES) - np.eye(5)).max() <= 1e-15
    x = np.linspace(-1, 1, 17)
    assert np.abs(_lagrange5(x).sum(axis=1) - 1.0).max() <= 1e-14


def test_fitted_rule_reproduces_tensor_degree_4_and_not_5():
    P, w = _points()
    G, Wg = _fit(P, w)
    rng = np.random.default_rng(8)
    for _ in range(5):
        C = rng.uniform(-1, 1, (5, 5, 5))       # coefficients of x^a y^b z^c, a, b, c <= 4

        def f(X):
            return np.einsum("abc,qa,qb,qc->q", C, X[:, :1] ** np.arange(5), X[:, 1:2] ** np.arange(5),
                             X[:, 2:3] ** np.arange(5))
        exact = w @ f(P)
        assert abs(Wg @ f(G) - exact) <= 1e-14 * max(1.0, abs(exact)) * 10
    # degree 5 in one axis is not reproduced
    g5 = lambda X: X[:, 0] ** 5 * X[:, 1] ** 2
    assert abs(Wg @ g5(G) - w @ g5(P)) > 1e-6 * abs(w @ g5(P))


def test_hex27_moment_integrand_on_the_fitted_nodes():
    P, w = _points()
    G, Wg = _fit(P, w)
    _, dNq = O.shape(3, P, 27)          # [q][3][27]
    _, dNg = O.shape(3, G, 27)
    Aq = dNq.reshape(len(P), 81)
    Ag = dNg.reshape(len(G), 81)
    Mq = Aq.T @ (w[:, None] * Aq)
    Mg = Ag.T @ (Wg[:, None] * Ag)
    assert np.abs(Mg - Mq).max() <= 1e-13 * np.abs(Mq).max()


# ---- 3D lattice sets: 3 nodes per axis (the HEX_3x3x3 rule's), TET_11 points ----------------------

def _lagrange3(x, s):
    x = np.asarray(x, float)
    return np.stack([x * (x - s) / (2 * s * s), 1.0 - x * x / (s * s), x * (x + s) / (2 * s * s)], axis=1)


def test_hex8_fit_on_hex3_nodes_reproduces_tensor_degree_2():
    """With the HEX_3x3x3 rule's literal nodes s = 0.774596669241483, the weights W_g = sum_q w_q
    L_i(x_q) L_j(y_q) L_k(z_q) over TET_11 points reproduce every polynomial of degree <= 2 per
    axis, and hence the trilinear moment matrix sum w g g^T (the oracle's HEX8 shape functions);
    degree 3 in one axis is not reproduced."""
    x3, w3 = O.rule("HEX_3x3x3")
    s = np.abs(x3[:, 0]).max()
    xi11, w11 = O.rule("TET_11")
    rng = np.random.default_rng(9)
    P, W = [], []
    for t in range(10):
        V = rng.uniform(-1, 1, (4, 3))
        J = np.stack([V[0] - V[2], V[1] - V[2], V[3] - V[2]])
        det = abs(np.linalg.det(J)) / 6.0
        bary = np.stack([xi11[:, 0], xi11[:, 1], 1.0 - xi11.sum(axis=1), xi11[:, 2]], axis=1)
        P.append(bary @ V)
        W.append(w11 * det)
    P, W = np.concatenate(P), np.concatenate(W)
    Lx, Ly, Lz = _lagrange3(P[:, 0], s), _lagrange3(P[:, 1], s), _lagrange3(P[:, 2], s)
    Wt = np.einsum("q,qi,qj,qk->ijk", W, Lx, Ly, Lz)
    # the rule's points in its own order: node index per axis (-s, 0, s -> 0, 1, 2)
    idx = np.where(x3 < -0.5 * s, 0, np.where(x3 > 0.5 * s, 2, 1))
    Wg = Wt[idx[:, 0], idx[:, 1], idx[:, 2]]
    C = rng.uniform(-1, 1, (3, 3, 3))
    f = lambda X: np.einsum("abc,qa,qb,qc->q", C, X[:, :1] ** np.arange(3), X[:, 1:2] ** np.arange(3),
                            X[:, 2:3] ** np.arange(3))
    assert abs(Wg @ f(x3) - W @ f(P)) <= 1e-14 * max(1.0, abs(W @ f(P))) * 10
    g3 = lambda X: X[:, 0] ** 3 * X[:, 1]
    assert abs(Wg @ g3(x3) - W @ g3(P)) > 1e-6 * abs(W @ g3(P))
    _, dNq = O.shape(3, P)
    _, dNg = O.shape(3, x3)
    Aq, Ag = dNq.reshape(len(P), 24), dNg.reshape(len(x3), 24)
    Mq = Aq.T @ (W[:, None] * Aq)
    Mg = Ag.T @ (Wg[:, None] * Ag)
    assert np.abs(Mg - Mq).max() <= 1e-13 * np.abs(Mq).max()
