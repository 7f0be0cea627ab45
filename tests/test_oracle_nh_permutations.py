"""Exhaustive sweep of the oracle's node-hierarchy templates on a single IG cell.

Mirrors projects/XTK/test/xtk/UT_XTK_Child_Mesh_NH_Permutations.cpp:616-1290: every sign pattern of
the vertex level sets (vertices on the interface included, which select the one- and two-edge
templates) under every ordering of the vertex global ids (which sets the permutation id through the
order of the intersected edges' keys), with several edge-intersection positions.  For every case:

* the sub-cells of both phases tile the parent: volumes (areas) sum to the parent's;
* every sub-cell has the parent's orientation (positive detJ in the kernels' convention, rows
  X_r - X_2) and is non-degenerate;
* every sub-cell lies in one phase: the parent's linear level set at its vertices has one sign;
* the sub-cell count is 2 / 3 / 4 / 6 for 1 / 2 / 3 / 4 intersected edges (2 / 3 in 2D);
* the child faces on the parent's faces number 6 / 8 / 10 / 12 (the reference's
  tChildFacewithParentFaceRank checks at :1106, :1268, :770, :948) and cover the parent's
  surface; every other child face is shared by exactly two sub-cells (a conforming tiling);
* every new vertex lies on a parent edge where the linear level set vanishes.
"""
from itertools import permutations, product

import numpy as np
import pytest

from oracle import py as O

TET = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
TRI = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
TET_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
TRI_EDGES = [(0, 1), (1, 2), (2, 0)]


def _det(X, dim):
    if dim == 3:
        return np.linalg.det(np.stack([X[0] - X[2], X[1] - X[2], X[3] - X[2]]))
    return np.linalg.det(np.stack([X[0, :2] - X[2, :2], X[1, :2] - X[2, :2]]))


def _bary(P, X, dim):
    """Barycentric coordinates of points P wrt the parent simplex X."""
    A = np.vstack([X[:, :dim].T, np.ones(dim + 1)])
    B = np.vstack([P[:, :dim].T, np.ones(P.shape[0])])
    return np.linalg.solve(A, B).T


def _faces(cell, dim):
    n = dim + 1
    return [tuple(np.delete(np.arange(n), k)) for k in range(n)]


def _key(p):
    return tuple(np.round(p, 12))


def _face_measure(P, dim):
    if dim == 3:
        return 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    return np.linalg.norm(P[1] - P[0])


def _cases(dim):
    nv = dim + 1
    for signs in product((-1, 0, 1), repeat=nv):
        s = np.array(signs)
        nz = s[s != 0]
        if len(nz) == 0 or np.all(nz == nz[0]):
            continue
        n_edges = sum(1 for a, b in (TET_EDGES if dim == 3 else TRI_EDGES) if s[a] * s[b] < 0)
        yield s, n_edges


@pytest.mark.parametrize("dim", [3, 2])
def test_node_hierarchy_all_permutations(dim):
    X = TET if dim == 3 else TRI
    nv = dim + 1
    vol = abs(_det(X, dim)) / (6.0 if dim == 3 else 2.0)
    parent_faces = [np.delete(np.arange(nv), k) for k in range(nv)]
    parent_surface = sum(_face_measure(X[f], dim) for f in parent_faces)
    expect_cells = {1: 2, 2: 3, 3: 4, 4: 6}
    expect_bfaces = {1: 6, 2: 8, 3: 10, 4: 12} if dim == 3 else {1: 4, 2: 5}
    rng = np.random.default_rng(4)
    seen = set()
    n_checked = 0
    for s, n_edges in _cases(dim):
        for order in permutations(range(nv)):
            gid = np.array([10 * (o + 1) for o in order], dtype=np.uint32)
            for trial in range(3):
                mag = rng.uniform(0.15, 1.0, nv)
                f = s * mag
                cells = [O.nh_cell(dim, gid, f, X, ph) for ph in (0, 1)]
                allc = np.concatenate(cells)
                assert allc.shape[0] == expect_cells[n_edges], (s, order)
                # tiling: measure, orientation, non-degeneracy
                dets = np.array([_det(C, dim) for C in allc])
                assert np.all(dets > 1e-3 * abs(_det(X, dim)) * 1e-3), (s, order, dets)
                assert abs(dets.sum() / (6.0 if dim == 3 else 2.0) - vol) < 1e-14 * vol * 10
                # phase of every sub-cell from the parent's linear level set
                for ph, C in zip((0, 1), cells):
                    for cell in C:
                        fi = _bary(cell, X, dim) @ f
                        assert np.all(fi >= -1e-13) if ph == 1 else np.all(fi <= 1e-13), (s, order, ph, fi)
                # new vertices on parent edges, at the zero of the linear level set
                verts = {_key(p): p for p in allc.reshape(-1, 3)}
                new = [p for k, p in verts.items() if not any(np.allclose(p, x) for x in X)]
                assert len(new) == n_edges
                for p in new:
                    b = _bary(p[None, :], X, dim)[0]
                    on = np.flatnonzero(np.abs(b) > 1e-13)
                    assert len(on) == 2 and tuple(sorted(on)) in {tuple(sorted(e)) for e in
                                                                   (TET_EDGES if dim == 3 else TRI_EDGES)}
                    assert abs(b @ f) < 1e-13
                # child faces: on the parent surface once, inside exactly twice
                count = {}
                bmeasure, nb = 0.0, 0
                for cell in allc:
                    for fc in _faces(cell, dim):
                        P = cell[list(fc)]
                        k = tuple(sorted(_key(p) for p in P))
                        count[k] = count.get(k, 0) + 1
                        b = _bary(P, X, dim)
                        if np.any(np.all(np.abs(b) < 1e-13, axis=0)):
                            nb += 1
                            bmeasure += _face_measure(P, dim)
                assert nb == expect_bfaces[n_edges], (s, order, nb)
                assert abs(bmeasure - parent_surface) < 1e-13 * parent_surface
                inner = [c for k, c in count.items()
                         if not np.any(np.all(np.abs(_bary(np.array(k), X, dim)) < 1e-13, axis=0))]
                assert all(c == 2 for c in inner), (s, order)
                n_checked += 1
                seen.add((tuple(s), order))
    # every sign pattern with a cut, under all id orderings
    n_patterns = sum(1 for _ in _cases(dim))
    assert len(seen) == n_patterns * (24 if dim == 3 else 6)
    assert n_checked == 3 * len(seen)


def test_node_hierarchy_uncut_cell_is_kept_whole():
    for f in ([1.0, 0.5, 0.3, 0.2], [-1.0, -0.5, -0.3, -0.2], [0.0, 0.5, 0.3, 0.2]):
        f = np.array(f)
        ph = 1 if f.max() > 0 else 0
        C = O.nh_cell(3, np.array([1, 2, 3, 4], np.uint32), f, TET, ph)
        assert C.shape[0] == 1 and np.array_equal(C[0], TET)
        assert O.nh_cell(3, np.array([1, 2, 3, 4], np.uint32), f, TET, 1 - ph).shape[0] == 0
