"""Elemental extraction operators (the reference's T-matrix output,
cl_MTK_Integration_Mesh_write_T_matrices.cpp:38-160, 684-809) written by
mgpu_write_extraction_operators and read back.

Pins:
  * layout: one record per integration cell (a cut element's sub-cells, an uncut element's box),
    ids ascending and equal to the element's adofs;
  * the weights equal N(xi_v) T computed here from the oracle's shape functions and the descriptor's
    T rows (an independent evaluation of the same product);
  * partition of unity: every vertex row sums to 1 (p = 1, p = 2, HEX27, THB and enriched rows);
  * linear reproduction: with the adofs' coefficient points (lattice nodes for p = 1, Greville
    abscissae s - 1/2 for p = 2 splines) the weights reproduce the vertex's physical position.
"""
import numpy as np
import pytest

import moris_amd as M
from helpers import two_bodies
from oracle import py as O

CASES = {
    "p1_hex8_sphere_elast": lambda: M.make_spec(3, (6, 5, 4), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                spheres=[[0.5, 0.5, 0.5, 0.3]]),
    "p1_quad4_circle": lambda: M.make_spec(2, (9, 7), spheres=[[0.5, 0.5, 0.0, 0.3]]),
    "p2_hex8_sphere": lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2),
    "p2_hex27_gyroid": lambda: M.make_spec(3, (4, 4, 4), ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi,
                                           bspline_order=2, lagrange_order=2),
    "thb_sphere": lambda: M.make_spec(3, (4, 4, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2,
                                      refine_levels=1),
    "enriched_two_spheres": lambda: M.make_spec(3, (10, 10, 10), spheres=two_bodies(3, 10), material_phase=0,
                                                enrich=1),
}


def _write_read(P, tmp_path):
    path = str(tmp_path / "ext.bin")
    n = M.write_extraction_operators(P.desc, path)
    head, cells = M.read_extraction_operators(path)
    assert head["n_cells"] == n == len(cells)
    return head, cells


@pytest.mark.parametrize("case", sorted(CASES))
def test_extraction_operators(case, tmp_path):
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    a = P.arrays()
    dim, nb = P.desc.spatial_dim, P.desc.n_bases
    head, cells = _write_read(P, tmp_path)
    assert (head["dim"], head["n_bases"], head["n_adofs_per_type"]) == (dim, nb, P.n_adofs_per_type)
    ig = np.diff(a["ig_ptr"])
    assert len(cells) == int(np.maximum(ig, 1).sum())
    nv_cut = dim + 1
    box_xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                       [-1, 1, 1]], float) if dim == 3 else np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
    c = 0
    for e in range(P.n_elements):
        q0, q1 = a["t_ptr"][e * nb], a["t_ptr"][(e + 1) * nb]
        ids = np.unique(a["t_adof"][q0:q1])
        T = np.zeros((nb, len(ids)))
        for b in range(nb):
            for q in range(a["t_ptr"][e * nb + b], a["t_ptr"][e * nb + b + 1]):
                T[b, np.searchsorted(ids, a["t_adof"][q])] += a["t_weight"][q]
        subs = range(a["ig_ptr"][e], a["ig_ptr"][e + 1]) if ig[e] else [None]
        for s in subs:
            ee, got_ids, E = cells[c]
            c += 1
            assert ee == e and np.array_equal(got_ids, ids)
            xi = box_xi if s is None else a["ig_xi"][s * nv_cut * dim:(s + 1) * nv_cut * dim].reshape(nv_cut, dim)
            N, _ = O.shape(dim, xi, nb if nb == 27 else None)
            assert np.abs(E - N @ T).max() <= 1e-14
            assert np.abs(E.sum(axis=1) - 1.0).max() <= 1e-13
    assert c == len(cells)


@pytest.mark.parametrize("order", [1, 2])
def test_extraction_reproduces_vertex_positions(order, tmp_path):
    n = (6, 5, 4)
    spec, keep = M.make_spec(3, n, spheres=[[0.45, 0.5, 0.55, 0.3]], bspline_order=order)
    P = M.Problem(spec, keep)
    a = P.arrays()
    h = np.array([1.0 / k for k in n])
    ec = a["elem_cell"].astype(np.int64)
    ijk = np.stack([ec % n[0], (ec // n[0]) % n[1], ec // (n[0] * n[1])], axis=1)
    if order == 1:
        nodes_per_axis = np.array(n) + 1
        offs = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
        used = np.unique(np.concatenate([(ijk + o) @ np.array([1, nodes_per_axis[0], nodes_per_axis[0] * nodes_per_axis[1]])
                                         for o in offs]))
        idx = np.stack([used % nodes_per_axis[0], (used // nodes_per_axis[0]) % nodes_per_axis[1],
                        used // (nodes_per_axis[0] * nodes_per_axis[1])], axis=1)
        G = idx * h
    else:
        # splines s = i .. i + 2 per axis over cell i; adofs are the used splines ascending
        sp_per_axis = np.array(n) + 2
        offs = np.array([[x, y, z] for z in range(3) for y in range(3) for x in range(3)])
        used = np.unique(np.concatenate([(ijk + o) @ np.array([1, sp_per_axis[0], sp_per_axis[0] * sp_per_axis[1]])
                                         for o in offs]))
        idx = np.stack([used % sp_per_axis[0], (used // sp_per_axis[0]) % sp_per_axis[1],
                        used // (sp_per_axis[0] * sp_per_axis[1])], axis=1)
        G = (idx - 0.5) * h   # Greville abscissae of the uniform quadratic splines
    assert len(G) == P.n_adofs_per_type
    _, cells = _write_read(P, tmp_path)
    ig = np.diff(a["ig_ptr"])
    box_xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                       [-1, 1, 1]], float)
    c = 0
    for e in range(P.n_elements):
        lo, ext = a["ip_box"][e * 6:e * 6 + 3], a["ip_box"][e * 6 + 3:e * 6 + 6]
        subs = range(a["ig_ptr"][e], a["ig_ptr"][e + 1]) if ig[e] else [None]
        for s in subs:
            _, ids, E = cells[c]
            c += 1
            xi = box_xi if s is None else a["ig_xi"][s * 12:(s + 1) * 12].reshape(4, 3)
            x = lo + 0.5 * (xi + 1.0) * ext
            assert np.abs(E @ G[ids] - x).max() <= 1e-13
