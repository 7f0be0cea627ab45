"""Product set-up (moris_amd host: XTK cut, adof numbering, materials) against the oracle's
independent restatement: element lists, cut-cell topology (sub-cell counts, every vertex coordinate)
and T rows must be bit-exact."""
import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, random_spheres, spec_circle2, spec_sphere3
from oracle import py as O

CASES = {
    "c1_circle_64x64_elast": lambda: spec_circle2(64),                     # lattice vertices on the circle
    "c2_sphere_24_diff": lambda: spec_sphere3(24),
    "sphere_16_elast": lambda: spec_sphere3(16, iwg=M.IWG_STRUC_LINEAR_BULK),
    "sphere_phase0": lambda: spec_sphere3(12, phase=0),
    "random_spheres_20": lambda: M.make_spec(3, (20, 20, 20), iwg=M.IWG_STRUC_LINEAR_BULK,
                                             spheres=random_spheres(16, SEED + 4)),
    "gyroid_12": lambda: M.make_spec(3, (12, 12, 12), ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi * 3.0),
    "plane_2d": lambda: M.make_spec(2, (9, 7), ls_kind=M.LS_PLANE, plane=[0.6, 0.8, 0.0, -0.55]),
    "no_levelset": lambda: M.make_spec(3, (3, 4, 5), iwg=M.IWG_STRUC_LINEAR_BULK),
    "single_cell_cut": lambda: M.make_spec(3, (1, 1, 1), spheres=[[0.2, 0.3, 0.1, 0.55]]),
    "p2_sphere_10_elast": lambda: M.make_spec(3, (10, 10, 10), iwg=M.IWG_STRUC_LINEAR_BULK,
                                              spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2),
    "p2_circle_16_diff": lambda: M.make_spec(2, (16, 12), spheres=[[0.5, 0.5, 0.0, 0.3]], bspline_order=2,
                                             material_phase=0, k_sigma=0.3),
    # HEX27 IP cells: p = 2 B-splines at the 27 Lagrange nodes, and the nodal (Lagrange) HEX27 mesh
    "hex27_p2_gyroid_8_elast": lambda: M.make_spec(3, (8, 8, 8), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                                                   gyroid_scale=2 * np.pi * 2.0, bspline_order=2, lagrange_order=2),
    "hex27_p2_sphere_6_diff": lambda: M.make_spec(3, (6, 5, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5,
                                                  bspline_order=2, lagrange_order=2),
    "hex27_lagrange_sphere_5_elast": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                         spheres=[[0.4, 0.5, 0.6, 0.35]], lagrange_order=2),
    # one THB level on the intersected cells (truncated hierarchical quadratic B-splines)
    "thb_sphere_6_hex8_elast": lambda: M.make_spec(3, (6, 6, 6), iwg=M.IWG_STRUC_LINEAR_BULK,
                                                   spheres=[[0.5, 0.5, 0.5, 0.3]], bspline_order=2, refine_levels=1),
    "thb_plane_5_hex27_elast": lambda: M.make_spec(3, (5, 5, 5), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_PLANE,
                                                   plane=[0.3, 0.5, 0.8, -0.6], bspline_order=2, lagrange_order=2,
                                                   refine_levels=1),
    "thb_gyroid_6_diff": lambda: M.make_spec(3, (6, 6, 6), ls_kind=M.LS_GYROID, gyroid_scale=2 * np.pi, k_sigma=0.4,
                                             bspline_order=2, refine_levels=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_setup_bit_exact(case):
    spec, keep = CASES[case]()
    P = M.Problem(spec, keep)
    Q = O.OracleProblem(spec)
    a, b = P.arrays(), Q.setup()
    assert (P.n_elements, P.n_cut, P.n_ig, P.n_adofs_per_type) == (Q.n_elements, Q.n_cut, Q.n_ig, Q.n_adofs_per_type)
    for k in ("elem_cell", "ig_ptr", "ig_xi", "material", "t_ptr", "t_adof", "t_weight"):
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_bad_specs_are_rejected_with_status():
    spec, keep = M.make_spec(4 if False else 3, (0, 2, 2))
    with pytest.raises(M.MgpuError, match="bad grid"):
        M.Problem(spec, keep)
    spec, keep = M.make_spec(3, (2, 2, 2), iwg=7)
    with pytest.raises(M.MgpuError, match="iwg"):
        M.Problem(spec, keep)


def test_void_everywhere_gives_empty_set():
    # material phase 0 inside a tiny sphere between lattice nodes: no material cell at all
    spec, keep = M.make_spec(3, (4, 4, 4), spheres=[[0.3, 0.3, 0.3, 1e-3]], material_phase=0)
    P = M.Problem(spec, keep)
    assert P.n_elements == 0 and P.n_adofs_per_type == 0
