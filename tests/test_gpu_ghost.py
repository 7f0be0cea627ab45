"""GPU tests of the ghost stabilisation (double-sided side sets of IWG_*_Virtual_Work_Ghost on the
faces between cut elements and their neighbours, include/moris_gpu_sideset.h).

* the bulk set built with spec.ghost couples every facet's two elements: its pattern equals the
  oracle's graph, which includes those couplings, bit for bit;
* bulk + ghost K and R against the oracle's bulk and ghost assembly (dense B, D, n_flat, T per
  facet) at 1e-12: HEX8 and QUAD4, diffusion with per-element conductivity and elasticity, p = 1
  and p = 2 B-spline rows; a second call is bit-identical; the ghost part alone satisfies
  K_ghost u = R_ghost;
* a ghost side set on a bulk set without the couplings is refused.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import moris_amd as M
from helpers import SEED, rel
from moris_amd import sideset as SS
from oracle import py as O

pytestmark = pytest.mark.gpu

CASES = {
    "sphere3_elast": (lambda: M.make_spec(3, (7, 6, 5), spheres=[[0.45, 0.5, 0.55, 0.32]],
                                          iwg=M.IWG_STRUC_LINEAR_BULK, ghost=1), 0.01),
    "sphere3_diff": (lambda: M.make_spec(3, (8, 8, 8), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.4, ghost=1), 0.1),
    "circle2_elast": (lambda: M.make_spec(2, (12, 10), spheres=[[0.5, 0.5, 0.0, 0.3]], iwg=M.IWG_STRUC_LINEAR_BULK,
                                          ghost=1), 0.05),
    "circle2_diff_phase0": (lambda: M.make_spec(2, (14, 14), spheres=[[0.5, 0.5, 0.0, 0.33]], material_phase=0,
                                                k_sigma=0.5, ghost=1), 1.0),
    "p2_sphere3_elast": (lambda: M.make_spec(3, (6, 6, 6), spheres=[[0.5, 0.5, 0.5, 0.3]],
                                             iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2, ghost=1), 0.01),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_ghost_matches_oracle(case):
    print("case", case, flush=True)
    mk, alpha = CASES[case]
    spec, keep = mk()
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    Q = O.OracleProblem(spec)
    s = Q.setup()
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    L, F, xi, nrm = SS.ghost_facets(P)
    iwg = SS.IWG_STRUC_LINEAR_GHOST if P.desc.iwg == M.IWG_STRUC_LINEAR_BULK else SS.IWG_SPATIALDIFF_GHOST
    geom = SS.GEOM_QUAD4 if P.dim == 3 else SS.GEOM_LINE2
    G = SS.SideSet(S, iwg, geom, L, xi, nrm, alpha=alpha, follower=F)
    u = np.random.default_rng(SEED + 61).uniform(-1.0, 1.0, S.n_rows)
    vals, R = SS.assemble(S, [G], u)
    vo, Ro, _ = O.assemble_desc(P.desc, s["row_ptr"], s["col_idx"], u)
    O.ghost_assemble(P.desc, s["row_ptr"], s["col_idx"], L, F, xi, nrm, alpha, u, vo, Ro)
    print(case, "ghost facets", len(L), "|dK|/|K| %.2e |dR|/|R| %.2e" % (rel(vals, vo), rel(R, Ro)))
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12
    v2, R2 = SS.assemble(S, [G], u)
    assert np.array_equal(v2.view(np.uint64), vals.view(np.uint64)) and np.array_equal(R2.view(np.uint64), R.view(np.uint64))
    # the ghost part alone: K_ghost u = R_ghost, and it is not zero
    vb, Rb = SS.assemble(S, [], u)
    Kg = sp.csr_matrix((vals - vb, ci, rp), shape=(S.n_rows, S.n_rows))
    assert np.abs(vals - vb).max() > 0.0
    assert np.abs(Kg @ u - (R - Rb)).max() <= 1e-12 * np.abs(R).max()


def test_ghost_needs_a_coupled_bulk_set():
    spec, keep = M.make_spec(3, (6, 6, 6), spheres=[[0.5, 0.5, 0.5, 0.3]], iwg=M.IWG_STRUC_LINEAR_BULK, ghost=1)
    P = M.Problem(spec, keep)
    L, F, xi, nrm = SS.ghost_facets(P)
    spec.ghost = 0
    P0 = M.Problem(spec, keep)
    S0 = M.Set(P0.desc, keepalive=P0)   # lattice stencil path, no couplings
    with pytest.raises(M.MgpuError, match="couple"):
        SS.SideSet(S0, SS.IWG_STRUC_LINEAR_GHOST, SS.GEOM_QUAD4, L, xi, nrm, alpha=0.01, follower=F)
