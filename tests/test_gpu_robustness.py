"""GPU tests of the error and statistics paths.

* A host-built descriptor with one inverted IG sub-tet (two vertices swapped, detJ < -1e-6): its
  integration points are skipped, as Element_Bulk::compute_jacobian_and_residual does
  (projects/FEM/INT/src/ELEM/cl_FEM_Element_Bulk.cpp:291), by the lattice stencil path, the generic
  path (p = 2 rows), the HEX27 kernel (TET_35) and the learned path's moments; K and R match the
  oracle on the same descriptor, and mgpu_set_statistics reports the skipped points and sub-cell.
"""
import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, rel
from moris_amd import learned as L
from oracle import py as O

pytestmark = pytest.mark.gpu

CASES = [
    ("hex8 diffusion stencil", dict(iwg=M.IWG_SPATIALDIFF_BULK), 11, None),
    ("hex8 elasticity stencil", dict(iwg=M.IWG_STRUC_LINEAR_BULK), 11, None),
    ("hex8 elasticity learned", dict(iwg=M.IWG_STRUC_LINEAR_BULK), 11, "learned"),
    ("hex8 elasticity p=2 generic", dict(iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2), 11, None),
    ("hex27 elasticity p=2", dict(iwg=M.IWG_STRUC_LINEAR_BULK, bspline_order=2, lagrange_order=2), 35, None),
]


def _inverted(P):
    a = P.arrays()
    X = a["ig_xi"].reshape(-1, 4, 3).copy()
    box = a["ip_box"].reshape(-1, 6)
    ig = a["ig_ptr"]
    owner = np.repeat(np.arange(len(ig) - 1), np.diff(ig))
    h = box[owner, 3:]
    J = np.stack([X[:, 0] - X[:, 2], X[:, 1] - X[:, 2], X[:, 3] - X[:, 2]], axis=1) * (0.5 * h[:, None, :])
    vol = np.linalg.det(J) / 6.0
    assert np.all(vol > -1e-6)
    s0 = int(np.argmax(vol))                     # the largest sub-tet: its inverted detJ is < -1e-6
    X[s0, [0, 1]] = X[s0, [1, 0]]
    return dict(a, ig_xi=X.ravel()), s0


@pytest.mark.parametrize("name,kw,npts,mode", CASES, ids=[c[0] for c in CASES])
def test_inverted_subcell_is_skipped_and_counted(name, kw, npts, mode):
    spec, keep = M.make_spec(3, (6, 6, 6), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5, seed=SEED + 7, **kw)
    P = M.Problem(spec, keep)
    S0 = M.Set(P.desc, keepalive=P)
    assert S0.statistics()[5:7] == [0, 0]
    arrays, s0 = _inverted(P)
    d, k = M.desc_from_arrays(P.desc, arrays)
    S = M.Set(d, keepalive=(P, k))
    if mode == "learned":
        L.attach(S, L.MLP(seed=5))
        S.options(M.PREC_MIXED, M.QUAD_LEARNED)
    u = np.random.default_rng(SEED + 8).uniform(-1.0, 1.0, S.n_rows)
    vals, R = S.compute_jacobian_and_residual(u)
    st = S.statistics()
    assert st[5] == npts and st[6] == 1, st
    rp, ci = S.pattern()
    vo, Ro, _ = O.assemble_desc(d, rp, ci, u)
    bar = 1e-6 if mode == "learned" else 1e-12
    assert rel(vals, vo) <= bar and rel(R, Ro) <= bar, (rel(vals, vo), rel(R, Ro))
    # the skipped sub-tet did contribute on the original descriptor
    v_full, _ = S0.compute_jacobian_and_residual(u)
    assert rel(vals, v_full) > 1e-8
