"""GPU tests of the assembly at the sizes of the measured configurations.

* C5, one design variant at full size: 64^3 hex8 elasticity around the 4 x 4 x 4 sphere array
  (variant 0 of the bench batch).  The product's set-up (cut sub-cells, T rows, CSR pattern,
  element topology) equals the oracle's independent set-up bit for bit; the plain FP64 path matches
  the oracle's assembly at 1e-12; mixed precision + learned quadrature with the committed network
  matches it at 1e-6, with the acceptance rate printed and bounded below.
* C4's deployment set: 64^3 hex8 elasticity around 64 seeded random spheres, the committed network:
  set-up bit-exact, FP64 at 1e-12, mixed + learned at 1e-6, acceptance rate printed.
* C3's one-GPU share: the 48^3-base gyroid with one THB level (HEX27 cells, p = 2 truncated rows,
  scatter mode): size-independent properties of the whole set (R = K u, symmetry on sampled
  entries, rigid translations in the null space, mixed within 1e-6 of FP64, bit-identical repeated
  calls), and the oracle on a sub-set of its elements through a descriptor of those elements.
"""
import os

import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, rel
from moris_amd import learned as L
from moris_amd import train_lq
from moris_amd import workloads as W
from oracle import py as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _topology_from_rows(t_ptr, t_adof, nb, nf, N):
    """Sorted unique adof ids per element per dof type (cl_MSI_Equation_Object.cpp:199-291)."""
    ne = (len(t_ptr) - 1) // nb
    out = []
    for e in range(ne):
        ids = np.unique(t_adof[t_ptr[e * nb]:t_ptr[(e + 1) * nb]])
        out.append(np.concatenate([ids + f * N for f in range(nf)]))
    return out


def _lattice_case(spec, keep, label, min_accept):
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    print(label, "elements", P.n_elements, "cut", P.n_cut, "sub-cells", P.n_ig, "rows", S.n_rows, "nnz", S.nnz,
          flush=True)
    # set-up: bit-exact against the oracle's independent XTK cut, numbering and graph
    Q = O.OracleProblem(spec)
    s = Q.setup()
    a = P.arrays()
    rp, ci = S.pattern()
    assert np.array_equal(rp, s["row_ptr"]) and np.array_equal(ci, s["col_idx"])
    for k in ("elem_cell", "ig_ptr", "t_ptr", "t_adof"):
        assert np.array_equal(a[k], s[k]), k
    assert np.array_equal(a["ig_xi"].view(np.uint64), s["ig_xi"].view(np.uint64))
    topo = S.topology()
    sample = np.linspace(0, P.n_elements - 1, 2000).astype(np.int64)
    ref_topo = _topology_from_rows(s["t_ptr"], s["t_adof"], 8, 3, P.n_adofs_per_type)
    for e in sample:
        assert np.array_equal(topo[e], ref_topo[e]), e
    # FP64 path against the oracle's assembly
    u = np.random.default_rng(SEED + 61).uniform(-1.0, 1.0, S.n_rows)
    vals, R = S.compute_jacobian_and_residual(u)
    vo, Ro, nip = O.assemble_desc(P.desc, rp, ci, u, n_threads=THREADS)
    print(label, "FP64 |dK|/|K| %.3e |dR|/|R| %.3e (oracle points %d)" % (rel(vals, vo), rel(R, Ro), nip), flush=True)
    assert rel(vals, vo) <= 1e-12 and rel(R, Ro) <= 1e-12
    # mixed precision + learned quadrature with the committed network
    net = train_lq.load()
    assert net is not None, "network checkpoint missing"
    L.attach(S, net)
    S.options(M.PREC_MIXED, M.QUAD_LEARNED)
    vl, Rl = S.compute_jacobian_and_residual(u)
    flags = L.result(S)[2]
    acc = float(flags.mean())
    print(label, "mixed+learned |dK|/|K| %.3e |dR|/|R| %.3e accepted %.4f" % (rel(vl, vo), rel(Rl, Ro), acc),
          flush=True)
    assert rel(vl, vo) <= 1e-6 and rel(Rl, Ro) <= 1e-6
    assert acc >= min_accept
    st = S.stats()
    assert st[2] == int(flags.sum()) and st[1] == P.n_cut - int(flags.sum())
    return P, S, acc


def test_c5_full_size_variant():
    spec, keep = W.elasticity_spec(64, W.sphere_array(0))
    _lattice_case(spec, keep, "C5 variant 0", min_accept=0.9)


def test_c4_deployment_random_spheres():
    spec, keep = W.elasticity_spec(64, W.random_spheres(64))
    _lattice_case(spec, keep, "C4 deployment", min_accept=0.85)


def _sub_descriptor(P, elems):
    """Descriptor of a subset of P's elements (their T rows and sub-cells; adof numbering kept)."""
    a = P.arrays()
    nb = P.desc.n_bases
    dim = P.desc.spatial_dim
    nm = 2 if P.desc.iwg == M.IWG_STRUC_LINEAR_BULK else 1
    tp, ip = a["t_ptr"], a["ig_ptr"]
    t_ptr, t_adof, t_w, ig_ptr, ig_xi = [0], [], [], [0], []
    for e in elems:
        for b in range(nb):
            r = e * nb + b
            t_adof.append(a["t_adof"][tp[r]:tp[r + 1]])
            t_w.append(a["t_weight"][tp[r]:tp[r + 1]])
            t_ptr.append(t_ptr[-1] + tp[r + 1] - tp[r])
        ig_xi.append(a["ig_xi"][ip[e] * (dim + 1) * dim:ip[e + 1] * (dim + 1) * dim])
        ig_ptr.append(ig_ptr[-1] + ip[e + 1] - ip[e])
    arrays = {"t_ptr": np.array(t_ptr, np.int32), "t_adof": np.concatenate(t_adof), "t_weight": np.concatenate(t_w),
              "ig_ptr": np.array(ig_ptr, np.int32), "ig_xi": np.concatenate(ig_xi),
              "material": a["material"].reshape(-1, nm)[elems].ravel(),
              "ip_box": a["ip_box"].reshape(-1, 2 * dim)[elems].ravel(),
              "elem_cell": np.zeros(len(elems), np.int32), "vertex_phi": np.zeros(0)}
    d, keep = M.desc_from_arrays(P.desc, arrays)
    d.n_elements = len(elems)
    d.elem_cell = None
    d.lattice = M.Lattice()
    return d, keep


@pytest.mark.timeout(900)
def test_c3_thb_48_share_properties_and_sampled_oracle():
    n = 48
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, ls_kind=M.LS_GYROID,
                             gyroid_scale=2 * np.pi * n / 4.0, bspline_order=2, lagrange_order=2, refine_levels=1)
    P = M.Problem(spec, keep)
    print("C3 48^3 THB problem: elements", P.n_elements, "cut", P.n_cut, "sub-tets", P.n_ig, flush=True)
    S = M.Set(P.desc, keepalive=P)
    entries = int(np.sum(np.diff(S.topology_offsets()).astype(np.int64) ** 2))
    print("C3 set: rows", S.n_rows, "nnz", S.nnz, "K_e entries", entries, flush=True)
    assert entries > 2 ** 28                   # scatter mode
    rp, ci = S.pattern()
    N = S.n_rows // 3
    u = np.random.default_rng(SEED + 62).uniform(-1.0, 1.0, S.n_rows)
    ub, vb, rb = M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)
    S.assemble_device(ub, vb, rb)
    M.synchronize()
    vals = vb.to_array(np.float64, S.nnz)
    R = rb.to_array(np.float64, S.n_rows)
    print("C3 assembled", flush=True)
    import scipy.sparse as sp
    K = sp.csr_matrix((vals, ci, rp), shape=(S.n_rows, S.n_rows))
    kmax = np.abs(vals).max()
    assert rel(R, K @ u) <= 1e-12
    for f in range(3):
        t = np.zeros(S.n_rows)
        t[f * N:(f + 1) * N] = 1.0
        assert np.abs(K @ t).max() <= 1e-12 * kmax * np.sqrt(S.n_rows)
    # symmetry on 200,000 sampled entries: K[c, r] located by binary search in row c
    rng = np.random.default_rng(5)
    idx = rng.integers(0, S.nnz, 200_000)
    rows = np.searchsorted(rp, idx, side="right") - 1
    cols = ci[idx]
    pos = np.array([rp[c] + np.searchsorted(ci[rp[c]:rp[c + 1]], r) for r, c in zip(rows, cols)])
    assert np.array_equal(ci[pos], rows)
    assert np.abs(vals[idx] - vals[pos]).max() <= 1e-13 * kmax
    print("C3 properties done", flush=True)
    # repeated call bit-identical
    S.assemble_device(ub, vb, rb)
    M.synchronize()
    assert np.array_equal(vb.to_array(np.float64, S.nnz).view(np.uint64), vals.view(np.uint64))
    # mixed precision within 1e-6
    S.options(M.PREC_MIXED, M.QUAD_SUBDIVISION)
    S.assemble_device(ub, vb, rb)
    M.synchronize()
    assert rel(vb.to_array(np.float64, S.nnz), vals) <= 1e-6
    del K, vb
    # oracle on 64 sampled elements of both levels, cut and uncut, through a descriptor of those elements
    ig = np.diff(P.arrays()["ig_ptr"])
    cut = np.flatnonzero(ig > 0)
    uncut = np.flatnonzero(ig == 0)
    elems = np.concatenate([cut[:: max(1, len(cut) // 48)][:48], uncut[:: max(1, len(uncut) // 16)][:16]])
    d, k = _sub_descriptor(P, elems)
    Ss = M.Set(d, keepalive=(P, k))
    rps, cis = Ss.pattern()
    us = np.random.default_rng(SEED + 63).uniform(-1.0, 1.0, Ss.n_rows)
    vs, Rs = Ss.compute_jacobian_and_residual(us)
    vo, Ro, nip = O.assemble_desc(d, rps, cis, us, n_threads=THREADS)
    print("C3 sampled elements: |dK|/|K| %.3e |dR|/|R| %.3e (oracle points %d)" % (rel(vs, vo), rel(Rs, Ro), nip),
          flush=True)
    assert rel(vs, vo) <= 1e-12 and rel(Rs, Ro) <= 1e-12
