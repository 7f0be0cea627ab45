"""GPU tests of launch shapes, the autotuner and graph-captured batches.

* every stencil variant and cut-element order gives bit-identical K and R (each row and each cut
  element is computed by one thread / wavefront in a fixed order);
* the autotuner times every shape, keeps one of them and leaves the results unchanged;
* a batch of design variants (sphere-array radii) captured into one HIP graph: each replay, with
  the outputs cleared by the host and synchronisations between replays, reproduces the eager
  per-set results bit for bit, for 1 and 2 graph branches and the autotuned branch count, on the
  FP64 path and on mixed precision + learned quadrature; the FP64 variants also match the oracle.
"""
import numpy as np
import pytest

import moris_amd as M
from helpers import SEED, rel
from moris_amd import learned as L
from moris_amd import workloads as W
from oracle import py as O

pytestmark = pytest.mark.gpu


def _set(spec_keep, quad=None, net=None, prec=M.PREC_FP64):
    spec, keep = spec_keep
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    if net is not None:
        L.attach(S, net)
    S.options(prec, quad if quad is not None else M.QUAD_SUBDIVISION)
    return spec, S


def _buffers(S, seed):
    u = np.random.default_rng(seed).uniform(-1.0, 1.0, S.n_rows)
    return u, M.DeviceBuffer.from_array(u), M.DeviceBuffer(8 * S.nnz), M.DeviceBuffer(8 * S.n_rows)


def _run(S, ub, vb, rb):
    S.assemble_device(ub, vb, rb)
    M.synchronize()
    return vb.to_array(np.float64, S.nnz), rb.to_array(np.float64, S.n_rows)


@pytest.mark.parametrize("iwg,n", [(M.IWG_SPATIALDIFF_BULK, 14), (M.IWG_STRUC_LINEAR_BULK, 10)])
def test_launch_shapes_bit_identical(iwg, n):
    spec, S = _set(M.make_spec(3, (n, n, n), iwg=iwg, spheres=W.sphere_array(0, per_axis=2, radius=0.2)))
    assert S.n_cut > 0
    u, ub, vb, rb = _buffers(S, 1)
    v0, r0 = _run(S, ub, vb, rb)
    rows = S.tuning_variants()
    assert len(rows) >= 2
    for var in range(len(rows)):
        for order in (0, 1):
            S.tuning(var, order)
            v, r = _run(S, ub, vb, rb)
            assert np.array_equal(v, v0) and np.array_equal(r, r0), (var, order)
    ms = S.autotune(ub, vb, rb, iters=2, ignore_cache=True)
    assert len(ms) == 2 * len(rows) and all(t > 0 for t in ms.values())
    var, order = S.tuning()
    assert ms[(var, order)] == min(ms.values())
    v, r = _run(S, ub, vb, rb)
    assert np.array_equal(v, v0) and np.array_equal(r, r0)
    vo, Ro = O.OracleProblem(spec).assemble(u)
    assert rel(v, vo) <= 1e-12 and rel(r, Ro) <= 1e-12


def test_tuning_rejects_bad_shapes():
    _, S = _set(M.make_spec(3, (6, 6, 6), spheres=[[0.5, 0.5, 0.5, 0.3]]))
    with pytest.raises(M.MgpuError):
        S.tuning(len(S.tuning_variants()), 0)
    with pytest.raises(M.MgpuError):
        S.tuning(0, 2)


@pytest.mark.parametrize("learned", [False, True])
def test_batch_graph_replays_match_eager(learned):
    n, n_var = 10, 4
    net = None
    if learned:
        x, y, _ = L.dataset(20000, SEED + 40)
        net = L.MLP(seed=9)
        net.train(x, y, epochs=3, batch=512, lr=3e-3)
    sets, specs, bufs, eager = [], [], [], []
    for k in range(n_var):
        sk = W.elasticity_spec(n, W.sphere_array(k, per_axis=2, radius=0.2))
        spec, S = _set(sk, quad=M.QUAD_LEARNED if learned else None, net=net,
                       prec=M.PREC_MIXED if learned else M.PREC_FP64)
        u, ub, vb, rb = _buffers(S, 10 + k)
        eager.append(_run(S, ub, vb, rb))
        sets.append(S)
        specs.append((spec, u))
        bufs.append((ub, vb, rb))
    for nb in (1, 2, 0):
        B = M.Batch(sets, [b[0] for b in bufs], [b[1] for b in bufs], [b[2] for b in bufs], n_branches=nb)
        for replay in range(3):
            zeros_v = [np.zeros(S.nnz) for S in sets]
            for S, (ub, vb, rb), z in zip(sets, bufs, zeros_v):
                M.load_library().mgpu_memcpy_htod(vb.ptr, z.ctypes.data, z.nbytes)
                M.load_library().mgpu_memcpy_htod(rb.ptr, np.zeros(S.n_rows).ctypes.data, 8 * S.n_rows)
            M.synchronize()
            B.launch()
            B.synchronize()
            for S, (ub, vb, rb), (ve, re) in zip(sets, bufs, eager):
                v = vb.to_array(np.float64, S.nnz)
                r = rb.to_array(np.float64, S.n_rows)
                assert np.array_equal(v, ve) and np.array_equal(r, re), (nb, replay)
        per, used = B.time(2)
        assert per > 0 and 1 <= used <= n_var
        B.close()
    bar = 1e-6 if learned else 1e-12
    for (spec, u), (ve, re) in zip(specs, eager):
        vo, Ro = O.OracleProblem(spec).assemble(u)
        assert rel(ve, vo) <= bar and rel(re, Ro) <= bar
