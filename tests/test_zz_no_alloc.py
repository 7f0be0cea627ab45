"""mgpu_cg_solve_noalloc (work space in a fixed device array, no allocation) against mgpu_cg_solve.

Both are called through the C ABI on the same device buffers: two calls in a row, the second on a
smaller system, each bit-identical to the allocating entry point (solution, iteration count,
residuals).  A system beyond the fixed work space is refused before any device work.
"""
import ctypes

import numpy as np
import pytest

import moris_amd as M
from helpers import SEED
from moris_amd import solve as SV

pytestmark = pytest.mark.gpu

MAX_ROWS = 2500000


def _system(n):
    spec, keep = M.make_spec(3, (n, n, n), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.5, 0.3]],
                             seed=SEED + 70)
    P = M.Problem(spec, keep)
    S = M.Set(P.desc, keepalive=P)
    K, _ = S.compute_jacobian_and_residual(np.zeros(S.n_rows))
    rp, ci = S.pattern()
    # clamp the first adofs of each field so that K is definite
    fixed = np.zeros(S.n_rows, bool)
    N = S.n_rows // 3
    for f in range(3):
        fixed[f * N:f * N + max(4, N // 20)] = True
    A = SV.DeviceCSR(rp, ci, K)
    b = M.DeviceBuffer.from_array(np.random.default_rng(n).uniform(-1.0, 1.0, S.n_rows))
    SV.apply_dirichlet(A, b, fixed)
    return A, b


def _call(fn, A, b, rtol=1e-11, max_iter=5000):
    lib = SV._lib()
    getattr(lib, fn).argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
                                 ctypes.POINTER(SV.CgInfo)]
    getattr(lib, fn).restype = ctypes.c_int32
    x = M.DeviceBuffer.from_array(np.zeros(A.n_rows))
    info = SV.CgInfo()
    rc = getattr(lib, fn)(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, A.vals.ptr, b.ptr, x.ptr, rtol, max_iter, None,
                          ctypes.byref(info))
    assert rc == 0, M.load_library().mgpu_last_error_string()
    return x.to_array(np.float64, A.n_rows), info


def test_noalloc_cg_equals_allocating_cg_on_two_calls():
    for n in (14, 9):                       # the second system is smaller than the first
        A, b = _system(n)
        x_ref, i_ref = _call("mgpu_cg_solve", A, b)
        x, info = _call("mgpu_cg_solve_noalloc", A, b)
        print("rows", A.n_rows, "iterations", info.iterations, "rel residual %.2e" % info.rel_residual, flush=True)
        assert i_ref.converged == 1 and info.converged == 1
        assert np.array_equal(x.view(np.uint64), x_ref.view(np.uint64))
        assert info.iterations == i_ref.iterations
        assert info.rel_residual == i_ref.rel_residual and info.true_rel_residual == i_ref.true_rel_residual


def test_noalloc_cg_refuses_a_system_beyond_its_work_space():
    lib = SV._lib()
    lib.mgpu_cg_solve_noalloc.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 5 + [ctypes.c_double, ctypes.c_int32,
                                                                                      ctypes.c_void_p, ctypes.c_void_p]
    lib.mgpu_cg_solve_noalloc.restype = ctypes.c_int32
    dummy = M.DeviceBuffer(64)
    rc = lib.mgpu_cg_solve_noalloc(MAX_ROWS + 1, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr, 1e-8, 10,
                                   None, None)
    assert rc == 1
    assert b"exceed" in M.load_library().mgpu_last_error_string()
