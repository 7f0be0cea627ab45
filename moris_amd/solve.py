"""Python view of include/moris_gpu_solve.h: the assembled system's solve on the GPU.

The reference solves K u = R after assembly with AztecOO (projects/SOL/DLA/src/
cl_DLA_Linear_Solver_Aztec.cpp:88-137); here conjugate gradients with a Jacobi preconditioner run
on the CSR values the assembly left in device memory (restarted GMRES, the reference's default, for
the non-symmetric systems of unsymmetric Nitsche side sets, with a Jacobi or an ILU(k)
preconditioner), with Aztec's default convergence test ||r_k|| <= rtol ||r_0||.  No CPU fallback:
every call goes through libmoris_gpu.so.
"""
import ctypes

import numpy as np

from . import DeviceBuffer, _check, load_library

_vp = ctypes.c_void_p


class CgInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_residual", ctypes.c_double),
                ("true_rel_residual", ctypes.c_double)]


def _lib():
    lib = load_library()
    if not getattr(lib, "_solve_bound", False):
        lib.mgpu_csr_apply_dirichlet.argtypes = [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        lib.mgpu_csr_apply_dirichlet.restype = ctypes.c_int32
        lib.mgpu_cg_solve.argtypes = [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int32, _vp,
                                      ctypes.POINTER(CgInfo)]
        lib.mgpu_cg_solve.restype = ctypes.c_int32
        lib.mgpu_gmres_solve.argtypes = [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int32,
                                         ctypes.c_int32, _vp, ctypes.POINTER(CgInfo)]
        lib.mgpu_gmres_solve.restype = ctypes.c_int32
        lib.mgpu_ilu_create.argtypes = [ctypes.c_int32, _vp, _vp, ctypes.c_int32, _vp, ctypes.POINTER(_vp)]
        lib.mgpu_ilu_factor.argtypes = [_vp, _vp, _vp]
        lib.mgpu_ilu_apply.argtypes = [_vp, _vp, _vp, _vp]
        lib.mgpu_ilu_factors.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(ctypes.c_int32)]
        lib.mgpu_ilu_destroy.argtypes = [_vp]
        lib.mgpu_gmres_solve_ilu.argtypes = [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int32,
                                             ctypes.c_int32, _vp, _vp, ctypes.POINTER(CgInfo)]
        for f in ("mgpu_ilu_create", "mgpu_ilu_factor", "mgpu_ilu_apply", "mgpu_ilu_factors", "mgpu_ilu_destroy",
                  "mgpu_gmres_solve_ilu"):
            getattr(lib, f).restype = ctypes.c_int32
        lib._solve_bound = True
    return lib


class DeviceCSR:
    """A CSR matrix in device memory: the pattern (int32) and a values buffer."""

    def __init__(self, row_ptr, col_idx, vals=None, vals_buf=None):
        self.n_rows = len(row_ptr) - 1
        self.nnz = len(col_idx)
        self.row_ptr = DeviceBuffer.from_array(np.ascontiguousarray(row_ptr, np.int32))
        self.col_idx = DeviceBuffer.from_array(np.ascontiguousarray(col_idx, np.int32))
        if vals_buf is not None:
            self.vals = vals_buf
        else:
            self.vals = DeviceBuffer.from_array(np.ascontiguousarray(vals, np.float64))

    def values(self):
        return self.vals.to_array(np.float64, self.nnz)


def apply_dirichlet(A, rhs_buf, fixed, value=None):
    """Strong Dirichlet rows/columns on A (in place) and rhs (device buffer); fixed: bool per row."""
    f = DeviceBuffer.from_array(np.ascontiguousarray(fixed, np.uint8))
    v = DeviceBuffer.from_array(np.ascontiguousarray(value, np.float64)) if value is not None else None
    _check(_lib().mgpu_csr_apply_dirichlet(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, A.vals.ptr, rhs_buf.ptr, f.ptr,
                                           v.ptr if v is not None else None, None), "mgpu_csr_apply_dirichlet")
    load_library().mgpu_synchronize()


def cg(A, b, x0=None, rtol=1e-10, max_iter=20000):
    """Solve A x = b (b: host array or DeviceBuffer) by Jacobi-preconditioned CG on the GPU;
    returns (x as a host array, CgInfo)."""
    bb = b if isinstance(b, DeviceBuffer) else DeviceBuffer.from_array(np.ascontiguousarray(b, np.float64))
    x = DeviceBuffer.from_array(np.zeros(A.n_rows) if x0 is None else np.ascontiguousarray(x0, np.float64))
    info = CgInfo()
    _check(_lib().mgpu_cg_solve(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, A.vals.ptr, bb.ptr, x.ptr, rtol, max_iter,
                                None, ctypes.byref(info)), "mgpu_cg_solve")
    return x.to_array(np.float64, A.n_rows), info


def gmres(A, b, x0=None, rtol=1e-10, restart=30, max_iter=20000, ilu=None):
    """Solve A x = b (any non-singular A) by restarted GMRES with a right preconditioner on the GPU:
    Jacobi, or the factors of an ILU object (factored from A); returns (x as a host array, CgInfo
    with iterations = Arnoldi steps)."""
    bb = b if isinstance(b, DeviceBuffer) else DeviceBuffer.from_array(np.ascontiguousarray(b, np.float64))
    x = DeviceBuffer.from_array(np.zeros(A.n_rows) if x0 is None else np.ascontiguousarray(x0, np.float64))
    info = CgInfo()
    if ilu is None:
        _check(_lib().mgpu_gmres_solve(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, A.vals.ptr, bb.ptr, x.ptr, rtol,
                                       restart, max_iter, None, ctypes.byref(info)), "mgpu_gmres_solve")
    else:
        _check(_lib().mgpu_gmres_solve_ilu(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, A.vals.ptr, bb.ptr, x.ptr, rtol,
                                           restart, max_iter, ilu._h, None, ctypes.byref(info)),
               "mgpu_gmres_solve_ilu")
    return x.to_array(np.float64, A.n_rows), info


class ILU:
    """ILU(k) factors of a DeviceCSR on the GPU (mgpu_ilu_*): the pattern and level schedules at
    construction, the numeric factors by factor() (called once here)."""

    def __init__(self, A, fill_level=0, factor=True):
        self._h = _vp()
        _check(_lib().mgpu_ilu_create(A.n_rows, A.row_ptr.ptr, A.col_idx.ptr, fill_level, None, ctypes.byref(self._h)),
               "mgpu_ilu_create")
        self.fill_level = fill_level
        if factor:
            self.factor(A)

    def factor(self, A):
        _check(_lib().mgpu_ilu_factor(self._h, A.vals.ptr, None), "mgpu_ilu_factor")

    def apply(self, v):
        """(L U)^-1 v for a host vector; returns a host array."""
        n = self.factors(values=False)[0]
        vb = DeviceBuffer.from_array(np.ascontiguousarray(v, np.float64))
        zb = DeviceBuffer(8 * max(n, 1))
        _check(_lib().mgpu_ilu_apply(self._h, vb.ptr, zb.ptr, None), "mgpu_ilu_apply")
        load_library().mgpu_synchronize()
        return zb.to_array(np.float64, n)

    def factors(self, values=True):
        """(n, row_ptr, col_idx, lu, (forward levels, backward levels)) as host arrays."""
        n, nnz = ctypes.c_int32(), ctypes.c_int64()
        rp, ci, lu = _vp(), _vp(), _vp()
        lev = (ctypes.c_int32 * 2)()
        _check(_lib().mgpu_ilu_factors(self._h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci),
                                       ctypes.byref(lu), lev), "mgpu_ilu_factors")
        if not values:
            return n.value, None, None, None, (lev[0], lev[1])
        lib = load_library()
        out_rp = np.zeros(n.value + 1, np.int32)
        out_ci = np.zeros(nnz.value, np.int32)
        out_lu = np.zeros(nnz.value, np.float64)
        for arr, src in ((out_rp, rp), (out_ci, ci), (out_lu, lu)):
            if arr.nbytes and src.value:
                _check(lib.mgpu_memcpy_dtoh(arr.ctypes.data, src.value, arr.nbytes), "mgpu_memcpy_dtoh")
        return n.value, out_rp, out_ci, out_lu, (lev[0], lev[1])

    def close(self):
        if self._h:
            _lib().mgpu_ilu_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

