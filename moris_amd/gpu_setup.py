"""Python view of include/moris_gpu_setup.h: the set-up of a lattice set on the GPU (level set,
XTK cut of every intersected cell, element list, adof numbering, T rows, topology, CSR pattern),
the products of Problem + Set's host set-up computed by kernels.  No CPU fallback.
"""
import ctypes

import numpy as np

from . import IWG_STRUC_LINEAR_BULK, SetDesc, _check, load_library

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)

_NAMES = ("mgpu_setup_build", "mgpu_setup_destroy", "mgpu_setup_counts", "mgpu_setup_time", "mgpu_setup_copy",
          "mgpu_setup_desc", "mgpu_set_create_from_setup")


def _lib():
    lib = load_library()
    if not getattr(lib, "_setup_bound", False):
        from . import ProblemSpec
        lib.mgpu_setup_build.argtypes = [ctypes.POINTER(ProblemSpec), _vp, ctypes.POINTER(_vp)]
        lib.mgpu_setup_destroy.argtypes = [_vp]
        lib.mgpu_setup_counts.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
        lib.mgpu_setup_time.argtypes = [_vp, _f64p]
        lib.mgpu_setup_copy.argtypes = [_vp, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _i32p, _i32p, _f64p, _i32p,
                                        _i32p, _i32p]
        lib.mgpu_setup_desc.argtypes = [_vp, ctypes.POINTER(SetDesc)]
        lib.mgpu_set_create_from_setup.argtypes = [_vp, ctypes.POINTER(_vp)]
        for n in _NAMES:
            getattr(lib, n).restype = ctypes.c_int32
        lib._setup_bound = True
    return lib


class GpuSetup:
    """Set-up of one workload on the GPU (HEX8 / QUAD4 cells, bspline_order 1 or 2, no refinement).

    ``desc`` is a SetDesc over host copies owned by the handle (feeds Set); ``ms`` the device time
    of the build; ``arrays()`` / ``pattern()`` / ``topology()`` numpy copies of the device arrays.
    """

    def __init__(self, spec, keepalive=None, stream=None):
        lib = _lib()
        self._keep = keepalive
        self.spec = spec
        self._h = _vp()
        _check(lib.mgpu_setup_build(ctypes.byref(spec), stream, ctypes.byref(self._h)), "mgpu_setup_build")
        c = (ctypes.c_int64 * 8)()
        _check(lib.mgpu_setup_counts(self._h, c), "mgpu_setup_counts")
        (self.n_elements, self.n_cut, self.n_ig, self.n_adofs_per_type, self.n_rows, self.nnz, self.n_t,
         self.n_bases) = [int(v) for v in c]
        ms = ctypes.c_double()
        _check(lib.mgpu_setup_time(self._h, ctypes.byref(ms)), "mgpu_setup_time")
        self.ms = ms.value
        self._desc = None

    @property
    def desc(self):
        """SetDesc over host copies of the device arrays (made on first use)."""
        if self._desc is None:
            d = SetDesc()
            _check(_lib().mgpu_setup_desc(self._h, ctypes.byref(d)), "mgpu_setup_desc")
            self._desc = d
        return self._desc

    def make_set(self):
        """A Set made from the device arrays with no host round trip (bspline_order 1); it keeps
        this set-up alive."""
        from . import Set
        h = _vp()
        _check(_lib().mgpu_set_create_from_setup(self._h, ctypes.byref(h)), "mgpu_set_create_from_setup")
        return Set(None, keepalive=self, handle=h, n_elements=self.n_elements, n_cut=self.n_cut)

    @property
    def dim(self):
        return self.spec.dim

    @property
    def n_fields(self):
        return self.dim if self.spec.iwg == IWG_STRUC_LINEAR_BULK else 1

    def _copy(self):
        ne, dim, nb, nf = self.n_elements, self.dim, self.n_bases, self.n_fields
        order = 2 if self.spec.bspline_order == 2 else 1
        # adofs per type and element: (p + 1)^dim splines or nodes, 27 on HEX27 cells (p = 2 splines
        # or the 27 nodes)
        na = 27 if nb == 27 else (order + 1) ** dim
        out = {
            "elem_cell": np.zeros(ne, np.int32), "ig_ptr": np.zeros(ne + 1, np.int32),
            "ig_xi": np.zeros(self.n_ig * (dim + 1) * dim), "material": np.zeros(ne * (2 if nf > 1 else 1)),
            "ip_box": np.zeros(ne * 2 * dim), "vertex_phi": np.zeros(ne * nb), "t_ptr": np.zeros(ne * nb + 1, np.int32),
            "t_adof": np.zeros(self.n_t, np.int32), "t_weight": np.zeros(self.n_t), "topo": np.zeros(ne * nf * na, np.int32),
            "row_ptr": np.zeros(self.n_rows + 1, np.int32), "col_idx": np.zeros(self.nnz, np.int32),
        }
        keys = ("elem_cell", "ig_ptr", "ig_xi", "material", "ip_box", "vertex_phi", "t_ptr", "t_adof", "t_weight",
                "topo", "row_ptr", "col_idx")
        ptrs = [out[k].ctypes.data_as(_i32p if out[k].dtype == np.int32 else _f64p) for k in keys]
        _check(_lib().mgpu_setup_copy(self._h, *ptrs), "mgpu_setup_copy")
        out["topo"] = out["topo"].reshape(ne, nf * na)
        return out

    def arrays(self):
        """The descriptor arrays (the keys of Problem.arrays())."""
        a = self._copy()
        return {k: a[k] for k in ("t_ptr", "t_weight", "elem_cell", "ig_ptr", "ig_xi", "material", "t_adof", "ip_box",
                                  "vertex_phi")}

    def pattern(self):
        a = self._copy()
        return a["row_ptr"], a["col_idx"]

    def topology(self):
        return self._copy()["topo"]

    def close(self):
        if self._h:
            _lib().mgpu_setup_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
