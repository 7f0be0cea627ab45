"""moris_amd -- MI355X-native element K+R assembly for XTK-cut HMR meshes.

Python view of the C ABI in ``include/moris_gpu.h`` (libmoris_gpu.so, built in-tree by
``moris_amd.build``).  The classes mirror the reference's operator surface for the bulk-assembly
path: a :class:`Problem` is the HMR/GEN/XTK/MSI set-up that produces element-major set data, a
:class:`Set` is ``fem::Set`` + the SOL graph, and :meth:`Set.compute_jacobian_and_residual`
replaces ``Solver_Interface::fill_matrix_and_RHS`` over the set's equation objects
(projects/SOL/DLA/src/cl_DLA_Solver_Interface.cpp:317-365).

No PyTorch here: the library owns device memory through its own helpers, so the process holds a
single HIP runtime.  There is no CPU fallback: every compute call goes through the HIP library and
raises if it is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmoris_gpu.so")

# constants (include/moris_gpu.h)
MGPU_OK = 0
IWG_SPATIALDIFF_BULK = 0
IWG_STRUC_LINEAR_BULK = 1
GEOM_QUAD4, GEOM_HEX8, GEOM_TRI3, GEOM_TET4, GEOM_HEX27 = 0, 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
QUAD_SUBDIVISION, QUAD_LEARNED, QUAD_MOMENT_FIT = 0, 1, 2
LS_NONE, LS_SPHERES, LS_GYROID, LS_PLANE = 0, 1, 2, 3

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_f64p = ctypes.POINTER(ctypes.c_double)


class ProblemSpec(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("iwg", ctypes.c_int32),
        ("ls_kind", ctypes.c_int32),
        ("n_spheres", ctypes.c_int32),
        ("spheres", _c_f64p),
        ("gyroid_scale", ctypes.c_double),
        ("plane", ctypes.c_double * 4),
        ("material_phase", ctypes.c_int32),
        ("E", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("k_mu", ctypes.c_double),
        ("k_sigma", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("bspline_order", ctypes.c_int32),
        ("lagrange_order", ctypes.c_int32),
        ("refine_levels", ctypes.c_int32),
        ("enrich", ctypes.c_int32),
        ("ghost", ctypes.c_int32),
    ]


class Lattice(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("lo", ctypes.c_double * 3), ("h", ctypes.c_double * 3)]


class SetDesc(ctypes.Structure):
    _fields_ = [
        ("spatial_dim", ctypes.c_int32),
        ("iwg", ctypes.c_int32),
        ("ip_geometry", ctypes.c_int32),
        ("ig_geometry", ctypes.c_int32),
        ("n_elements", ctypes.c_int32),
        ("n_bases", ctypes.c_int32),
        ("n_adofs_per_type", ctypes.c_int32),
        ("ip_box", _c_f64p),
        ("material", _c_f64p),
        ("t_ptr", _c_i32p),
        ("t_adof", _c_i32p),
        ("t_weight", _c_f64p),
        ("ig_ptr", _c_i32p),
        ("ig_xi", _c_f64p),
        ("lattice", Lattice),
        ("elem_cell", _c_i32p),
        ("vertex_phi", _c_f64p),
        ("n_couplings", ctypes.c_int32),
        ("couplings", _c_i32p),
    ]


# every symbol include/moris_gpu.h declares
EXPORTED = [
    "mgpu_problem_build", "mgpu_problem_desc", "mgpu_problem_counts", "mgpu_problem_destroy",
    "mgpu_write_extraction_operators",
    "mgpu_set_create", "mgpu_set_destroy", "mgpu_set_options", "mgpu_set_stream", "mgpu_pattern_get",
    "mgpu_element_topology", "mgpu_element_topology_csr", "mgpu_set_assemble", "mgpu_set_assemble_host", "mgpu_set_stats",
    "mgpu_set_statistics",
    "mgpu_device_count", "mgpu_malloc", "mgpu_free", "mgpu_memcpy_htod", "mgpu_memcpy_dtoh",
    "mgpu_synchronize", "mgpu_set_time_assemble", "mgpu_last_error_string",
    "mgpu_set_tuning", "mgpu_set_get_tuning", "mgpu_set_tuning_variants", "mgpu_set_autotune",
    "mgpu_batch_create", "mgpu_batch_destroy", "mgpu_batch_launch", "mgpu_batch_synchronize", "mgpu_batch_time",
]

_lib = None


class MgpuError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libmoris_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MgpuError("libmoris_gpu.so not built at %s (run python -m moris_amd.build)" % path)
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.mgpu_last_error_string.restype = ctypes.c_char_p
    lib.mgpu_problem_build.argtypes = [ctypes.POINTER(ProblemSpec), ctypes.POINTER(vp)]
    lib.mgpu_problem_desc.argtypes = [vp, ctypes.POINTER(SetDesc)]
    lib.mgpu_problem_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.mgpu_problem_destroy.argtypes = [vp]
    lib.mgpu_set_create.argtypes = [ctypes.POINTER(SetDesc), ctypes.POINTER(vp)]
    lib.mgpu_set_destroy.argtypes = [vp]
    lib.mgpu_set_options.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.mgpu_set_stream.argtypes = [vp, vp]
    lib.mgpu_pattern_get.argtypes = [vp, _c_i32p, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(_c_i32p), ctypes.POINTER(_c_i32p)]
    lib.mgpu_element_topology.argtypes = [vp, _c_i32p, ctypes.POINTER(_c_i32p)]
    lib.mgpu_element_topology_csr.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int64)),
                                              ctypes.POINTER(_c_i32p)]
    lib.mgpu_set_assemble.argtypes = [vp, vp, vp, vp]
    lib.mgpu_set_assemble_host.argtypes = [vp, _c_f64p, _c_f64p, _c_f64p]
    lib.mgpu_set_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.mgpu_set_statistics.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    lib.mgpu_device_count.argtypes = [_c_i32p]
    lib.mgpu_malloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    lib.mgpu_free.argtypes = [vp]
    lib.mgpu_memcpy_htod.argtypes = [vp, vp, ctypes.c_size_t]
    lib.mgpu_memcpy_dtoh.argtypes = [vp, vp, ctypes.c_size_t]
    lib.mgpu_synchronize.argtypes = []
    lib.mgpu_set_time_assemble.argtypes = [vp, vp, vp, vp, ctypes.c_int32, _c_f64p, _c_f64p]
    lib.mgpu_set_tuning.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.mgpu_set_get_tuning.argtypes = [vp, _c_i32p, _c_i32p]
    lib.mgpu_set_tuning_variants.argtypes = [vp, _c_i32p, _c_i32p]
    lib.mgpu_set_autotune.argtypes = [vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32, _c_f64p]
    lib.mgpu_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                      ctypes.POINTER(vp), ctypes.c_int32, vp, ctypes.POINTER(vp)]
    lib.mgpu_batch_destroy.argtypes = [vp]
    lib.mgpu_batch_launch.argtypes = [vp]
    lib.mgpu_batch_synchronize.argtypes = [vp]
    lib.mgpu_batch_time.argtypes = [vp, ctypes.c_int32, _c_f64p, _c_i32p]
    for name in EXPORTED:
        getattr(lib, name).restype = ctypes.c_int32 if name != "mgpu_last_error_string" else ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc, what):
    if rc != MGPU_OK:
        msg = _lib.mgpu_last_error_string().decode(errors="replace")
        raise MgpuError("%s failed (status %d): %s" % (what, rc, msg))


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def make_spec(dim, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), iwg=IWG_SPATIALDIFF_BULK, spheres=None,
              ls_kind=None, gyroid_scale=1.0, plane=(0.0, 0.0, 0.0, 0.0), material_phase=1, E=1.0, nu=0.3,
              k_mu=0.0, k_sigma=0.0, seed=20261202, bspline_order=1, lagrange_order=1, refine_levels=0,
              enrich=0, ghost=0):
    """Build a ProblemSpec; returns (spec, keepalive) -- keep the second value alive with the spec."""
    s = ProblemSpec()
    s.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    s.n[:] = nn[:3]
    s.lo[:] = (list(lo) + [0.0] * 3)[:3]
    s.hi[:] = (list(hi) + [1.0] * 3)[:3]
    s.iwg = iwg
    keep = None
    if spheres is not None:
        keep = np.ascontiguousarray(np.asarray(spheres, dtype=np.float64).reshape(-1, 4))
        s.n_spheres = keep.shape[0]
        s.spheres = _ptr(keep, ctypes.c_double)
        s.ls_kind = LS_SPHERES if ls_kind is None else ls_kind
    else:
        s.ls_kind = LS_NONE if ls_kind is None else ls_kind
    s.gyroid_scale = gyroid_scale
    s.plane[:] = list(plane)
    s.material_phase = material_phase
    s.E, s.nu, s.k_mu, s.k_sigma = E, nu, k_mu, k_sigma
    s.seed = seed
    s.bspline_order = bspline_order
    s.lagrange_order = lagrange_order
    s.refine_levels = refine_levels
    s.enrich = enrich
    s.ghost = ghost
    return s, keep


class Problem:
    """Host set-up of one workload (lattice, level set, XTK cut, adof numbering)."""

    def __init__(self, spec, keepalive=None):
        lib = load_library()
        self._keep = keepalive
        self.spec = spec
        self._h = ctypes.c_void_p()
        _check(lib.mgpu_problem_build(ctypes.byref(spec), ctypes.byref(self._h)), "mgpu_problem_build")
        self.desc = SetDesc()
        _check(lib.mgpu_problem_desc(self._h, ctypes.byref(self.desc)), "mgpu_problem_desc")
        c = (ctypes.c_int64 * 6)()
        _check(lib.mgpu_problem_counts(self._h, c), "mgpu_problem_counts")
        (self.n_elements, self.n_cut, self.n_ig, self.n_adofs_per_type, self.n_cells, self.n_void) = list(c)

    @property
    def dim(self):
        return self.desc.spatial_dim

    def arrays(self):
        """numpy views (copies) of the descriptor arrays."""
        d = self.desc
        ne, nb, dim = d.n_elements, d.n_bases, d.spatial_dim
        nm = 2 if d.iwg == IWG_STRUC_LINEAR_BULK else 1
        nv = dim + 1

        def view(p, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        ig_ptr = view(d.ig_ptr, ne + 1, np.int32)
        t_ptr = view(d.t_ptr, ne * nb + 1, np.int32)
        n_t = int(t_ptr[-1]) if ne else 0
        return {
            "t_ptr": t_ptr,
            "t_weight": view(d.t_weight, n_t, np.float64),
            "elem_cell": view(d.elem_cell, ne, np.int32),
            "ig_ptr": ig_ptr,
            "ig_xi": view(d.ig_xi, int(ig_ptr[-1]) * nv * dim, np.float64),
            "material": view(d.material, ne * nm, np.float64),
            "t_adof": view(d.t_adof, n_t, np.int32),
            "ip_box": view(d.ip_box, ne * 2 * dim, np.float64),
            "vertex_phi": view(d.vertex_phi, ne * nb, np.float64),
        }

    def close(self):
        if self._h:
            load_library().mgpu_problem_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_extraction_operators(desc, path):
    """mgpu_write_extraction_operators: the elemental extraction operators (T-matrices from the
    integration cells to the adofs) of a set descriptor to `path`; returns the cell count."""
    lib = load_library()
    lib.mgpu_write_extraction_operators.argtypes = [ctypes.POINTER(SetDesc), ctypes.c_char_p,
                                                    ctypes.POINTER(ctypes.c_int64)]
    n = ctypes.c_int64()
    _check(lib.mgpu_write_extraction_operators(ctypes.byref(desc), os.fsencode(path), ctypes.byref(n)),
           "mgpu_write_extraction_operators")
    return n.value


def read_extraction_operators(path):
    """Read a file of write_extraction_operators: (header dict, list of (element, adof ids [na],
    weights [n_vertices][na]))."""
    raw = open(path, "rb").read()
    if raw[:8] != b"MGPUEXT1":
        raise MgpuError("not an extraction-operator file: %s" % path)
    dim, nb, n_adofs, _ = np.frombuffer(raw, np.int32, 4, 8)
    n_cells = int(np.frombuffer(raw, np.int64, 1, 24)[0])
    off, cells = 32, []
    for _ in range(n_cells):
        e, nv, na = (int(x) for x in np.frombuffer(raw, np.int32, 3, off))
        off += 12
        ids = np.frombuffer(raw, np.int32, na, off).copy()
        off += 4 * na
        E = np.frombuffer(raw, np.float64, nv * na, off).reshape(nv, na).copy()
        off += 8 * nv * na
        cells.append((e, ids, E))
    return {"dim": int(dim), "n_bases": int(nb), "n_adofs_per_type": int(n_adofs), "n_cells": n_cells}, cells


def desc_from_arrays(desc, arrays):
    """A copy of set descriptor `desc` whose arrays are taken from `arrays` (numpy, the keys of
    Problem.arrays()); returns (SetDesc, keepalive) -- keep the second value alive with the first."""
    d = SetDesc()
    ctypes.pointer(d)[0] = desc
    keep = {}
    ctypes_of = {"t_ptr": ctypes.c_int32, "t_adof": ctypes.c_int32, "ig_ptr": ctypes.c_int32,
                 "elem_cell": ctypes.c_int32, "t_weight": ctypes.c_double, "ig_xi": ctypes.c_double,
                 "material": ctypes.c_double, "ip_box": ctypes.c_double, "vertex_phi": ctypes.c_double}
    for k, a in arrays.items():
        ct = ctypes_of[k]
        a = np.ascontiguousarray(a, dtype=np.int32 if ct is ctypes.c_int32 else np.float64)
        keep[k] = a
        setattr(d, k, _ptr(a, ct) if a.size else None)
    return d, keep


class DeviceBuffer:
    """Device allocation owned by the library's HIP runtime."""

    def __init__(self, nbytes):
        lib = load_library()
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _check(lib.mgpu_malloc(ctypes.byref(self.ptr), max(self.nbytes, 8)), "mgpu_malloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            _check(load_library().mgpu_memcpy_htod(b.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "htod")
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(load_library().mgpu_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes), "dtoh")
        return out

    def free(self):
        if self.ptr:
            load_library().mgpu_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Set:
    """fem::Set of the bulk blocks of a problem, on the GPU."""

    def __init__(self, desc, keepalive=None, handle=None, n_elements=None, n_cut=None):
        """A set from a descriptor (mgpu_set_create), or around a handle made elsewhere (the GPU
        set-up's mgpu_set_create_from_setup), given its element and cut-element counts."""
        lib = load_library()
        self._keep = keepalive
        self.desc = desc
        if handle is None:
            self._h = ctypes.c_void_p()
            _check(lib.mgpu_set_create(ctypes.byref(desc), ctypes.byref(self._h)), "mgpu_set_create")
            self.n_elements = desc.n_elements
            ne = desc.n_elements
            ig = np.ctypeslib.as_array(desc.ig_ptr, shape=(ne + 1,)) if ne else np.zeros(1, np.int32)
            self.n_cut = int(np.count_nonzero(np.diff(ig)))
        else:
            self._h = handle
            self.n_elements = n_elements
            self.n_cut = n_cut
        nr = ctypes.c_int32()
        nnz = ctypes.c_int64()
        _check(lib.mgpu_pattern_get(self._h, ctypes.byref(nr), ctypes.byref(nnz), None, None), "mgpu_pattern_get")
        self.n_rows = nr.value
        self.nnz = nnz.value
        self._rp = self._ci = None

    def _pattern_ptrs(self):
        if self._rp is None:
            rp, ci = _c_i32p(), _c_i32p()
            _check(load_library().mgpu_pattern_get(self._h, None, None, ctypes.byref(rp), ctypes.byref(ci)),
                   "mgpu_pattern_get")
            self._rp, self._ci = rp, ci
        return self._rp, self._ci

    def options(self, precision=PREC_FP64, quadrature=QUAD_SUBDIVISION):
        _check(load_library().mgpu_set_options(self._h, precision, quadrature), "mgpu_set_options")

    def pattern(self):
        prp, pci = self._pattern_ptrs()
        rp = np.ctypeslib.as_array(prp, shape=(self.n_rows + 1,)).copy()
        ci = np.ctypeslib.as_array(pci, shape=(max(self.nnz, 1),))[: self.nnz].copy()
        return rp, ci

    def topology(self):
        n = ctypes.c_int32()
        p = _c_i32p()
        _check(load_library().mgpu_element_topology(self._h, ctypes.byref(n), ctypes.byref(p)), "topology")
        ne = self.n_elements
        if ne == 0:
            return np.zeros((0, n.value), dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(ne * n.value,)).reshape(ne, n.value).copy()

    def topology_offsets(self):
        """Offsets of the element topologies (int64, n_elements + 1)."""
        off = ctypes.POINTER(ctypes.c_int64)()
        p = _c_i32p()
        _check(load_library().mgpu_element_topology_csr(self._h, ctypes.byref(off), ctypes.byref(p)), "topology_csr")
        ne = self.n_elements
        if ne == 0:
            return np.zeros(1, np.int64)
        return np.ctypeslib.as_array(off, shape=(ne + 1,)).copy()

    def topology_lists(self):
        """Element topologies as a list of arrays (elements may differ in size)."""
        off = ctypes.POINTER(ctypes.c_int64)()
        p = _c_i32p()
        _check(load_library().mgpu_element_topology_csr(self._h, ctypes.byref(off), ctypes.byref(p)), "topology_csr")
        ne = self.n_elements
        if ne == 0:
            return []
        o = np.ctypeslib.as_array(off, shape=(ne + 1,)).copy()
        ids = np.ctypeslib.as_array(p, shape=(int(o[-1]),)).copy() if o[-1] > 0 else np.zeros(0, np.int32)
        return [ids[o[e]:o[e + 1]] for e in range(ne)]

    def compute_jacobian_and_residual(self, u_adof):
        """K values (CSR, pattern()) and R for adof vector u (host arrays)."""
        u = np.ascontiguousarray(u_adof, dtype=np.float64)
        if u.shape != (self.n_rows,):
            raise ValueError("u must have %d entries" % self.n_rows)
        vals = np.empty(self.nnz, dtype=np.float64)
        rhs = np.empty(self.n_rows, dtype=np.float64)
        _check(load_library().mgpu_set_assemble_host(self._h, _ptr(u, ctypes.c_double), _ptr(vals, ctypes.c_double),
                                                     _ptr(rhs, ctypes.c_double)), "mgpu_set_assemble_host")
        return vals, rhs

    def assemble_device(self, u_buf, vals_buf, rhs_buf):
        _check(load_library().mgpu_set_assemble(self._h, u_buf.ptr, vals_buf.ptr, rhs_buf.ptr), "mgpu_set_assemble")

    def time_assemble(self, u_buf, vals_buf, rhs_buf, iters):
        per = ctypes.c_double()
        km = (ctypes.c_double * 5)()
        _check(load_library().mgpu_set_time_assemble(self._h, u_buf.ptr, vals_buf.ptr, rhs_buf.ptr, iters,
                                                     ctypes.byref(per), km), "mgpu_set_time_assemble")
        return per.value, list(km)

    def tuning_variants(self):
        """Rows per workgroup of each stencil variant (second half: direct stores, no LDS staging)."""
        n = ctypes.c_int32()
        rows = (ctypes.c_int32 * 8)()
        _check(load_library().mgpu_set_tuning_variants(self._h, ctypes.byref(n), rows), "mgpu_set_tuning_variants")
        return list(rows)[: n.value]

    def tuning(self, stencil_variant=None, cut_order=None):
        """Get, or set and return, the launch shape (stencil variant, cut order)."""
        lib = load_library()
        v, o = ctypes.c_int32(), ctypes.c_int32()
        _check(lib.mgpu_set_get_tuning(self._h, ctypes.byref(v), ctypes.byref(o)), "mgpu_set_get_tuning")
        if stencil_variant is not None or cut_order is not None:
            nv = v.value if stencil_variant is None else stencil_variant
            no = o.value if cut_order is None else cut_order
            _check(lib.mgpu_set_tuning(self._h, nv, no), "mgpu_set_tuning")
            return nv, no
        return v.value, o.value

    def autotune(self, u_buf, vals_buf, rhs_buf, iters=5, ignore_cache=False):
        """Time every launch shape and keep the fastest; returns {(variant, order): ms}."""
        nv = len(self.tuning_variants())
        ms = np.zeros(2 * nv, dtype=np.float64)
        _check(load_library().mgpu_set_autotune(self._h, u_buf.ptr, vals_buf.ptr, rhs_buf.ptr, iters,
                                                1 if ignore_cache else 0, _ptr(ms, ctypes.c_double)),
               "mgpu_set_autotune")
        return {(k // 2, k % 2): float(ms[k]) for k in range(2 * nv) if ms[k] >= 0.0}

    def stats(self):
        s = (ctypes.c_int64 * 5)()
        _check(load_library().mgpu_set_stats(self._h, s), "mgpu_set_stats")
        return list(s)

    def statistics(self):
        """mgpu_set_statistics: stats() plus [5] points and [6] sub-cells skipped for detJ < -1e-6."""
        s = (ctypes.c_int64 * 7)()
        _check(load_library().mgpu_set_statistics(self._h, s, 7), "mgpu_set_statistics")
        return list(s)

    def close(self):
        if self._h:
            load_library().mgpu_set_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """One assemble call of each set captured into a HIP graph, replayed with one launch."""

    def __init__(self, sets, u_bufs, vals_bufs, rhs_bufs, n_branches=0):
        lib = load_library()
        n = len(sets)
        if not (len(u_bufs) == len(vals_bufs) == len(rhs_bufs) == n):
            raise ValueError("one u, vals and rhs buffer per set")
        vp = ctypes.c_void_p
        hs = (vp * n)(*[s._h for s in sets])
        us = (vp * n)(*[b.ptr for b in u_bufs])
        vs = (vp * n)(*[b.ptr for b in vals_bufs])
        rs = (vp * n)(*[b.ptr for b in rhs_bufs])
        self._keep = (list(sets), list(u_bufs), list(vals_bufs), list(rhs_bufs))
        self._h = vp()
        _check(lib.mgpu_batch_create(hs, n, us, vs, rs, n_branches, None, ctypes.byref(self._h)), "mgpu_batch_create")

    def launch(self):
        _check(load_library().mgpu_batch_launch(self._h), "mgpu_batch_launch")

    def synchronize(self):
        _check(load_library().mgpu_batch_synchronize(self._h), "mgpu_batch_synchronize")

    def time(self, iters):
        per = ctypes.c_double()
        nb = ctypes.c_int32()
        _check(load_library().mgpu_batch_time(self._h, iters, ctypes.byref(per), ctypes.byref(nb)), "mgpu_batch_time")
        return per.value, nb.value

    def close(self):
        if self._h:
            load_library().mgpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    _check(load_library().mgpu_synchronize(), "mgpu_synchronize")


def device_count():
    n = ctypes.c_int32()
    load_library().mgpu_device_count(ctypes.byref(n))
    return n.value
