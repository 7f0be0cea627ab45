"""TEST INFRASTRUCTURE ONLY -- ctypes binding of oracle/liboracle.so, the FP64 CPU restatement of the
reference assembly.  Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
only; the product never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "all"], check=True)


BENCH_LIB_PATH = os.path.join(_HERE, "liboracle_bench.so")


def use_bench_build():
    """Load the -O3 -march=x86-64-v3 build (CPU baseline timing) instead of the checker build;
    must be called before any other use of the oracle in the process."""
    global LIB_PATH
    if _lib is not None and LIB_PATH != BENCH_LIB_PATH:
        raise RuntimeError("the oracle checker build is already loaded")
    LIB_PATH = BENCH_LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle not built at %s (make -C oracle)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_build.argtypes = [vp]
    lib.oracle_build.restype = vp
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_counts.argtypes = [vp, vp]
    lib.oracle_get_setup.argtypes = [vp] * 10
    lib.oracle_assemble.argtypes = [vp, vp, vp, vp]
    lib.oracle_assemble.restype = ctypes.c_int
    lib.oracle_assemble_desc.argtypes = [vp, vp, vp, ctypes.c_int32, vp, vp, vp, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int32]
    lib.oracle_assemble_desc.restype = ctypes.c_int64
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OracleProblem:
    """Independent set-up (XTK cut, adofs, graph) + FP64 assembly of one workload spec."""

    def __init__(self, spec):
        lib = load()
        self._h = lib.oracle_build(ctypes.byref(spec))
        if not self._h:
            raise RuntimeError("oracle_build: " + lib.oracle_last_error().decode())
        c = np.zeros(7, dtype=np.int64)
        lib.oracle_counts(self._h, _p(c))
        (self.n_elements, self.n_cut, self.n_ig, self.n_adofs_per_type, self.n_rows, self.nnz,
         self.n_t) = [int(x) for x in c]
        self.dim = spec.dim
        self.iwg = spec.iwg
        self.nb = 27 if getattr(spec, "lagrange_order", 1) == 2 else (8 if spec.dim == 3 else 4)

    def setup(self):
        dim = self.dim
        nb = self.nb
        nm = 2 if self.iwg == 1 else 1
        out = {
            "elem_cell": np.zeros(self.n_elements, np.int32),
            "ig_ptr": np.zeros(self.n_elements + 1, np.int32),
            "ig_xi": np.zeros(self.n_ig * (dim + 1) * dim, np.float64),
            "material": np.zeros(self.n_elements * nm, np.float64),
            "t_adof": np.zeros(self.n_t, np.int32),
            "row_ptr": np.zeros(self.n_rows + 1, np.int32),
            "col_idx": np.zeros(self.nnz, np.int32),
            "t_ptr": np.zeros(self.n_elements * nb + 1, np.int32),
            "t_weight": np.zeros(self.n_t, np.float64),
        }
        load().oracle_get_setup(self._h, *[_p(out[k]) for k in
                                           ("elem_cell", "ig_ptr", "ig_xi", "material", "t_adof", "row_ptr",
                                            "col_idx", "t_ptr", "t_weight")])
        out["ip_box"] = np.zeros(self.n_elements * 2 * dim, np.float64)
        lib = load()
        lib.oracle_get_ip_box.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_get_ip_box(self._h, _p(out["ip_box"]))
        return out

    def assemble(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        vals = np.zeros(self.nnz)
        rhs = np.zeros(self.n_rows)
        if load().oracle_assemble(self._h, _p(u), _p(vals), _p(rhs)) != 0:
            raise RuntimeError("oracle_assemble: " + load().oracle_last_error().decode())
        return vals, rhs

    def __del__(self):
        try:
            if self._h:
                load().oracle_destroy(self._h)
                self._h = None
        except Exception:
            pass


def assemble_desc(desc, row_ptr, col_idx, u, e0=0, e1=None, n_threads=1):
    """Reference algorithm on an external set descriptor + graph; returns (vals, rhs, n_ip)."""
    lib = load()
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n_rows = row_ptr.shape[0] - 1
    vals = np.zeros(col_idx.shape[0])
    rhs = np.zeros(n_rows)
    if e1 is None:
        e1 = desc.n_elements
    nip = lib.oracle_assemble_desc(ctypes.byref(desc), _p(row_ptr), _p(col_idx), n_rows, _p(u), _p(vals), _p(rhs),
                                   e0, e1, n_threads)
    if nip < 0:
        raise RuntimeError("oracle_assemble_desc: " + lib.oracle_last_error().decode())
    return vals, rhs, nip


def shape(dim, xi, nb=None):
    """Lagrange QUAD4/HEX8 (or HEX27 with nb=27) N [n][nb] and dN/dxi [n][dim][nb] at points xi [n][dim]."""
    lib = load()
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, dim)
    n = xi.shape[0]
    if nb == 27:
        lib.oracle_shape_hex27.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        N = np.zeros((n, 27))
        dN = np.zeros((n, 3, 27))
        lib.oracle_shape_hex27(n, _p(xi), _p(N), _p(dN))
        return N, dN
    lib.oracle_shape.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    nb = 8 if dim == 3 else 4
    N = np.zeros((n, nb))
    dN = np.zeros((n, dim, nb))
    lib.oracle_shape(dim, n, _p(xi), _p(N), _p(dN))
    return N, dN


RULES = {"QUAD_2x2": (0, 2), "HEX_2x2x2": (1, 3), "TRI_7": (2, 2), "TET_11": (3, 3), "HEX_3x3x3": (4, 3),
         "TET_35": (5, 3)}


def rule(name):
    """Points [n][dim] and weights [n] of a quadrature rule of the oracle."""
    lib = load()
    lib.oracle_rule.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_rule.restype = ctypes.c_int
    rid, dim = RULES[name]
    xi = np.zeros(64 * 3)
    w = np.zeros(64)
    n = lib.oracle_rule(rid, _p(xi), _p(w))
    return xi[: n * dim].reshape(n, dim), w[:n]


def element(problem, e, u):
    """(ids, K_e, R_e) of element e of an OracleProblem, before the scatter."""
    lib = load()
    lib.oracle_element.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_element.restype = ctypes.c_int
    u = np.ascontiguousarray(u, dtype=np.float64)
    nmax = 3 * 64   # fields x adofs per type of the largest element
    ids = np.zeros(nmax, np.int32)
    Ke = np.zeros(nmax * nmax)
    Re = np.zeros(nmax)
    na = lib.oracle_element(problem._h, e, _p(u), _p(ids), _p(Ke), _p(Re))
    if na < 0:
        raise RuntimeError("oracle_element: " + lib.oracle_last_error().decode())
    return ids[:na].copy(), Ke[: na * na].reshape(na, na).copy(), Re[:na].copy()


def assemble_list(desc, row_ptr, col_idx, u, elems, n_threads=1):
    """Reference algorithm over an explicit element list; returns (vals, rhs, n_ip)."""
    lib = load()
    lib.oracle_assemble_list.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_int32]
    lib.oracle_assemble_list.restype = ctypes.c_int64
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    elems = np.ascontiguousarray(elems, dtype=np.int64)
    n_rows = row_ptr.shape[0] - 1
    vals = np.zeros(col_idx.shape[0])
    rhs = np.zeros(n_rows)
    nip = lib.oracle_assemble_list(ctypes.byref(desc), _p(row_ptr), _p(col_idx), n_rows, _p(u), _p(vals), _p(rhs),
                                   _p(elems), elems.shape[0], n_threads)
    if nip < 0:
        raise RuntimeError("oracle_assemble_list: " + lib.oracle_last_error().decode())
    return vals, rhs, nip


def nh_cell(dim, gid, phi, x, phase):
    """Node hierarchy of one IG cell (tet / tri): sub-cells of `phase` as [n][dim+1][3]."""
    lib = load()
    lib.oracle_nh_cell.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_int]
    lib.oracle_nh_cell.restype = ctypes.c_int
    g = np.ascontiguousarray(gid, dtype=np.uint32)
    f = np.ascontiguousarray(phi, dtype=np.float64)
    xx = np.zeros((dim + 1, 3))
    xx[:, :dim] = np.asarray(x, dtype=np.float64)[:, :dim]
    out = np.zeros(16 * (dim + 1) * 3)
    n = lib.oracle_nh_cell(dim, _p(g), _p(f), _p(xx), phase, _p(out), 16)
    if n < 0:
        raise RuntimeError("oracle_nh_cell: " + lib.oracle_last_error().decode())
    return out[: n * (dim + 1) * 3].reshape(n, dim + 1, 3)


def facets(problem, where, nvs):
    """Facets of an OracleProblem (its own extraction): (elem, xi [n][nvs][dim], normal [n][dim])."""
    lib = load()
    lib.oracle_facets.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int64]
    lib.oracle_facets.restype = ctypes.c_int64
    n = lib.oracle_facets(problem._h, where, nvs, None, None, None, 0)
    if n < 0:
        raise RuntimeError("oracle_facets: " + lib.oracle_last_error().decode())
    dim = problem.dim
    elem = np.zeros(max(n, 1), np.int32)
    xi = np.zeros(max(n, 1) * nvs * dim)
    nr = np.zeros(max(n, 1) * dim)
    lib.oracle_facets(problem._h, where, nvs, _p(elem), _p(xi), _p(nr), n)
    return elem[:n].copy(), xi[: n * nvs * dim].reshape(n, nvs, dim), nr[: n * dim].reshape(n, dim)


def side_assemble(desc, row_ptr, col_idx, iwg, nvs, elem, xi, normal, alpha, value, u, vals, rhs):
    """Reference algorithm of a side set on descriptor `desc`, ADDED into vals / rhs (numpy, in place)."""
    lib = load()
    lib.oracle_side_assemble.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
    lib.oracle_side_assemble.restype = ctypes.c_int
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    elem = np.ascontiguousarray(elem, np.int32)
    xi = np.ascontiguousarray(xi, np.float64)
    normal = np.ascontiguousarray(normal, np.float64)
    value = np.ascontiguousarray(list(value) + [0.0] * (3 - len(value)), np.float64)
    u = np.ascontiguousarray(u, np.float64)
    assert vals.dtype == np.float64 and rhs.dtype == np.float64 and vals.flags.c_contiguous and rhs.flags.c_contiguous
    if lib.oracle_side_assemble(ctypes.byref(desc), _p(row_ptr), _p(col_idx), iwg, nvs, len(elem), _p(elem), _p(xi),
                                _p(normal), alpha, _p(value), _p(u), _p(vals), _p(rhs)) != 0:
        raise RuntimeError("oracle_side_assemble: " + lib.oracle_last_error().decode())


def ghost_facets(problem):
    """Ghost facets of an OracleProblem (its own extraction): (leader, follower, axis) arrays."""
    lib = load()
    lib.oracle_ghost_facets.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    lib.oracle_ghost_facets.restype = ctypes.c_int64
    n = lib.oracle_ghost_facets(problem._h, None, 0)
    trip = np.zeros(max(n, 1) * 3, np.int32)
    lib.oracle_ghost_facets(problem._h, _p(trip), n)
    trip = trip[: 3 * n].reshape(n, 3)
    return trip[:, 0].copy(), trip[:, 1].copy(), trip[:, 2].copy()


def ghost_assemble(desc, row_ptr, col_idx, leader, follower, xi, normal, alpha, u, vals, rhs):
    """Reference algorithm of ghost facets on descriptor `desc` (xi: the leader's face vertices,
    normal: its outward normal), ADDED into vals / rhs (numpy, in place)."""
    lib = load()
    lib.oracle_ghost_assemble.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_ghost_assemble.restype = ctypes.c_int
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    leader = np.ascontiguousarray(leader, np.int32)
    follower = np.ascontiguousarray(follower, np.int32)
    xi = np.ascontiguousarray(xi, np.float64)
    normal = np.ascontiguousarray(normal, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    assert vals.dtype == np.float64 and rhs.dtype == np.float64 and vals.flags.c_contiguous and rhs.flags.c_contiguous
    if lib.oracle_ghost_assemble(ctypes.byref(desc), _p(row_ptr), _p(col_idx), len(leader), _p(leader), _p(follower),
                                 _p(xi), _p(normal), alpha, _p(u), _p(vals), _p(rhs)) != 0:
        raise RuntimeError("oracle_ghost_assemble: " + lib.oracle_last_error().decode())
