"""Shared test helpers: workload specs of the configurations, lattice geometry, CSR utilities."""
import ctypes
import os
import struct

import numpy as np

import moris_amd as M
from oracle import py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = 20261202


def load_bin(name):
    """Reference fixture: uint32 rows, uint32 cols, FP64 column-major
    (projects/LINALG/src/fn_load_matrix_from_binary_file.hpp)."""
    with open(os.path.join(GOLDEN, name + ".bin"), "rb") as f:
        b = f.read()
    r, c = struct.unpack("<II", b[:8])
    a = np.frombuffer(b[8:8 + 8 * r * c], dtype="<f8")
    return a.reshape(c, r).T.copy()


def spec_sphere3(n, iwg=M.IWG_SPATIALDIFF_BULK, r=0.3, k_sigma=0.5, phase=1):
    return M.make_spec(3, (n, n, n), iwg=iwg, spheres=[[0.5, 0.5, 0.5, r]], k_sigma=k_sigma, seed=SEED + 1,
                       material_phase=phase)


def spec_circle2(n, iwg=M.IWG_STRUC_LINEAR_BULK, r=0.25, phase=1):
    return M.make_spec(2, (n, n), iwg=iwg, spheres=[[0.5, 0.5, 0.0, r]], seed=SEED, material_phase=phase)


def random_spheres(n_spheres, seed, rmin=0.05, rmax=0.12):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(n_spheres, 3))
    r = rng.uniform(rmin, rmax, size=(n_spheres, 1))
    return np.hstack([c, r])


def two_bodies(dim, n, r_cells=2.5, gap_cells=0.5):
    """Two spheres (circles) on the diagonal through the centre of cell n//2 (each side), material
    inside: the cell holds one corner cap of each, so it splits into two subphases."""
    h = 1.0 / n
    ctr = np.full(dim, (n // 2 + 0.5) * h)
    u = np.ones(dim) / np.sqrt(dim)
    r = r_cells * h
    a = ctr - (gap_cells * h + r) * u
    b = ctr + (gap_cells * h + r) * u
    pad = [0.0] * (3 - dim)
    return [list(a) + pad + [r], list(b) + pad + [r]]


def adof_coords(problem_arrays, dim, n):
    """Coordinates of every adof (used lattice nodes in ascending node order) of a unit-box lattice."""
    ec = problem_arrays["elem_cell"].astype(np.int64)
    nx, ny = n[0], n[1]
    i = ec % nx
    j = (ec // nx) % ny
    k = ec // (nx * ny) if dim == 3 else np.zeros_like(ec)
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)] if dim == 3 else \
        [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    nodes = []
    for o in offs:
        nodes.append((i + o[0]) + (nx + 1) * ((j + o[1]) + (ny + 1) * (k + o[2])))
    nodes = np.unique(np.concatenate(nodes))
    ii = nodes % (nx + 1)
    jj = (nodes // (nx + 1)) % (ny + 1)
    kk = nodes // ((nx + 1) * (ny + 1))
    h = [1.0 / nn for nn in n]
    X = np.stack([ii * h[0], jj * h[1]] + ([kk * h[2]] if dim == 3 else []), axis=1)
    return nodes, X


def csr_matvec(row_ptr, col_idx, vals, x):
    rows = np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr))
    y = np.zeros(len(row_ptr) - 1)
    np.add.at(y, rows, vals * x[col_idx])
    return y


def csr_to_scipy(row_ptr, col_idx, vals, n):
    import scipy.sparse as sp
    return sp.csr_matrix((vals, col_idx, row_ptr), shape=(n, n))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# facet vertex counts of the side geometries (MGPU_GEOM_TRI3, MGPU_GEOM_QUAD4, MGPU_GEOM_LINE2)
NV = {2: 3, 0: 4, 5: 2}


def oracle_side_system(spec, sides, u):
    """Oracle bulk assembly plus side sets [(iwg, where, geom, alpha, value)] at u."""
    Q = O.OracleProblem(spec)
    vals, R = Q.assemble(u)
    s = Q.setup()
    desc = oracle_desc(Q, s)
    for iwg, where, geom, alpha, value in sides:
        e, x, n = O.facets(Q, where, NV[geom])
        O.side_assemble(desc[0], s["row_ptr"], s["col_idx"], iwg, NV[geom], e, x, n, alpha, value, u, vals, R)
    return Q, s, vals, R


def oracle_desc(Q, s):
    """A set descriptor over the oracle's own set-up arrays (for the side assembly)."""
    d = M.SetDesc()
    dim = Q.dim
    d.spatial_dim = dim
    d.iwg = Q.iwg
    d.ip_geometry = (M.GEOM_HEX27 if Q.nb == 27 else M.GEOM_HEX8) if dim == 3 else M.GEOM_QUAD4
    d.ig_geometry = M.GEOM_TET4 if dim == 3 else M.GEOM_TRI3
    d.n_elements = Q.n_elements
    d.n_bases = Q.nb
    d.n_adofs_per_type = Q.n_adofs_per_type
    keep = {k: np.ascontiguousarray(s[k]) for k in ("ip_box", "material", "t_ptr", "t_adof", "t_weight", "ig_ptr",
                                                      "ig_xi", "elem_cell")}
    for k, a in keep.items():
        ct = ctypes.c_int32 if a.dtype == np.int32 else ctypes.c_double
        setattr(d, k, a.ctypes.data_as(ctypes.POINTER(ct)))
    return d, keep
