// gpu_setup.hip -- the set-up of a lattice set on the GPU: level set at the lattice nodes, XTK cut of
// every intersected cell, element list, adof numbering, T rows (p = 1 nodal, p = 2 level-0
// B-splines), element topology and CSR pattern.  It produces the same arrays as the host path
// (setup/problem.cpp + set/set_host.cpp) for HEX8 / QUAD4 IP cells without refinement, and for
// HEX27 IP cells with p = 2 rows (the cut on the cell's corners, the rows at the 27 nodes).
//
//   intersected cells          Geometry_Engine::is_intersected, cl_GEN_Geometry_Engine.cpp:278-300
//   cut of one cell            common/xtk_cut.hpp (regular subdivision + node hierarchy,
//                              cl_XTK_Regular_Subdivision_Interface.hpp:110-360,
//                              cl_XTK_Node_Hierarchy_Interface.cpp:160-733)
//   adof ids                   Dof_Manager::set_owned_adofs_ids_by_type, cl_MSI_Dof_Manager.cpp:703-745
//   T rows                     T_Matrix_Base::evaluate, cl_HMR_T_Matrix_Base.cpp:306-457 (level 0)
//   pattern                    union of element topologies, cl_Sparse_Matrix_EpetraFECrs.cpp:195-240
//
// Kernels (one thread per node, cell, element or row; scans by hipCUB):
//   k_node_phi     level set at the lattice nodes
//   k_cell_class   0 void, 1 uncut material, 2 intersected
//   k_cut          one thread per intersected cell: the 15 subdivision vertices (face nodes from the
//                  lowest intersected cell sharing the face, as the host does), the cut into a slot
//   k_elem_fill    element list in lattice order, IP boxes, material, vertex level set, sub-cell counts
//   k_copy_sub     slots -> ig_xi, one block per intersected cell
//   k_basis_used, k_basis_adof   bases some element uses, numbered in ascending basis index
//   k_trows        T rows and element topology
//   k_rows         CSR row lengths, then columns (union of the topologies around each basis)
// The file is compiled without FMA contraction (as the host set-up) so that the level set and the
// cut's intersection points round as on the host.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/moris_gpu_setup.h"
#include "../common/fem_tables.hpp"
#include "../common/levelset.hpp"
#include "../common/xtk_cut.hpp"
#include "../set/set_state.hpp"
#include "internal.hpp"

#define SU_TRY(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace mgpu {
namespace su {

constexpr int kT = 256;
constexpr int kHex27RowEntries = 343;   // T entries of the 27 rows of a HEX27 element, p = 2

struct Lat {
    int dim, nb, order, phase, iwg;
    int nb_ip;   // IP-cell bases: nb, or 27 for HEX27 cells
    // bases of a cell per axis (p + 1, or 3 for nodal HEX27 rows) and the basis-index step of a
    // cell (1; 2 for nodal HEX27 rows, whose bases are the nodes of the half-step grid)
    int bper, bstride;
    int n[3];
    double lo[3], h[3];
    int64_t n_cells, n_nodes, n_faces;
    LevelSet ls;   // spheres point to device memory
    // material
    double E, nu, k_mu, k_sigma;
    uint64_t seed;
    // basis lattice: p = 1 the nodes (n+1 per axis), p = 2 the splines (n+2 per axis)
    int nbs[3];
};

__device__ inline int64_t node_id(const Lat& L, int i, int j, int k) {
    return (int64_t)i + (int64_t)(L.n[0] + 1) * ((int64_t)j + (int64_t)(L.n[1] + 1) * k);
}

__device__ inline void node_x(const Lat& L, int i, int j, int k, double* x) {
    const int ijk[3] = {i, j, k};
    for (int d = 0; d < L.dim; ++d) x[d] = L.lo[d] + (double)ijk[d] * L.h[d];
    if (L.dim == 2) x[2] = 0.0;
}

// physical point of parametric xi in cell (i, j, k): N(xi) . X, in the host's order of operations
__device__ inline void cell_point(const Lat& L, int i, int j, int k, const double* xi, double* x) {
    double N[8];
    if (L.dim == 3) {
        for (int v = 0; v < 8; ++v)
            N[v] = 0.125 * (1.0 + xi[0] * kHex8XiHat[v][0]) * (1.0 + xi[1] * kHex8XiHat[v][1]) *
                   (1.0 + xi[2] * kHex8XiHat[v][2]);
    } else {
        for (int v = 0; v < 4; ++v) N[v] = 0.25 * (1.0 + xi[0] * kQuad4XiHat[v][0]) * (1.0 + xi[1] * kQuad4XiHat[v][1]);
    }
    for (int d = 0; d < 3; ++d) x[d] = 0.0;
    for (int v = 0; v < L.nb; ++v) {
        double xv[3];
        if (L.dim == 3) node_x(L, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2], xv);
        else node_x(L, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0, xv);
        for (int d = 0; d < L.dim; ++d) x[d] += N[v] * xv[d];
    }
}

__device__ inline void cell_ijk(const Lat& L, int64_t c, int* ijk) {
    ijk[0] = (int)(c % L.n[0]);
    ijk[1] = (int)((c / L.n[0]) % L.n[1]);
    ijk[2] = L.dim == 3 ? (int)(c / ((int64_t)L.n[0] * L.n[1])) : 0;
}

__device__ inline int64_t cell_node(const Lat& L, int64_t c, int v) {
    int ijk[3];
    cell_ijk(L, c, ijk);
    return L.dim == 3 ? node_id(L, ijk[0] + kHex8Off[v][0], ijk[1] + kHex8Off[v][1], ijk[2] + kHex8Off[v][2])
                      : node_id(L, ijk[0] + kQuad4Off[v][0], ijk[1] + kQuad4Off[v][1], 0);
}

__global__ void k_node_phi(Lat L, double* phi) {
    const int64_t nid = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (nid >= L.n_nodes) return;
    const int i = (int)(nid % (L.n[0] + 1));
    const int j = (int)((nid / (L.n[0] + 1)) % (L.n[1] + 1));
    const int k = L.dim == 3 ? (int)(nid / ((int64_t)(L.n[0] + 1) * (L.n[1] + 1))) : 0;
    double x[3];
    node_x(L, i, j, k, x);
    phi[nid] = L.ls(x);
}

__global__ void k_cell_class(Lat L, const double* phi, int8_t* cls, int32_t* is_cut) {
    const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (c >= L.n_cells) return;
    int r0 = -1;
    bool diff = false;
    for (int v = 0; v < L.nb; ++v) {
        const int r = ls_region(phi[cell_node(L, c, v)]);
        if (v == 0) r0 = r;
        else if (r != r0) diff = true;
    }
    cls[c] = diff ? 2 : (r0 == L.phase ? 1 : 0);
    is_cut[c] = diff ? 1 : 0;
}

__global__ void k_compact_cut(Lat L, const int32_t* is_cut, const int32_t* cut_pos, int32_t* cut_cell) {
    const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (c >= L.n_cells) return;
    if (is_cut[c]) cut_cell[cut_pos[c]] = (int32_t)c;
}

// One thread per intersected cell: the cut into its slot (kMaxSubTets sub-tets of 12 doubles).
__global__ void k_cut(Lat L, const double* phi, const int8_t* cls, const int32_t* cut_cell, int32_t n_int,
                      double* slots, int32_t* n_sub, int32_t* status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_int) return;
    const int64_t c = cut_cell[t];
    int ijk[3];
    cell_ijk(L, c, ijk);
    const int i = ijk[0], j = ijk[1], k = ijk[2];
    auto inter = [&](int a, int b, int cc) { return cls[a + (int64_t)L.n[0] * (b + (int64_t)L.n[1] * cc)] == 2; };
    int ns = 0, rc;
    if (L.dim == 3) {
        HexSubdivVertex hv[15];
        for (int v = 0; v < 8; ++v) {
            for (int d = 0; d < 3; ++d) hv[v].xi[d] = kHex8XiHat[v][d];
            const int64_t nid = node_id(L, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2]);
            hv[v].phi = phi[nid];
            hv[v].gid = (uint32_t)nid;
        }
        const int64_t nxf = (int64_t)(L.n[0] + 1) * L.n[1] * L.n[2];
        const int64_t nyf = (int64_t)L.n[0] * (L.n[1] + 1) * L.n[2];
        for (int f = 0; f < 6; ++f) {
            HexSubdivVertex& fv = hv[8 + f];
            for (int d = 0; d < 3; ++d) fv.xi[d] = kHexFaceXi[f][d];
            int pi = i, pj = j, pk = k, pf = f;
            int64_t fid;
            switch (f) {
                case 0: if (j > 0 && inter(i, j - 1, k)) { pj = j - 1; pf = 2; }
                        fid = nxf + i + (int64_t)L.n[0] * (j + (int64_t)(L.n[1] + 1) * k); break;
                case 1: fid = (i + 1) + (int64_t)(L.n[0] + 1) * (j + (int64_t)L.n[1] * k); break;
                case 2: fid = nxf + i + (int64_t)L.n[0] * ((j + 1) + (int64_t)(L.n[1] + 1) * k); break;
                case 3: if (i > 0 && inter(i - 1, j, k)) { pi = i - 1; pf = 1; }
                        fid = i + (int64_t)(L.n[0] + 1) * (j + (int64_t)L.n[1] * k); break;
                case 4: if (k > 0 && inter(i, j, k - 1)) { pk = k - 1; pf = 5; }
                        fid = nxf + nyf + i + (int64_t)L.n[0] * (j + (int64_t)L.n[1] * k); break;
                default: fid = nxf + nyf + i + (int64_t)L.n[0] * (j + (int64_t)L.n[1] * (k + 1)); break;
            }
            double x[3];
            const double fxi[3] = {kHexFaceXi[pf][0], kHexFaceXi[pf][1], kHexFaceXi[pf][2]};
            cell_point(L, pi, pj, pk, fxi, x);
            fv.phi = L.ls(x);
            fv.gid = (uint32_t)(L.n_nodes + fid);
        }
        HexSubdivVertex& cv = hv[14];
        const double zero[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; ++d) cv.xi[d] = 0.0;
        double x[3];
        cell_point(L, i, j, k, zero, x);
        cv.phi = L.ls(x);
        cv.gid = (uint32_t)(L.n_nodes + L.n_faces + c);
        rc = cut_hex8(hv, L.phase, reinterpret_cast<double(*)[4][3]>(slots + (size_t)t * kMaxSubTets * 12), &ns);
    } else {
        QuadSubdivVertex qv[5];
        for (int v = 0; v < 4; ++v) {
            for (int d = 0; d < 2; ++d) qv[v].xi[d] = kQuad4XiHat[v][d];
            const int64_t nid = node_id(L, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0);
            qv[v].phi = phi[nid];
            qv[v].gid = (uint32_t)nid;
        }
        const double zero[2] = {0.0, 0.0};
        double x[3];
        cell_point(L, i, j, 0, zero, x);
        qv[4].xi[0] = qv[4].xi[1] = 0.0;
        qv[4].phi = L.ls(x);
        qv[4].gid = (uint32_t)(L.n_nodes + c);
        rc = cut_quad4(qv, L.phase, reinterpret_cast<double(*)[3][2]>(slots + (size_t)t * kMaxSubTets * 12), &ns);
    }
    n_sub[t] = rc == kCutOk ? ns : 0;
    if (rc != kCutOk) atomicOr(status, 1);
}

// is_elem per cell (uncut material, or intersected with material sub-cells)
__global__ void k_elem_flag(Lat L, const int8_t* cls, const int32_t* cut_pos, const int32_t* n_sub, int32_t* is_elem) {
    const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (c >= L.n_cells) return;
    is_elem[c] = cls[c] == 1 || (cls[c] == 2 && n_sub[cut_pos[c]] > 0) ? 1 : 0;
}

__global__ void k_elem_fill(Lat L, const double* phi, const int8_t* cls, const int32_t* cut_pos, const int32_t* n_sub,
                            const int32_t* is_elem, const int32_t* elem_pos, int32_t* elem_cell, int32_t* cell_elem,
                            int32_t* elem_nsub, int32_t* elem_slot, double* ip_box, double* material,
                            double* vertex_phi) {
    const int64_t c = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (c >= L.n_cells) return;
    if (!is_elem[c]) {
        cell_elem[c] = -1;
        return;
    }
    const int32_t e = elem_pos[c];
    cell_elem[c] = e;
    elem_cell[e] = (int32_t)c;
    const bool cut = cls[c] == 2;
    elem_nsub[e] = cut ? n_sub[cut_pos[c]] : 0;
    elem_slot[e] = cut ? cut_pos[c] : -1;
    int ijk[3];
    cell_ijk(L, c, ijk);
    const int d2 = 2 * L.dim;
    for (int d = 0; d < L.dim; ++d) {
        ip_box[(size_t)e * d2 + d] = L.lo[d] + (double)ijk[d] * L.h[d];
        ip_box[(size_t)e * d2 + L.dim + d] = L.h[d];
    }
    if (L.nb_ip == 27) {
        // level set at the 27 HEX27 nodes (the cut itself uses the 8 corner values)
        for (int v = 0; v < 27; ++v) {
            double x[3];
            for (int d = 0; d < 3; ++d) x[d] = L.lo[d] + ((double)ijk[d] + 0.5 * kHex27Half[v][d]) * L.h[d];
            vertex_phi[(size_t)e * 27 + v] = L.ls(x);
        }
    } else {
        for (int v = 0; v < L.nb; ++v) vertex_phi[(size_t)e * L.nb + v] = phi[cell_node(L, c, v)];
    }
    if (L.iwg == MGPU_IWG_SPATIALDIFF_BULK) {
        material[e] = L.k_sigma == 0.0 ? exp(L.k_mu) : exp(L.k_mu + L.k_sigma * normal_at(L.seed, (uint64_t)c));
    } else {
        material[2 * (size_t)e] = L.E;
        material[2 * (size_t)e + 1] = L.nu;
    }
}

// One block per intersected cell that is an element: its sub-cells from the slot into ig_xi, the
// block's lanes striding over the values (coalesced on both sides).
__global__ void k_copy_sub(int dim, const int32_t* cut_cell, const int32_t* cell_elem, const int32_t* n_sub,
                           const int32_t* ig_ptr, const double* slots, double* ig_xi) {
    const int t = blockIdx.x;
    const int32_t e = cell_elem[cut_cell[t]];
    if (e < 0) return;
    const int per = (dim + 1) * dim;
    const double* src = slots + (size_t)t * kMaxSubTets * 12;
    double* dst = ig_xi + (size_t)ig_ptr[e] * per;
    const int n = n_sub[t] * per;
    for (int q = threadIdx.x; q < n; q += blockDim.x) dst[q] = src[q];
}

// Basis lattice: p = 1 the lattice nodes (n + 1 per axis), p = 2 the quadratic B-splines (n + 2 per
// axis, spline s supported on cells s - 2 .. s).  Cell i of an axis carries the bases i .. i + per - 1
// (per = p + 1).  Index x fastest.
__device__ inline int64_t basis_index(const Lat& L, int a, int b, int c) {
    return (int64_t)a + (int64_t)L.nbs[0] * ((int64_t)b + (int64_t)L.nbs[1] * c);
}

__global__ void k_basis_used(Lat L, int32_t n_el, const int32_t* elem_cell, int32_t* used) {
    const int32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= n_el) return;
    int ijk[3];
    cell_ijk(L, elem_cell[e], ijk);
    const int per = L.bper, pz = L.dim == 3 ? per : 1, st = L.bstride;
    for (int c = 0; c < pz; ++c)
        for (int b = 0; b < per; ++b)
            for (int a = 0; a < per; ++a) used[basis_index(L, st * ijk[0] + a, st * ijk[1] + b, st * ijk[2] + c)] = 1;
}

__global__ void k_basis_adof(int64_t n, const int32_t* used, const int32_t* pos, int32_t* adof, int32_t* basis_of) {
    const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (s >= n) return;
    adof[s] = used[s] ? pos[s] : -1;
    if (used[s]) basis_of[pos[s]] = (int32_t)s;
}

// T rows and topology of element e.  p = 1: one entry of weight 1 per vertex (its node).  p = 2:
// a vertex at lattice node m of an axis carries the splines m and m + 1 with 1/2 each; the row is
// the tensor product, x fastest, weight the product of the axis weights (1/8 in 3D, 1/4 in 2D).
// HEX27 cells (p = 2): a node at half-step coordinate g of an axis carries the splines m, m + 1 with
// 1/2 each when g = 2 m, and m, m + 1, m + 2 with 1/8, 3/4, 1/8 when g = 2 m + 1 (the cell centre);
// the 27 rows of an element hold (2 + 3 + 2)^3 = 343 entries.
// Topology: the element's bases in ascending index per dof type, type f offset by f N.
__global__ void k_trows(Lat L, int32_t n_el, const int32_t* elem_cell, const int32_t* basis_adof, int32_t N, int nf,
                        int32_t* t_ptr, int32_t* t_adof, double* t_weight, int32_t* topo) {
    const int32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= n_el) return;
    int ijk[3];
    cell_ijk(L, elem_cell[e], ijk);
    if (L.nb_ip == 27 && L.order == 1) {
        // nodal HEX27 rows: node v of the element is the half-step grid node 2 ijk + kHex27Half[v]
        for (int v = 0; v < 27; ++v) {
            const int64_t r = (int64_t)e * 27 + v;
            t_ptr[r] = (int32_t)r;
            t_adof[r] = basis_adof[basis_index(L, 2 * ijk[0] + kHex27Half[v][0], 2 * ijk[1] + kHex27Half[v][1],
                                               2 * ijk[2] + kHex27Half[v][2])];
            t_weight[r] = 1.0;
        }
        if (e == n_el - 1) t_ptr[(int64_t)n_el * 27] = (int32_t)((int64_t)n_el * 27);
    } else if (L.nb_ip == 27) {
        int64_t pos = (int64_t)e * kHex27RowEntries;
        for (int v = 0; v < 27; ++v) {
            t_ptr[(int64_t)e * 27 + v] = (int32_t)pos;
            int na[3], sm[3];
            double wa[3][3];
            for (int d = 0; d < 3; ++d) {
                const int g = 2 * ijk[d] + kHex27Half[v][d];
                sm[d] = g / 2;
                if ((g & 1) == 0) {
                    na[d] = 2;
                    wa[d][0] = wa[d][1] = 0.5;
                    wa[d][2] = 0.0;
                } else {
                    na[d] = 3;
                    wa[d][0] = 0.125;
                    wa[d][1] = 0.75;
                    wa[d][2] = 0.125;
                }
            }
            for (int c = 0; c < na[2]; ++c)
                for (int b = 0; b < na[1]; ++b)
                    for (int a = 0; a < na[0]; ++a) {
                        t_adof[pos] = basis_adof[basis_index(L, sm[0] + a, sm[1] + b, sm[2] + c)];
                        t_weight[pos] = wa[0][a] * wa[1][b] * wa[2][c];
                        ++pos;
                    }
        }
        if (e == n_el - 1) t_ptr[(int64_t)n_el * 27] = (int32_t)((int64_t)n_el * kHex27RowEntries);
    } else {
    const int per_row = L.order == 1 ? 1 : (L.dim == 3 ? 8 : 4);
    for (int v = 0; v < L.nb; ++v) {
        const int ox = L.dim == 3 ? kHex8Off[v][0] : kQuad4Off[v][0];
        const int oy = L.dim == 3 ? kHex8Off[v][1] : kQuad4Off[v][1];
        const int oz = L.dim == 3 ? kHex8Off[v][2] : 0;
        const int64_t r = (int64_t)e * L.nb + v;
        t_ptr[r] = (int32_t)(r * per_row);
        if (L.order == 1) {
            t_adof[r] = basis_adof[basis_index(L, ijk[0] + ox, ijk[1] + oy, ijk[2] + oz)];
            t_weight[r] = 1.0;
            continue;
        }
        int q = 0;
        const int cz = L.dim == 3 ? 2 : 1;
        for (int c = 0; c < cz; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    t_adof[r * per_row + q] = basis_adof[basis_index(L, ijk[0] + ox + a, ijk[1] + oy + b, ijk[2] + oz + c)];
                    t_weight[r * per_row + q] = L.dim == 3 ? 0.5 * 0.5 * 0.5 : 0.5 * 0.5 * 1.0;
                    ++q;
                }
    }
    if (e == n_el - 1) t_ptr[(int64_t)n_el * L.nb] = (int32_t)((int64_t)n_el * L.nb * per_row);
    }
    const int per = L.bper, pz = L.dim == 3 ? per : 1, st = L.bstride;
    const int na_t = per * per * pz;
    int32_t* out = topo + (int64_t)e * nf * na_t;
    int q = 0;
    for (int c = 0; c < pz; ++c)
        for (int b = 0; b < per; ++b)
            for (int a = 0; a < per; ++a, ++q) {
                const int32_t id = basis_adof[basis_index(L, st * ijk[0] + a, st * ijk[1] + b, st * ijk[2] + c)];
                for (int f = 0; f < nf; ++f) out[f * na_t + q] = id + f * N;
            }
}

// Row of basis s (every dof type): the bases sharing a material element with s, i.e. the union of
// the topologies of the elements around s.  Cell i carries the bases st i .. st i + per - 1 (st the
// basis step of a cell), so the cells around s are those with st i in s - per + 1 .. s, and their
// bases lie in the window s - per + 1 .. s + per - 1 (w = 2 per - 1 per axis, w^dim <= 125 bits).
// COUNT: row lengths; else the columns, ascending: type g's window bases offset by g N.
template <bool COUNT>
__global__ void k_rows(Lat L, const int32_t* basis_of, int32_t N, const int32_t* cell_elem, const int32_t* basis_adof,
                       int nf, int32_t* row_len, const int32_t* row_ptr, int32_t* col_idx) {
    const int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (a >= N) return;
    const int64_t s = basis_of[a];
    const int si[3] = {(int)(s % L.nbs[0]), (int)((s / L.nbs[0]) % L.nbs[1]),
                       L.dim == 3 ? (int)(s / ((int64_t)L.nbs[0] * L.nbs[1])) : 0};
    const int per = L.bper, pz = L.dim == 3 ? per : 1, st = L.bstride;
    const int w = 2 * per - 1, wz = L.dim == 3 ? w : 1;
    uint32_t mask[4] = {0u, 0u, 0u, 0u};
    for (int dz = 0; dz < pz; ++dz)
        for (int dy = 0; dy < per; ++dy)
            for (int dx = 0; dx < per; ++dx) {
                // the cell whose first basis is s - d per axis (when a cell starts there)
                const int bx = si[0] - dx, by = si[1] - dy, bz = si[2] - dz;
                if (bx < 0 || by < 0 || bz < 0 || bx % st || by % st || bz % st) continue;
                const int ci = bx / st, cj = by / st, ck = bz / st;
                if (ci >= L.n[0] || cj >= L.n[1] || ck >= L.n[2]) continue;
                if (cell_elem[ci + (int64_t)L.n[0] * (cj + (int64_t)L.n[1] * ck)] < 0) continue;
                // the cell's bases ci + q .. in window coordinates (origin s - per + 1)
                for (int c = 0; c < pz; ++c)
                    for (int b = 0; b < per; ++b)
                        for (int q = 0; q < per; ++q) {
                            const int bit = (q - dx + per - 1) + w * ((b - dy + per - 1) + w * (L.dim == 3 ? c - dz + per - 1 : 0));
                            mask[bit >> 5] |= 1u << (bit & 31);
                        }
            }
    const int m = __popc(mask[0]) + __popc(mask[1]) + __popc(mask[2]) + __popc(mask[3]);
    if (COUNT) {
        for (int f = 0; f < nf; ++f) row_len[(int64_t)f * N + a] = nf * m;
        return;
    }
    for (int f = 0; f < nf; ++f) {
        int32_t* out = col_idx + row_ptr[(int64_t)f * N + a];
        int q = 0;
        for (int z = 0; z < wz; ++z)
            for (int y = 0; y < w; ++y)
                for (int x = 0; x < w; ++x) {
                    const int bit = x + w * (y + w * z);
                    if (!(mask[bit >> 5] & (1u << (bit & 31)))) continue;
                    const int32_t id = basis_adof[basis_index(L, si[0] - per + 1 + x, si[1] - per + 1 + y,
                                                              L.dim == 3 ? si[2] - per + 1 + z : 0)];
                    for (int g = 0; g < nf; ++g) out[g * m + q] = id + g * N;
                    ++q;
                }
    }
}

// Sub-cells the assembly kernels skip (detJ < -1e-6, cl_FEM_Element_Bulk.cpp:291), counted with the
// host set-up's formula: rows X_r - X_2 scaled by half the cell extent.
__global__ void k_count_skip(Lat L, int32_t n_ig, const double* ig_xi, unsigned long long* count) {
    const int32_t t = blockIdx.x * kT + threadIdx.x;
    if (t >= n_ig) return;
    double detJ;
    if (L.dim == 3) {
        const double* X = ig_xi + (size_t)t * 12;
        double J[3][3];
        const int rr[3] = {0, 1, 3};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) J[r][c] = (X[rr[r] * 3 + c] - X[2 * 3 + c]) * (0.5 * L.h[c]);
        detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
    } else {
        const double* X = ig_xi + (size_t)t * 6;
        double J[2][2];
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) J[r][c] = (X[r * 2 + c] - X[2 * 2 + c]) * (0.5 * L.h[c]);
        detJ = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / 2.0;
    }
    if (detJ < kDetJLowerLimit) atomicAdd(count, 1ull);
}

}  // namespace su
}  // namespace mgpu

using namespace mgpu;

struct mgpu_gpu_setup {
    su::Lat L{};
    int nf = 1;
    hipStream_t stream = nullptr;
    int32_t n_el = 0, n_cut = 0, n_ig = 0, N = 0;
    int64_t n_rows = 0, nnz = 0, n_t = 0, na_t = 0;
    double ms = 0.0;
    // device arrays
    int32_t *elem_cell = nullptr, *ig_ptr = nullptr, *t_ptr = nullptr, *t_adof = nullptr, *topo = nullptr;
    int32_t *row_ptr = nullptr, *col_idx = nullptr;
    int32_t *cell_elem = nullptr, *node_adof = nullptr, *basis_of = nullptr;   // lattice maps (p = 1: nodes)
    int64_t n_skip = 0;                                                           // sub-cells with detJ < -1e-6
    double *ig_xi = nullptr, *material = nullptr, *ip_box = nullptr, *t_weight = nullptr, *vertex_phi = nullptr;
    std::vector<void*> allocs;
    // host copies behind mgpu_setup_desc
    std::vector<int32_t> h_elem_cell, h_ig_ptr, h_t_ptr, h_t_adof;
    std::vector<double> h_ig_xi, h_material, h_ip_box, h_t_weight, h_vertex_phi;
    mgpu_gpu_setup() = default;
    mgpu_gpu_setup(const mgpu_gpu_setup&) = delete;
    mgpu_gpu_setup& operator=(const mgpu_gpu_setup&) = delete;
    ~mgpu_gpu_setup() {
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {
template <class T>
mgpu_status dalloc(mgpu_gpu_setup* S, size_t n, T** p) {
    void* q = nullptr;
    SU_TRY(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    S->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return MGPU_OK;
}

// exclusive prefix sum of n int32 values into out (n + 1 entries: the total at out[n])
mgpu_status scan(mgpu_gpu_setup* S, const int32_t* in, int32_t* out, int64_t n, int32_t* total) {
    size_t bytes = 0;
    SU_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, S->stream));
    void* tmp = nullptr;
    SU_TRY(hipMalloc(&tmp, std::max<size_t>(bytes, 8)));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, S->stream);
    int32_t last_in = 0, last_out = 0;
    if (e == hipSuccess && n > 0) e = hipMemcpyAsync(&last_in, in + n - 1, 4, hipMemcpyDeviceToHost, S->stream);
    if (e == hipSuccess && n > 0) e = hipMemcpyAsync(&last_out, out + n - 1, 4, hipMemcpyDeviceToHost, S->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return set_error(MGPU_ERR_HIP, std::string("scan: ") + hipGetErrorString(e));
    *total = n > 0 ? last_in + last_out : 0;
    SU_TRY(hipMemcpyAsync(out + n, total, 4, hipMemcpyHostToDevice, S->stream));
    SU_TRY(hipStreamSynchronize(S->stream));
    return MGPU_OK;
}

unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + su::kT - 1) / su::kT); }
}  // namespace

extern "C" mgpu_status mgpu_setup_build(const mgpu_problem_spec* spec, void* stream, mgpu_gpu_setup** out) {
    MGPU_GUARD_BEGIN
    if (!spec || !out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (spec->dim != 2 && spec->dim != 3) return set_error(MGPU_ERR_INVALID_ARG, "dim must be 2 or 3");
    if (spec->iwg != MGPU_IWG_SPATIALDIFF_BULK && spec->iwg != MGPU_IWG_STRUC_LINEAR_BULK)
        return set_error(MGPU_ERR_UNSUPPORTED, "iwg not supported");
    if (spec->lagrange_order > 2 || spec->refine_levels != 0)
        return set_error(MGPU_ERR_UNSUPPORTED, "the GPU set-up covers HEX8 / QUAD4 and HEX27 IP cells without refinement");
    if (spec->lagrange_order == 2 && spec->dim != 3)
        return set_error(MGPU_ERR_UNSUPPORTED, "the GPU set-up takes HEX27 cells in 3D");
    if (spec->enrich != 0)
        return set_error(MGPU_ERR_UNSUPPORTED, "the GPU set-up numbers one element per cell (no enrichment)");
    const int order = spec->bspline_order <= 1 ? 1 : spec->bspline_order;
    if (order > 2) return set_error(MGPU_ERR_UNSUPPORTED, "bspline_order must be 1 or 2");
    for (int d = 0; d < spec->dim; ++d)
        if (spec->n[d] < 1 || !(spec->hi[d] > spec->lo[d])) return set_error(MGPU_ERR_INVALID_ARG, "bad grid extent");
    if (spec->ls_kind == MGPU_LS_SPHERES && (spec->n_spheres < 1 || !spec->spheres))
        return set_error(MGPU_ERR_INVALID_ARG, "sphere level set without spheres");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(MGPU_ERR_NO_DEVICE, "no HIP device visible");
    auto S = std::make_unique<mgpu_gpu_setup>();
    S->stream = static_cast<hipStream_t>(stream);
    su::Lat& L = S->L;
    L.dim = spec->dim;
    L.nb = L.dim == 3 ? 8 : 4;
    L.nb_ip = spec->lagrange_order == 2 ? 27 : L.nb;
    L.order = order;
    const bool nodal27 = L.nb_ip == 27 && order == 1;
    L.bper = nodal27 ? 3 : order + 1;
    L.bstride = nodal27 ? 2 : 1;
    L.phase = spec->material_phase;
    L.iwg = spec->iwg;
    for (int d = 0; d < 3; ++d) {
        L.n[d] = d < L.dim ? spec->n[d] : 1;
        L.lo[d] = d < L.dim ? spec->lo[d] : 0.0;
        L.h[d] = d < L.dim ? (spec->hi[d] - spec->lo[d]) / (double)spec->n[d] : 1.0;
        L.nbs[d] = d < L.dim ? (nodal27 ? 2 * L.n[d] + 1 : L.n[d] + (order == 1 ? 1 : 2)) : 1;
    }
    L.n_cells = (int64_t)L.n[0] * L.n[1] * L.n[2];
    L.n_nodes = L.dim == 3 ? (int64_t)(L.n[0] + 1) * (L.n[1] + 1) * (L.n[2] + 1) : (int64_t)(L.n[0] + 1) * (L.n[1] + 1);
    L.n_faces = L.dim == 3 ? (int64_t)(L.n[0] + 1) * L.n[1] * L.n[2] + (int64_t)L.n[0] * (L.n[1] + 1) * L.n[2] +
                                 (int64_t)L.n[0] * L.n[1] * (L.n[2] + 1)
                           : 0;
    if (L.n_nodes + L.n_faces + L.n_cells >= (int64_t)UINT32_MAX || L.n_cells >= INT32_MAX)
        return set_error(MGPU_ERR_INVALID_ARG, "grid too large for 32-bit ids");
    L.ls.kind = spec->ls_kind;
    L.ls.dim = L.dim;
    L.ls.n_spheres = spec->ls_kind == MGPU_LS_SPHERES ? spec->n_spheres : 0;
    L.ls.gyroid_scale = spec->gyroid_scale;
    for (int d = 0; d < 4; ++d) L.ls.plane[d] = spec->plane[d];
    L.E = spec->E;
    L.nu = spec->nu;
    L.k_mu = spec->k_mu;
    L.k_sigma = spec->k_sigma;
    L.seed = spec->seed;
    S->nf = spec->iwg == MGPU_IWG_STRUC_LINEAR_BULK ? L.dim : 1;
    const int nf = S->nf;
    mgpu_status rc;
    hipStream_t st = S->stream;
    hipEvent_t e0, e1;
    SU_TRY(hipEventCreate(&e0));
    SU_TRY(hipEventCreate(&e1));
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } evg{e0, e1};
    if (L.ls.n_spheres > 0) {
        double* sp = nullptr;
        if ((rc = dalloc(S.get(), (size_t)4 * L.ls.n_spheres, &sp)) != MGPU_OK) return rc;
        SU_TRY(hipMemcpy(sp, spec->spheres, (size_t)4 * L.ls.n_spheres * sizeof(double), hipMemcpyHostToDevice));
        L.ls.spheres = sp;
    }
    SU_TRY(hipEventRecord(e0, st));
    // 1. level set, classes, intersected cells
    double* phi;
    int8_t* cls;
    int32_t *is_cut, *cut_pos, *cut_cell, *status;
    if ((rc = dalloc(S.get(), (size_t)L.n_nodes, &phi)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells, &cls)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells, &is_cut)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells + 1, &cut_pos)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), 1, &status)) != MGPU_OK) return rc;
    SU_TRY(hipMemsetAsync(status, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(su::k_node_phi, dim3(blocks(L.n_nodes)), dim3(su::kT), 0, st, L, phi);
    hipLaunchKernelGGL(su::k_cell_class, dim3(blocks(L.n_cells)), dim3(su::kT), 0, st, L, phi, cls, is_cut);
    SU_TRY(hipGetLastError());
    int32_t n_int = 0;
    if ((rc = scan(S.get(), is_cut, cut_pos, L.n_cells, &n_int)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_int, &cut_cell)) != MGPU_OK) return rc;
    hipLaunchKernelGGL(su::k_compact_cut, dim3(blocks(L.n_cells)), dim3(su::kT), 0, st, L, is_cut, cut_pos, cut_cell);
    // 2. the cut of every intersected cell into its slot
    double* slots;
    int32_t* n_sub;
    if ((rc = dalloc(S.get(), (size_t)n_int * kMaxSubTets * 12, &slots)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_int, &n_sub)) != MGPU_OK) return rc;
    // one thread per intersected cell in blocks of 64, so that every CU takes some of the cells
    hipLaunchKernelGGL(su::k_cut, dim3((unsigned)std::max<int64_t>(1, (n_int + 63) / 64)), dim3(64), 0, st, L, phi, cls,
                       cut_cell, n_int, slots, n_sub,
                       status);
    SU_TRY(hipGetLastError());
    // 3. elements in lattice order, their boxes, material and sub-cells
    int32_t *is_elem, *elem_pos, *elem_nsub, *elem_slot, *cell_elem;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells, &is_elem)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells + 1, &elem_pos)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)L.n_cells, &cell_elem)) != MGPU_OK) return rc;
    S->cell_elem = cell_elem;
    hipLaunchKernelGGL(su::k_elem_flag, dim3(blocks(L.n_cells)), dim3(su::kT), 0, st, L, cls, cut_pos, n_sub, is_elem);
    int32_t n_el = 0;
    if ((rc = scan(S.get(), is_elem, elem_pos, L.n_cells, &n_el)) != MGPU_OK) return rc;
    S->n_el = n_el;
    const int d2 = 2 * L.dim;
    if ((rc = dalloc(S.get(), (size_t)n_el, &S->elem_cell)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el + 1, &elem_nsub)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el, &elem_slot)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el * d2, &S->ip_box)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el * (nf == 1 ? 1 : 2), &S->material)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el + 1, &S->ig_ptr)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el * L.nb_ip, &S->vertex_phi)) != MGPU_OK) return rc;
    hipLaunchKernelGGL(su::k_elem_fill, dim3(blocks(L.n_cells)), dim3(su::kT), 0, st, L, phi, cls, cut_pos, n_sub,
                       is_elem, elem_pos, S->elem_cell, cell_elem, elem_nsub, elem_slot, S->ip_box, S->material,
                       S->vertex_phi);
    int32_t n_ig = 0;
    if ((rc = scan(S.get(), elem_nsub, S->ig_ptr, n_el, &n_ig)) != MGPU_OK) return rc;
    S->n_ig = n_ig;
    const int per_sub = (L.dim + 1) * L.dim;
    if ((rc = dalloc(S.get(), (size_t)n_ig * per_sub, &S->ig_xi)) != MGPU_OK) return rc;
    if (n_int > 0)
        hipLaunchKernelGGL(su::k_copy_sub, dim3((unsigned)n_int), dim3(64), 0, st, L.dim, cut_cell, cell_elem, n_sub,
                           S->ig_ptr, slots, S->ig_xi);
    SU_TRY(hipGetLastError());
    unsigned long long* skip_count;
    if ((rc = dalloc(S.get(), 1, &skip_count)) != MGPU_OK) return rc;
    SU_TRY(hipMemsetAsync(skip_count, 0, sizeof(unsigned long long), st));
    if (n_ig > 0)
        hipLaunchKernelGGL(su::k_count_skip, dim3(blocks(n_ig)), dim3(su::kT), 0, st, L, n_ig, S->ig_xi, skip_count);
    // 4. adof numbering of the used bases, T rows, topology
    const int64_t n_basis = (int64_t)L.nbs[0] * L.nbs[1] * L.nbs[2];
    if (n_basis >= INT32_MAX) return set_error(MGPU_ERR_INVALID_ARG, "too many bases for 32-bit ids");
    int32_t *used, *bpos, *badof, *basis_of;
    if ((rc = dalloc(S.get(), (size_t)n_basis, &used)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_basis + 1, &bpos)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_basis, &badof)) != MGPU_OK) return rc;
    SU_TRY(hipMemsetAsync(used, 0, (size_t)n_basis * sizeof(int32_t), st));
    hipLaunchKernelGGL(su::k_basis_used, dim3(blocks(n_el)), dim3(su::kT), 0, st, L, n_el, S->elem_cell, used);
    int32_t N = 0;
    if ((rc = scan(S.get(), used, bpos, n_basis, &N)) != MGPU_OK) return rc;
    S->N = N;
    if ((rc = dalloc(S.get(), (size_t)N, &basis_of)) != MGPU_OK) return rc;
    S->node_adof = badof;
    S->basis_of = basis_of;
    hipLaunchKernelGGL(su::k_basis_adof, dim3(blocks(n_basis)), dim3(su::kT), 0, st, n_basis, used, bpos, badof,
                       basis_of);
    const int per_row = order == 1 ? 1 : (L.dim == 3 ? 8 : 4);
    S->n_t = nodal27 ? (int64_t)n_el * 27
                     : (L.nb_ip == 27 ? (int64_t)n_el * su::kHex27RowEntries : (int64_t)n_el * L.nb * per_row);
    const int per = L.bper;
    S->na_t = (int64_t)per * per * (L.dim == 3 ? per : 1);
    const int64_t w = 2 * per - 1;
    if (S->n_t >= INT32_MAX || (int64_t)nf * N * nf * w * w * (L.dim == 3 ? w : 1) >= INT32_MAX)
        return set_error(MGPU_ERR_INVALID_ARG, "T rows or pattern exceed 32-bit offsets");
    if ((rc = dalloc(S.get(), (size_t)n_el * L.nb_ip + 1, &S->t_ptr)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)S->n_t, &S->t_adof)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)S->n_t, &S->t_weight)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)n_el * nf * S->na_t, &S->topo)) != MGPU_OK) return rc;
    if (n_el > 0) {
        hipLaunchKernelGGL(su::k_trows, dim3(blocks(n_el)), dim3(su::kT), 0, st, L, n_el, S->elem_cell, badof, N, nf,
                           S->t_ptr, S->t_adof, S->t_weight, S->topo);
    } else {
        SU_TRY(hipMemsetAsync(S->t_ptr, 0, sizeof(int32_t), st));
    }
    SU_TRY(hipGetLastError());
    // 5. pattern: row lengths, offsets, columns
    S->n_rows = (int64_t)nf * N;
    int32_t* row_len;
    if ((rc = dalloc(S.get(), (size_t)S->n_rows, &row_len)) != MGPU_OK) return rc;
    if ((rc = dalloc(S.get(), (size_t)S->n_rows + 1, &S->row_ptr)) != MGPU_OK) return rc;
    hipLaunchKernelGGL(su::k_rows<true>, dim3(blocks(N)), dim3(su::kT), 0, st, L, basis_of, N, cell_elem, badof, nf,
                       row_len, nullptr, nullptr);
    int32_t nnz = 0;
    // (the int32 scan bounds the pattern at 2^31 - 1 values, as the CSR offsets of the ABI)
    if ((rc = scan(S.get(), row_len, S->row_ptr, S->n_rows, &nnz)) != MGPU_OK) return rc;
    S->nnz = nnz;
    if ((rc = dalloc(S.get(), (size_t)nnz, &S->col_idx)) != MGPU_OK) return rc;
    hipLaunchKernelGGL(su::k_rows<false>, dim3(blocks(N)), dim3(su::kT), 0, st, L, basis_of, N, cell_elem, badof, nf,
                       nullptr, S->row_ptr, S->col_idx);
    SU_TRY(hipGetLastError());
    SU_TRY(hipEventRecord(e1, st));
    SU_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    SU_TRY(hipEventElapsedTime(&ms, e0, e1));
    S->ms = ms;
    int32_t bad = 0;
    SU_TRY(hipMemcpy(&bad, status, sizeof(int32_t), hipMemcpyDeviceToHost));
    unsigned long long nsk = 0;
    SU_TRY(hipMemcpy(&nsk, skip_count, sizeof(nsk), hipMemcpyDeviceToHost));
    S->n_skip = (int64_t)nsk;
    if (bad) return set_error(MGPU_ERR_GEOMETRY, "node-hierarchy template not found in a cell");
    {   // cut elements: elements with sub-cells
        std::vector<int32_t> ns((size_t)n_el + 1);
        SU_TRY(hipMemcpy(ns.data(), S->ig_ptr, ns.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        int32_t nc = 0;
        for (int32_t e = 0; e < n_el; ++e) nc += ns[(size_t)e + 1] > ns[(size_t)e];
        S->n_cut = nc;
    }
    *out = S.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_setup_destroy(mgpu_gpu_setup* s) {
    delete s;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_setup_counts(const mgpu_gpu_setup* s, int64_t c[8]) {
    if (!s || !c) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    c[0] = s->n_el;
    c[1] = s->n_cut;
    c[2] = s->n_ig;
    c[3] = s->N;
    c[4] = s->n_rows;
    c[5] = s->nnz;
    c[6] = s->n_t;
    c[7] = s->L.nb_ip;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_setup_time(const mgpu_gpu_setup* s, double* ms) {
    if (!s || !ms) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *ms = s->ms;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_setup_copy(const mgpu_gpu_setup* s, int32_t* elem_cell, int32_t* ig_ptr, double* ig_xi,
                                       double* material, double* ip_box, double* vertex_phi, int32_t* t_ptr,
                                       int32_t* t_adof, double* t_weight, int32_t* topo, int32_t* row_ptr,
                                       int32_t* col_idx) {
    MGPU_GUARD_BEGIN
    if (!s) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    const int dim = s->L.dim, nm = s->nf == 1 ? 1 : 2;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> mgpu_status {
        if (dst && bytes) SU_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return MGPU_OK;
    };
    mgpu_status rc;
    if ((rc = cp(elem_cell, s->elem_cell, (size_t)s->n_el * 4)) != MGPU_OK) return rc;
    if ((rc = cp(ig_ptr, s->ig_ptr, ((size_t)s->n_el + 1) * 4)) != MGPU_OK) return rc;
    if ((rc = cp(ig_xi, s->ig_xi, (size_t)s->n_ig * (dim + 1) * dim * 8)) != MGPU_OK) return rc;
    if ((rc = cp(material, s->material, (size_t)s->n_el * nm * 8)) != MGPU_OK) return rc;
    if ((rc = cp(ip_box, s->ip_box, (size_t)s->n_el * 2 * dim * 8)) != MGPU_OK) return rc;
    if ((rc = cp(vertex_phi, s->vertex_phi, (size_t)s->n_el * s->L.nb_ip * 8)) != MGPU_OK) return rc;
    if ((rc = cp(t_ptr, s->t_ptr, ((size_t)s->n_el * s->L.nb_ip + 1) * 4)) != MGPU_OK) return rc;
    if ((rc = cp(t_adof, s->t_adof, (size_t)s->n_t * 4)) != MGPU_OK) return rc;
    if ((rc = cp(t_weight, s->t_weight, (size_t)s->n_t * 8)) != MGPU_OK) return rc;
    if ((rc = cp(topo, s->topo, (size_t)s->n_el * s->nf * s->na_t * 4)) != MGPU_OK) return rc;
    if ((rc = cp(row_ptr, s->row_ptr, ((size_t)s->n_rows + 1) * 4)) != MGPU_OK) return rc;
    if ((rc = cp(col_idx, s->col_idx, (size_t)s->nnz * 4)) != MGPU_OK) return rc;
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_setup_desc(mgpu_gpu_setup* s, mgpu_set_desc* d) {
    MGPU_GUARD_BEGIN
    if (!s || !d) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    const int dim = s->L.dim, nm = s->nf == 1 ? 1 : 2;
    s->h_elem_cell.resize((size_t)s->n_el);
    s->h_ig_ptr.resize((size_t)s->n_el + 1);
    s->h_ig_xi.resize((size_t)s->n_ig * (dim + 1) * dim);
    s->h_material.resize((size_t)s->n_el * nm);
    s->h_ip_box.resize((size_t)s->n_el * 2 * dim);
    s->h_vertex_phi.resize((size_t)s->n_el * s->L.nb_ip);
    s->h_t_ptr.resize((size_t)s->n_el * s->L.nb_ip + 1);
    s->h_t_adof.resize((size_t)s->n_t);
    s->h_t_weight.resize((size_t)s->n_t);
    mgpu_status rc = mgpu_setup_copy(s, s->h_elem_cell.data(), s->h_ig_ptr.data(), s->h_ig_xi.data(), s->h_material.data(),
                                     s->h_ip_box.data(), s->h_vertex_phi.data(), s->h_t_ptr.data(), s->h_t_adof.data(), s->h_t_weight.data(),
                                     nullptr, nullptr, nullptr);
    if (rc != MGPU_OK) return rc;
    std::memset(d, 0, sizeof(*d));
    d->spatial_dim = dim;
    d->iwg = s->L.iwg;
    d->ip_geometry = s->L.nb_ip == 27 ? MGPU_GEOM_HEX27 : (dim == 3 ? MGPU_GEOM_HEX8 : MGPU_GEOM_QUAD4);
    d->ig_geometry = dim == 3 ? MGPU_GEOM_TET4 : MGPU_GEOM_TRI3;
    d->n_elements = s->n_el;
    d->n_bases = s->L.nb_ip;
    d->n_adofs_per_type = s->N;
    d->ip_box = s->h_ip_box.data();
    d->material = s->h_material.data();
    d->t_ptr = s->h_t_ptr.data();
    d->t_adof = s->h_t_adof.data();
    d->t_weight = s->h_t_weight.data();
    d->ig_ptr = s->h_ig_ptr.data();
    d->ig_xi = s->h_ig_xi.data();
    for (int k = 0; k < 3; ++k) {
        d->lattice.n[k] = s->L.n[k];
        d->lattice.lo[k] = s->L.lo[k];
        d->lattice.h[k] = s->L.h[k];
    }
    d->elem_cell = s->h_elem_cell.data();
    d->vertex_phi = s->h_vertex_phi.data();
    return MGPU_OK;
    MGPU_GUARD_END
}

namespace {
template <class T>
mgpu_status set_upload(mgpu_set* S, const T* src, size_t n, T** dst) {
    *dst = nullptr;
    if (n == 0) return MGPU_OK;
    void* p = nullptr;
    SU_TRY(hipMalloc(&p, n * sizeof(T)));
    S->allocs.push_back(p);
    SU_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<T*>(p);
    return MGPU_OK;
}
}  // namespace

extern "C" mgpu_status mgpu_set_create_from_setup(mgpu_gpu_setup* G, mgpu_set** out) {
    MGPU_GUARD_BEGIN
    if (!G || !out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (G->L.order != 1 || G->L.nb_ip != G->L.nb)
        return set_error(MGPU_ERR_UNSUPPORTED, "a set made from the GPU set-up needs HEX8 / QUAD4 cells with nodal T "
                                               "rows (bspline_order 1); use mgpu_setup_desc and mgpu_set_create");
    const su::Lat& L = G->L;
    auto S = std::make_unique<mgpu_set>();
    SetHost& h = S->h;
    h.dim = L.dim;
    h.iwg = L.iwg;
    h.nf = G->nf;
    h.nb = L.nb;
    h.na = h.nf * h.nb;
    h.nm = h.nf == 1 ? 1 : 2;
    h.n_el = G->n_el;
    h.n_adofs_type = G->N;
    h.n_rows = (int32_t)G->n_rows;
    h.nnz = G->nnz;
    h.lattice = true;
    for (int k = 0; k < 3; ++k) {
        h.lat.n[k] = L.n[k];
        h.lat.lo[k] = L.lo[k];
        h.lat.h[k] = L.h[k];
    }
    h.n_cells = L.n_cells;
    h.n_nodes = L.n_nodes;
    const int npts = L.dim == 3 ? kTet11Points : kTri7Points;
    h.n_skip_subcells = G->n_skip;
    h.n_skip_points = G->n_skip * npts;
    // cut elements from the sub-cell offsets (the one host copy the set needs up front)
    std::vector<int32_t> igp((size_t)G->n_el + 1);
    SU_TRY(hipMemcpy(igp.data(), G->ig_ptr, igp.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    h.elem_cut.assign((size_t)G->n_el, -1);
    for (int32_t e = 0; e < G->n_el; ++e)
        if (igp[(size_t)e + 1] > igp[(size_t)e]) {
            h.elem_cut[(size_t)e] = (int32_t)h.cut_elems.size();
            h.cut_elems.push_back(e);
        }
    DevView& v = S->v;
    v.n_el = h.n_el;
    v.n_cut = (int32_t)h.cut_elems.size();
    v.N = h.n_adofs_type;
    v.n_rows = h.n_rows;
    v.nx = L.n[0];
    v.ny = L.n[1];
    v.nz = L.dim == 3 ? L.n[2] : 1;
    for (int d = 0; d < 3; ++d) v.h[d] = L.h[d];
    S->n_ip_cut = (int64_t)G->n_ig * npts;
    // the set-up's device arrays, used in place (the set-up outlives the set)
    v.material = G->material;
    v.cell_elem = G->cell_elem;
    v.node_adof = G->node_adof;
    v.row_node = G->basis_of;
    v.row_ptr = G->row_ptr;
    v.ig_ptr = G->ig_ptr;
    v.ig_xi = G->ig_xi;
    mgpu_status rc;
    if ((rc = set_upload(S.get(), h.elem_cut.data(), h.elem_cut.size(), const_cast<int32_t**>(&v.elem_cut))) != MGPU_OK)
        return rc;
    if ((rc = set_upload(S.get(), h.cut_elems.data(), h.cut_elems.size(), const_cast<int32_t**>(&v.cut_elems))) != MGPU_OK)
        return rc;
    if (L.dim == 3 && v.n_cut > 0) {
        std::vector<double> vphi((size_t)G->n_el * 8);
        SU_TRY(hipMemcpy(vphi.data(), G->vertex_phi, vphi.size() * sizeof(double), hipMemcpyDeviceToHost));
        S->vphi_cut.resize((size_t)v.n_cut * 8);
        for (int32_t c = 0; c < v.n_cut; ++c)
            for (int k = 0; k < 8; ++k) S->vphi_cut[(size_t)c * 8 + k] = vphi[(size_t)h.cut_elems[(size_t)c] * 8 + k];
    }
    if ((rc = finish_lattice_set(S.get(), igp.data())) != MGPU_OK) return rc;
    // host copies of the pattern, topology and cell map only when a host entry point asks
    S->fetch_host = [G](mgpu_set* T) -> mgpu_status {
        SetHost& hh = T->h;
        hh.row_ptr.resize((size_t)hh.n_rows + 1);
        hh.col_idx.resize((size_t)hh.nnz);
        hh.topo.resize((size_t)hh.n_el * hh.na);
        hh.cell_elem.resize((size_t)hh.n_cells);
        SU_TRY(hipMemcpy(hh.row_ptr.data(), G->row_ptr, hh.row_ptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (hh.nnz > 0)
            SU_TRY(hipMemcpy(hh.col_idx.data(), G->col_idx, hh.col_idx.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (!hh.topo.empty())
            SU_TRY(hipMemcpy(hh.topo.data(), G->topo, hh.topo.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        SU_TRY(hipMemcpy(hh.cell_elem.data(), G->cell_elem, hh.cell_elem.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        hh.topo_ptr.resize((size_t)hh.n_el + 1);
        for (int64_t e = 0; e <= hh.n_el; ++e) hh.topo_ptr[(size_t)e] = e * hh.na;
        // lattice vertex -> position of its adof among the element's adofs (ascending node index)
        static const uint8_t pos3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
        hh.vert_pos.resize((size_t)hh.n_el * hh.nb);
        for (int64_t e = 0; e < hh.n_el; ++e)
            for (int b = 0; b < hh.nb; ++b) hh.vert_pos[(size_t)e * hh.nb + b] = pos3[b];
        return MGPU_OK;
    };
    *out = S.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

