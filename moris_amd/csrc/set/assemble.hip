// assemble.hip -- HIP kernels of the FP64 element K+R assembly and the mgpu_set entry points that
// drive them.  Target: gfx950 (MI355X), 64-wide wavefronts.
//
// Path of one mgpu_set_assemble call on a lattice set:
//   k_cut_diffusion / k_cut_elasticity
//                   one wavefront per cut element.  For every IG sub-cell and every point of its
//                   rule (TET_11 / TRI_7): map the point to the IP cell (Space_Interpolator::
//                   map_integration_point, cl_MTK_Space_Interpolator.cpp:1381-1397), take detJ of
//                   the sub-cell (:1180-1210 tet, :1120-1128 tri), skip if detJ < -1e-6
//                   (cl_FEM_Element_Bulk.cpp:291), w* = w * detJ (:297), evaluate dN/dx of the IP
//                   cell (Field_Interpolator::eval_d1Ndx1, cl_FEM_Field_Interpolator.cpp:464-496,
//                   diagonal inverse Jacobian of a rectangular cell, cl_MTK_Space_Interpolator.cpp
//                   :676-700) and accumulate the moment matrix M = sum_ip w* g g^T, g = dN/dx.
//                   The element matrix K_p = B^T D B contracted with M (IWG compute_jacobian,
//                   cl_FEM_IWG_Isotropic_Struc_Linear_Bulk.cpp:200, cl_FEM_IWG_Diffusion_Bulk.cpp
//                   :280) is projected with the element's T rows into sorted adof order
//                   (cl_MSI_Equation_Object.cpp:810-841) and stored.
//   k_prep          u gathered per lattice node, one material record per lattice cell.
//   k_stencil       owner-computes scatter: one thread per CSR row.  The row sums the K_e rows of
//                   the (up to 8 / 4) lattice cells around its node -- uncut cells from the
//                   compile-time reference moments (HEX_2x2x2 / QUAD_2x2 on the IP cell) and the
//                   cell record, cut cells from the stored K_e -- and writes every value of the
//                   row exactly once; the row
//                   of R is the same row applied to u (R = sum_e K_e u_e, linear IWGs).  Replaces
//                   Epetra_FECrsMatrix::SumIntoGlobalValues (cl_Sparse_Matrix_EpetraFECrs.cpp
//                   :103-133); deterministic, no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../common/fem_tables.hpp"
#include "../setup/internal.hpp"
#include "set_state.hpp"
#include "../learned/quad.hpp"

namespace mgpu {

// ------------------------------------------------------------------------------------------------
// k_cut_diffusion: one wavefront per cut element.  Each lane integrates every 64th point and keeps
// the symmetric moments S[a][b] = sum_ip w* g_a . g_b in registers; one LDS reduction per element.
// ------------------------------------------------------------------------------------------------
template <int DIM, typename T>
__global__ __launch_bounds__(64) void k_cut_diffusion(DevView s) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NV = DIM + 1;
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    constexpr int NSYM = NB * (NB + 1) / 2;
    __shared__ double red[NSYM * 65];

    const int c = s.cut_order ? s.cut_order[blockIdx.x] : (int)blockIdx.x;
    if (s.skip && s.skip[c]) return;
    const int lane = threadIdx.x;
    const int e = s.cut_elems[c];
    const int ig0 = s.ig_ptr[e];
    const int nip = (s.ig_ptr[e + 1] - ig0) * NQ;
    double hh[DIM];
    for (int d = 0; d < DIM; ++d) hh[d] = s.h[d];
    (void)NV;

    // T = float: FP32 integrand with FP32 accumulation; the point map and the sub-cell Jacobian
    // (cut_point) stay FP64
    T acc[NSYM];
#pragma unroll
    for (int t = 0; t < NSYM; ++t) acc[t] = (T)0;
    for (int ip = lane; ip < nip; ip += 64) {
        double gd[DIM][NB];
        const T w = (T)cut_point<DIM>(s, ig0, ip, hh, gd);
        T g[DIM][NB];
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int a = 0; a < NB; ++a) g[i][a] = (T)gd[i][a];
        int t = 0;
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int b = a; b < NB; ++b) {
                T v = (T)0;
#pragma unroll
                for (int i = 0; i < DIM; ++i) v += g[i][a] * g[i][b];
                acc[t++] += w * v;
            }
    }
#pragma unroll
    for (int t = 0; t < NSYM; ++t) red[t * 65 + lane] = (double)acc[t];
    __syncthreads();
    T* ke = ke_store<T>(s) + (size_t)c * NB * NB;
    const double k = s.material[e];
    if (lane < NSYM) {
        double v = 0.0;
        for (int l = 0; l < 64; ++l) v += red[lane * 65 + l];
        // (a, b) of the packed upper triangle
        int a = 0, t = lane;
        while (t >= NB - a) { t -= NB - a; ++a; }
        const int b = a + t;
        ke[a * NB + b] = (T)(k * v);
        ke[b * NB + a] = (T)(k * v);
    }
}

// ------------------------------------------------------------------------------------------------
// k_cut_elasticity: one wavefront per cut element.  Points are evaluated 64 at a time into LDS; lane
// (a, b) accumulates the 3x3 (2x2) direction block M[(i,a),(j,b)] = sum w* g_ia g_jb over all points,
// then contracts it with D (B^T D B) and projects to sorted adof order.
// ------------------------------------------------------------------------------------------------
template <int DIM, typename T>
__global__ __launch_bounds__(64) void k_cut_elasticity(DevView s) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr int NA = DIM * NB;
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    constexpr int GPAD = NG + 1;
    __shared__ T sgw[64 * GPAD];
    __shared__ T sg[64 * GPAD];

    const int c = s.cut_order ? s.cut_order[blockIdx.x] : (int)blockIdx.x;
    if (s.skip && s.skip[c]) return;
    const int lane = threadIdx.x;
    const int e = s.cut_elems[c];
    const int ig0 = s.ig_ptr[e];
    const int nip = (s.ig_ptr[e + 1] - ig0) * NQ;
    double hh[DIM];
    for (int d = 0; d < DIM; ++d) hh[d] = s.h[d];
    const int a = (lane / NB) % NB, b = lane % NB;
    const bool active = lane < NB * NB;

    T M[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) M[i][j] = (T)0;

    for (int base = 0; base < nip; base += 64) {
        const int ip = base + lane;
        if (ip < nip) {
            double g[DIM][NB];
            const double w = cut_point<DIM>(s, ig0, ip, hh, g);
            for (int i = 0; i < DIM; ++i)
                for (int q = 0; q < NB; ++q) {
                    sg[lane * GPAD + i * NB + q] = (T)g[i][q];
                    sgw[lane * GPAD + i * NB + q] = (T)(w * g[i][q]);
                }
        } else {
            for (int k = 0; k < NG; ++k) {
                sg[lane * GPAD + k] = (T)0;
                sgw[lane * GPAD + k] = (T)0;
            }
        }
        __syncthreads();
        const int cnt = min(64, nip - base);
        if (active) {
            for (int p = 0; p < cnt; ++p) {
                T ga[DIM], gb[DIM];
#pragma unroll
                for (int i = 0; i < DIM; ++i) {
                    ga[i] = sgw[p * GPAD + i * NB + a];
                    gb[i] = sg[p * GPAD + i * NB + b];
                }
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) M[i][j] += ga[i] * gb[j];
            }
        }
        __syncthreads();
    }
    if (!active) return;
    constexpr int NS = DIM == 3 ? 6 : 3;
    double D[NS][NS];
    elastic_D<DIM>(s.material[2 * e], s.material[2 * e + 1], D);
    T* ke = ke_store<T>(s) + (size_t)c * NA * NA;
#pragma unroll
    for (int al = 0; al < DIM; ++al)
#pragma unroll
        for (int be = 0; be < DIM; ++be) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j)
                    if (d_nonzero<DIM>(voigt<DIM>(al, i), voigt<DIM>(be, j)))
                        v += D[voigt<DIM>(al, i)][voigt<DIM>(be, j)] * (double)M[i][j];
            ke[(a * DIM + al) * NA + (be * NB + b)] = (T)v;
        }
}

// ------------------------------------------------------------------------------------------------
// k_cut_elasticity_mfma: 3D elasticity, one wavefront per cut element, the moment matrix on MFMA.
//   M (24 x 24, component c = 8 i + a, i the direction, a the vertex) = sum_p w*_p g_p g_p^T is
//   symmetric; padded to 32 x 32 it is three 16 x 16 tiles (upper left, upper right, lower right).
//   Points are mapped 64 at a time into FP64 records (xi, w*) in LDS (one lane per point: IG -> IP
//   map and sub-cell detJ).  Each MFMA step takes 4 points: lane l handles point l / 16 of the step
//   and component column l % 16 = col, and forms in registers, from the point's record, the two
//   gradient entries it feeds -- g[col] (direction col / 8, vertex col % 8) and g[16 + col]
//   (direction 2, vertex col; zero padding for col >= 8) -- and w* times them: those are its A
//   (row col, k = point) and B (k = point, column col) operands of v_mfma_f64_16x16x4f64
//   (v_mfma_f32_16x16x4f32 with FP32 operands and FP32 accumulation on the mixed path; the records
//   are formed in FP64 there too).  The tiles go to LDS once, and the contraction with D and the
//   store of K_e follow as in k_cut_elasticity.
// ------------------------------------------------------------------------------------------------
typedef double mgpu_doublex4_t __attribute__((ext_vector_type(4)));
typedef float mgpu_floatx4_t __attribute__((ext_vector_type(4)));

template <typename T>
__device__ inline void mfma_16x16x4(T a, T b, typename std::conditional<sizeof(T) == 8, mgpu_doublex4_t,
                                                                        mgpu_floatx4_t>::type& acc) {
    if constexpr (sizeof(T) == 8) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    else acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

template <typename T>
__global__ __launch_bounds__(64) void k_cut_elasticity_mfma(DevView s) {
    constexpr int NB = 8, NG = 24, NA = 24;
    constexpr int MP = NG + 1;
    using acc_t = typename std::conditional<sizeof(T) == 8, mgpu_doublex4_t, mgpu_floatx4_t>::type;
    __shared__ double4 rec[64];
    __shared__ T Ms[NG * MP];

    const int c = s.cut_order ? s.cut_order[blockIdx.x] : (int)blockIdx.x;
    if (s.skip && s.skip[c]) return;
    const int lane = threadIdx.x;
    const int e = s.cut_elems[c];
    const int ig0 = s.ig_ptr[e];
    const int nip = (s.ig_ptr[e + 1] - ig0) * kTet11Points;
    double hh[3];
    for (int d = 0; d < 3; ++d) hh[d] = s.h[d];

    // the lane's component column and the constant factors of its two gradient entries
    const int col = lane & 15, grp = lane >> 4;
    const int a = col & 7;
    const double sx = kHex8XiHat[a][0], sy = kHex8XiHat[a][1], sz = kHex8XiHat[a][2];
    const bool lo_x = col < 8;                 // g[col]: d/dx (col < 8) or d/dy (col >= 8)
    const double c_lo = lo_x ? 0.25 * sx / hh[0] : 0.25 * sy / hh[1];
    const double c_hi = lo_x ? 0.25 * sz / hh[2] : 0.0;   // g[16 + col]: d/dz, or padding

    acc_t m00 = {0, 0, 0, 0}, m01 = {0, 0, 0, 0}, m11 = {0, 0, 0, 0};
    for (int base = 0; base < nip; base += 64) {
        {
            const int ip = base + lane;
            double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
            if (ip < nip) {
                double xi[3];
                r.w = cut_point_map<3>(s, ig0, ip, hh, xi);
                r.x = xi[0];
                r.y = xi[1];
                r.z = xi[2];
            }
            rec[lane] = r;
        }
        __syncthreads();
        const int steps = (min(64, nip - base) + 3) >> 2;
        for (int st = 0; st < steps; ++st) {
            const double4 r = rec[4 * st + grp];
            const double fx = 1.0 + sx * r.x, fy = 1.0 + sy * r.y, fz = 1.0 + sz * r.z;
            const double lo = c_lo * (lo_x ? fy : fx) * fz;
            const double hi = c_hi * fx * fy;
            const T blo = (T)lo, bhi = (T)hi;
            const T alo = (T)(r.w * lo), ahi = (T)(r.w * hi);
            mfma_16x16x4<T>(alo, blo, m00);
            mfma_16x16x4<T>(alo, bhi, m01);
            mfma_16x16x4<T>(ahi, bhi, m11);
        }
        __syncthreads();
    }
    // tiles -> M in LDS (result register j of a lane: row grp + 4 j for FP64, 4 grp + j for FP32;
    // column col), the lower left tile mirrored from the upper right one
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = sizeof(T) == 8 ? grp + 4 * j : 4 * grp + j;
        Ms[row * MP + col] = m00[j];
        if (col < 8) {
            Ms[row * MP + 16 + col] = m01[j];
            Ms[(16 + col) * MP + row] = m01[j];
            if (row < 8) Ms[(16 + row) * MP + 16 + col] = m11[j];
        }
    }
    __syncthreads();
    // B^T D B: lane (a, b) owns the 3 x 3 direction block of the vertex pair (a, b)
    const int va = lane >> 3, vb = lane & 7;
    double D[6][6];
    elastic_D<3>(s.material[2 * e], s.material[2 * e + 1], D);
    T* ke = ke_store<T>(s) + (size_t)c * NA * NA;
#pragma unroll
    for (int al = 0; al < 3; ++al)
#pragma unroll
        for (int be = 0; be < 3; ++be) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (d_nonzero<3>(voigt<3>(al, i), voigt<3>(be, j)))
                        v += D[voigt<3>(al, i)][voigt<3>(be, j)] * (double)Ms[(i * NB + va) * MP + j * NB + vb];
            ke[(va * 3 + al) * NA + (be * NB + vb)] = (T)v;
        }
}

// ------------------------------------------------------------------------------------------------
// k_prep: per-call inputs of the stencil kernel, one thread per lattice node / cell.  Each node
// gathers u of its adofs into the planar u_node (0 where the node has no adof); each cell writes
// one record (void / uncut material / cut index) with the material folded in: k
// (CM_Diffusion_Linear_Isotropic) or the three distinct entries of the isotropic D
// (CM_Struc_Linear_Isotropic::full_3d / full_plane_stress).  The reference moments of an uncut
// cell are compile-time constants (kRefGrad), scaled by the stencil itself.
// ------------------------------------------------------------------------------------------------
template <int DIM, int NF>
__global__ __launch_bounds__(256) void k_prep(DevView s, const double* u) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr bool DIFF = NF == 1;
    constexpr int NM = DIFF ? NB * NB : NG * NG;
    constexpr int NQ = DIM == 3 ? kHex2Points : kQuad2Points;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n_cells = (int64_t)s.nx * s.ny * s.nz;
    const int64_t n_nodes = (int64_t)(s.nx + 1) * (s.ny + 1) * (DIM == 3 ? s.nz + 1 : 1);
    if (c < n_nodes) {
        // u gathered per lattice node so that rows read it without the node -> adof hop
        const int a = s.node_adof[c];
#pragma unroll
        for (int f = 0; f < NF; ++f) s.u_node[(size_t)f * n_nodes + c] = a >= 0 ? u[(size_t)f * s.N + a] : 0.0;
    }
    if (c >= n_cells) return;
    const int e = s.cell_elem[c];
    if (DIFF) {
        double v = 0.0;
        if (e >= 0) {
            const int cut = s.elem_cut[e];
            v = cut >= 0 ? -(double)(cut + 1) : s.material[e];
        }
        s.cell_rec[c] = v;
    } else {
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        if (e >= 0) {
            const int cut = s.elem_cut[e];
            if (cut >= 0) {
                r[0] = -(double)(cut + 1);
            } else {
                constexpr int NS = DIM == 3 ? 6 : 3;
                double D[NS][NS];
                elastic_D<DIM>(s.material[2 * e], s.material[2 * e + 1], D);
                r[0] = 1.0;
                r[1] = D[0][0];
                r[2] = D[0][1];
                r[3] = D[NS - 1][NS - 1];
            }
        }
        for (int k = 0; k < 4; ++k) s.cell_rec[4 * c + k] = r[k];
    }
}

// ------------------------------------------------------------------------------------------------
// k_stencil: one thread per (CSR row, column dof type BE); the dof type AL of the row is uniform
// per block.  The thread sums the (AL, BE) block of the K_e rows of the lattice cells around its
// node: uncut cells from the reference moments and the cell record, cut cells from the stored K_e.
// ------------------------------------------------------------------------------------------------
// vertex order -> sorted position for lattice cells (adof ids ascend with lattice node index)
template <int DIM>
__device__ constexpr int lattice_sorted_pos(int b) {
    return DIM == 3 ? (b == 2 ? 3 : (b == 3 ? 2 : (b == 6 ? 7 : (b == 7 ? 6 : b)))) : (b == 2 ? 3 : (b == 3 ? 2 : b));
}

// Stencil variants (the autotuner's choices): rows per block (a block has ROWS * NF threads,
// <= 256) and whether the block stages its contiguous CSR range in LDS for coalesced writes or
// each thread stores its row directly.  Variants 0 .. n_rows_choices - 1 stage, the rest store
// directly with the same row counts; variant 1 is the default.
template <int NF>
__host__ __device__ constexpr int stencil_row_choices() {
    return NF == 1 ? 3 : 2;
}
template <int NF>
__host__ __device__ constexpr int stencil_variants() {
    return 2 * stencil_row_choices<NF>();
}
template <int NF>
__host__ __device__ constexpr int stencil_rows(int variant) {
    const int v = variant % stencil_row_choices<NF>();
    return NF == 3 ? (v == 0 ? 32 : 64) : (v == 0 ? 64 : (v == 1 ? 128 : 256));
}
template <int NF>
__host__ __device__ constexpr bool stencil_direct(int variant) {
    return variant >= stencil_row_choices<NF>();
}

// The (AL, BE) block of row r: its values in acc[q] for the stencil neighbours q set in mask, in
// ascending q from CSR index idx0; its part of R in *r_part.  False (and the error flag set) if the
// row's pattern does not match the stencil.
template <int DIM, int NF, int AL, int BE, typename KT>
__device__ inline bool stencil_row(const DevView& s, const double (*sc)[3], int r, double* acc, uint32_t& mask_out,
                                   int& idx0, double* r_part) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr int NA = NF * NB;
    constexpr int NS27 = DIM == 3 ? 27 : 9;
    constexpr bool DIFF = NF == 1;
    const int nx = s.nx, ny = s.ny, nz = s.nz;
    const int node = s.row_node[r];
    const int i = node % (nx + 1);
    const int j = (node / (nx + 1)) % (ny + 1);
    const int k = DIM == 3 ? node / ((nx + 1) * (ny + 1)) : 0;

    constexpr int NC = DIM == 3 ? 8 : 4;
    // the row's CSR range and the records of its (up to NC) cells, all requested before any is
    // used: out-of-lattice cells read cell 0 and are masked afterwards
    const int row = AL * s.N + r;
    const int start = s.row_ptr[row];
    const int next = s.row_ptr[row + 1];
    bool inside[NC];
    double recf[NC];
    double4 rec4[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        const int ci = i + ((cc & 1) ? 0 : -1);
        const int cj = j + ((cc & 2) ? 0 : -1);
        const int ck = DIM == 3 ? k + ((cc & 4) ? 0 : -1) : 0;
        inside[cc] = !(ci < 0 || ci >= nx || cj < 0 || cj >= ny || (DIM == 3 && (ck < 0 || ck >= nz)));
        const int cell = inside[cc] ? ci + nx * (cj + ny * ck) : 0;
        if (DIFF) recf[cc] = s.cell_rec[cell];
        else rec4[cc] = *reinterpret_cast<const double4*>(s.cell_rec + 4 * (size_t)cell);
    }

    // u at the 27 neighbours.  Where the registers allow (FP32 K_e, or diffusion) they are
    // requested together with the cell records, the address not waiting for the stencil mask (a
    // neighbour beyond the lattice is clamped into the node array and never used); with FP64 K_e
    // elasticity that would cost a wavefront per SIMD, so there they follow the mask (a neighbour
    // outside the stencil reads the row's own node).  The mask selects the use.
    // (node indices are int32: the lattice has fewer than 2^31 nodes)
    constexpr bool kEarlyU = sizeof(KT) == 4 || DIFF;
    const int n_nodes = (nx + 1) * (ny + 1) * (DIM == 3 ? nz + 1 : 1);
    const double* ub = s.u_node + (size_t)BE * n_nodes;
    double un[NS27];
    auto load_u = [&](uint32_t m) {
#pragma unroll
        for (int q = 0; q < NS27; ++q) {
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = DIM == 3 ? q / 9 - 1 : 0;
            const int nb = node + dx + (nx + 1) * (dy + (ny + 1) * dz);
            const int at = kEarlyU ? min(max(nb, 0), n_nodes - 1) : ((m & (1u << q)) ? nb : node);
            un[q] = ub[at];
        }
    };
    if constexpr (kEarlyU) load_u(0u);

    // stencil mask: the neighbours of the row node in its material cells
    uint32_t mask = 0;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        const int cx = (cc & 1) ? 0 : -1;
        const int cy = (cc & 2) ? 0 : -1;
        const int cz = DIM == 3 ? ((cc & 4) ? 0 : -1) : 0;
        const double flag = !inside[cc] ? 0.0 : (DIFF ? recf[cc] : rec4[cc].x);
        if (flag == 0.0) continue;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
            const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
            mask |= 1u << q;
        }
    }
    if constexpr (!kEarlyU) load_u(mask);

#pragma unroll
    for (int q = 0; q < NS27; ++q) acc[q] = 0.0;

    // the (AL, a_loc; BE, .) segment of a cut cell's stored K_e: NB contiguous values in vertex
    // order, 16-byte aligned, read as 16-byte vectors
    auto ke_seg = [&](int cc, double flag, double* kv) {
        constexpr int a_tab[8] = {6, 7, 5, 4, 2, 3, 1, 0};
        constexpr int a_tab2[4] = {2, 3, 1, 0};
        const int a_loc = DIM == 3 ? a_tab[cc] : a_tab2[cc];
        const int cut = (int)(-flag) - 1;
        const KT* base = ke_store<KT>(s) + (size_t)cut * NA * NA + (size_t)(a_loc * NF + AL) * NA + BE * NB;
        if constexpr (sizeof(KT) == 8) {
            const double2* p2 = reinterpret_cast<const double2*>(base);
#pragma unroll
            for (int h2 = 0; h2 < NB / 2; ++h2) {
                const double2 t = p2[h2];
                kv[2 * h2] = t.x;
                kv[2 * h2 + 1] = t.y;
            }
        } else {
            const float4* p4 = reinterpret_cast<const float4*>(base);
#pragma unroll
            for (int h4 = 0; h4 < NB / 4; ++h4) {
                const float4 t = p4[h4];
                kv[4 * h4] = t.x;
                kv[4 * h4 + 1] = t.y;
                kv[4 * h4 + 2] = t.z;
                kv[4 * h4 + 3] = t.w;
            }
        }
    };

#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        const int cx = (cc & 1) ? 0 : -1;
        const int cy = (cc & 2) ? 0 : -1;
        const int cz = DIM == 3 ? ((cc & 4) ? 0 : -1) : 0;
        // local vertex of the row node in this cell
        constexpr int a_tab[8] = {6, 7, 5, 4, 2, 3, 1, 0};   // vertex with offset (-cx,-cy,-cz)
        constexpr int a_tab2[4] = {2, 3, 1, 0};
        const int a_loc = DIM == 3 ? a_tab[cc] : a_tab2[cc];
        const double flag = !inside[cc] ? 0.0 : (DIFF ? recf[cc] : rec4[cc].x);
        if (flag == 0.0) continue;
        double c0 = 0.0, c1 = 0.0, cs = 0.0;
        if (!DIFF) {
            c0 = rec4[cc].y;
            c1 = rec4[cc].z;
            cs = rec4[cc].w;
        }
        if (flag < 0.0) {
            double kv[NB];
            ke_seg(cc, flag, kv);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                acc[q] += kv[b];
            }
        } else if (DIFF) {
            // k * sum_i M[(i,a),(i,b)] from the reference-cell integrals (compile-time constants)
            double ks[DIM];
#pragma unroll
            for (int ii = 0; ii < DIM; ++ii) ks[ii] = flag * sc[ii][ii];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                double v = 0.0;
#pragma unroll
                for (int ii = 0; ii < DIM; ++ii) v += ks[ii] * kRefGrad<DIM>.v[ii][a_loc][ii][b];
                acc[q] += v;
            }
        } else {
            // isotropic B^T D B block (AL, BE):
            //   D[AL][BE] M[(AL,a),(BE,b)] + D_s (delta_{AL,BE} sum_{i!=AL} M[(i,a),(i,b)]
            //                                      + (1 - delta_{AL,BE}) M[(BE,a),(AL,b)])
            // with M = G * scale from the reference-cell integrals (compile-time constants)
            const double cab = (AL == BE ? c0 : c1) * sc[AL][BE];
            double css[DIM];
#pragma unroll
            for (int ii = 0; ii < DIM; ++ii) css[ii] = cs * (AL == BE ? sc[ii][ii] : sc[BE][AL]);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                double v = cab * kRefGrad<DIM>.v[AL][a_loc][BE][b];
                if (AL == BE) {
#pragma unroll
                    for (int ii = 0; ii < DIM; ++ii)
                        if (ii != AL) v += css[ii] * kRefGrad<DIM>.v[ii][a_loc][ii][b];
                } else {
                    v += css[0] * kRefGrad<DIM>.v[BE][a_loc][AL][b];
                }
                acc[q] += v;
            }
        }
    }

    const int cnt = __popc(mask);
    if (next - start != NF * cnt) {
        atomicOr(s.err_flag, 1);
        *r_part = 0.0;
        return false;
    }
    double R = 0.0;
#pragma unroll
    for (int q = 0; q < NS27; ++q)
        if (mask & (1u << q)) R += acc[q] * un[q];
    *r_part = R;
    mask_out = mask;
    idx0 = start + BE * cnt;
    return true;
}

template <int DIM, int NF, int AL, typename KT>
__device__ inline bool stencil_dispatch_be(const DevView& s, const double (*sc)[3], int r, double* acc,
                                           uint32_t& mask, int& idx0, double* rp, int be) {
    switch (be) {
        case 0: return stencil_row<DIM, NF, AL, 0, KT>(s, sc, r, acc, mask, idx0, rp);
        case 1: if constexpr (NF > 1) return stencil_row<DIM, NF, AL, 1, KT>(s, sc, r, acc, mask, idx0, rp); break;
        case 2: if constexpr (NF > 2) return stencil_row<DIM, NF, AL, 2, KT>(s, sc, r, acc, mask, idx0, rp); break;
        default: break;
    }
    return false;
}

// Rows per staging round of a staged stencil block: the block's CSR range is written in rounds of
// this many rows, each a contiguous range, so the LDS copy holds one round only.
template <int NF>
__host__ __device__ constexpr int stencil_round_rows(int rows) {
    return rows < (NF == 3 ? 32 : 64) ? rows : (NF == 3 ? 32 : 64);
}

template <int DIM, int NF, int ROWS, bool DIRECT, typename KT>
__global__ __launch_bounds__(256) void k_stencil(DevView s, const double* u, double* vals, double* rhs) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr bool DIFF = NF == 1;
    constexpr int NS27 = DIM == 3 ? 27 : 9;
    constexpr int RR = stencil_round_rows<NF>(ROWS);
    static_assert(ROWS % RR == 0, "whole staging rounds");
    __shared__ double stage[DIRECT ? 1 : RR * NF * NS27];
    __shared__ double rpart[NF * ROWS];
    // scale of the reference-cell integrals: M[(i,a),(j,b)] = G[i][a][j][b] * sc[i][j]
    double sc[3][3];
    {
        double vol = 1.0;
        for (int d = 0; d < DIM; ++d) vol *= 0.5 * s.h[d];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) sc[i][j] = i < DIM && j < DIM ? 4.0 / (s.h[i] * s.h[j]) * vol : 0.0;
    }
    // block -> (row block, row field AL): the NF blocks of one row block are 8 IDs apart, so that
    // the dispatcher's round robin over the 8 XCDs puts them on one XCD at about the same time,
    // sharing in its L2 the cell records, u and the node-major rows of the stored cut K_e
    const int n_rb = (s.N + ROWS - 1) / ROWS;
    const int g8 = blockIdx.x >> 3;
    const int al = g8 % NF;
    const int rb = (g8 / NF) * 8 + (blockIdx.x & 7);
    if (rb >= n_rb) return;   // padding of the last group of 8 (the whole block, before any barrier)
    const int r0 = rb * ROWS;
    const int r1 = min(r0 + ROWS, s.N);
    const int be = threadIdx.x / ROWS;
    const int rr = threadIdx.x % ROWS;
    const int r = r0 + rr;
    double acc[NS27];
    uint32_t mask = 0;
    int idx0 = 0;
    bool ok = false;
    if (r < r1) {
        double* rp = rpart + be * ROWS + rr;
        switch (al) {
            case 0: ok = stencil_dispatch_be<DIM, NF, 0, KT>(s, sc, r, acc, mask, idx0, rp, be); break;
            case 1: if constexpr (NF > 1) ok = stencil_dispatch_be<DIM, NF, 1, KT>(s, sc, r, acc, mask, idx0, rp, be); break;
            case 2: if constexpr (NF > 2) ok = stencil_dispatch_be<DIM, NF, 2, KT>(s, sc, r, acc, mask, idx0, rp, be); break;
            default: break;
        }
    }
    if constexpr (DIRECT) {
        // each thread stores its row segment
        if (ok) {
            int idx = idx0;
#pragma unroll
            for (int q = 0; q < NS27; ++q)
                if (mask & (1u << q)) vals[idx++] = acc[q];
        }
    }
    __syncthreads();
    if (be == 0 && r < r1) {
        double R = 0.0;
#pragma unroll
        for (int f = 0; f < NF; ++f) R += rpart[f * ROWS + rr];
        rhs[al * s.N + r] = R;
    }
    if constexpr (!DIRECT) {
        // rounds of RR rows: their threads put their values in the block's LDS copy, then the
        // round's contiguous CSR range goes out in 16-byte stores (one value peeled at either end
        // where the range is not 16-byte aligned).  Every bound here is uniform over the block.
        for (int h = 0; h * RR < r1 - r0; ++h) {
            const int ra = r0 + h * RR, rb = min(ra + RR, r1);
            const int base = s.row_ptr[al * s.N + ra];
            const int n = s.row_ptr[al * s.N + rb] - base;
            if (ok && rr >= h * RR && rr < (h + 1) * RR && n <= RR * NF * NS27) {
                int idx = idx0 - base;
#pragma unroll
                for (int q = 0; q < NS27; ++q)
                    if (mask & (1u << q)) stage[idx++] = acc[q];
            }
            __syncthreads();
            if (n <= RR * NF * NS27) {   // a longer range is flagged by the row-length check
                double* dst = vals + base;
                const int head = (reinterpret_cast<uintptr_t>(dst) & 15) ? 1 : 0;
                const int n2 = (n - head) / 2;
                // nontemporal: the values (an output nothing reads back here) stream out without
                // displacing from L2 and the memory-side cache the cell records, u and the cut
                // K_e that the next blocks read (the cut K_e, written by the kernel before, is
                // meant to be found there, so its own stores stay ordinary)
                if (threadIdx.x == 0) {
                    if (head && n > 0) __builtin_nontemporal_store(stage[0], dst);
                    if (((n - head) & 1) && n > head) __builtin_nontemporal_store(stage[n - 1], dst + n - 1);
                }
                typedef double pair_t __attribute__((ext_vector_type(2)));
                pair_t* dst2 = reinterpret_cast<pair_t*>(dst + head);
                for (int p = threadIdx.x; p < n2; p += ROWS * NF)
                    __builtin_nontemporal_store(pair_t{stage[head + 2 * p], stage[head + 2 * p + 1]}, dst2 + p);
            }
            __syncthreads();
        }
    }
}

}  // namespace mgpu

// ================================================================================================
// host side of the set
// ================================================================================================
using namespace mgpu;

namespace {

thread_local std::string g_last_error;

template <class T>
mgpu_status upload(mgpu_set* S, const T* src, size_t n, T** dst) {
    *dst = nullptr;
    if (n == 0) return MGPU_OK;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    S->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<T*>(p);
    return MGPU_OK;
}

void launch_subdivision(mgpu_set* S, const int32_t* skip) {
    const int dim = S->h.dim, nf = S->h.nf;
    DevView v = S->v;
    v.skip = skip;
    dim3 g((unsigned)v.n_cut), b(64);
    const bool f32 = S->precision == MGPU_PREC_MIXED;
    if (dim == 3 && nf == 1) {
        if (f32) hipLaunchKernelGGL((k_cut_diffusion<3, float>), g, b, 0, S->stream, v);
        else hipLaunchKernelGGL((k_cut_diffusion<3, double>), g, b, 0, S->stream, v);
    } else if (dim == 3 && nf == 3) {
        if (f32) hipLaunchKernelGGL((k_cut_elasticity_mfma<float>), g, b, 0, S->stream, v);
        else hipLaunchKernelGGL((k_cut_elasticity_mfma<double>), g, b, 0, S->stream, v);
    } else if (dim == 2 && nf == 1) {
        if (f32) hipLaunchKernelGGL((k_cut_diffusion<2, float>), g, b, 0, S->stream, v);
        else hipLaunchKernelGGL((k_cut_diffusion<2, double>), g, b, 0, S->stream, v);
    } else {
        if (f32) hipLaunchKernelGGL((k_cut_elasticity<2, float>), g, b, 0, S->stream, v);
        else hipLaunchKernelGGL((k_cut_elasticity<2, double>), g, b, 0, S->stream, v);
    }
}

template <int DIM, int NF, int V>
void launch_stencil_variant(mgpu_set* S, const double* u, double* vals, double* rhs) {
    constexpr int ROWS = stencil_rows<NF>(V);
    const int n_rb = (S->v.N + ROWS - 1) / ROWS;
    const dim3 grid((unsigned)(((n_rb + 7) / 8) * 8 * NF));
    if (S->precision == MGPU_PREC_MIXED)
        hipLaunchKernelGGL((k_stencil<DIM, NF, ROWS, stencil_direct<NF>(V), float>), grid, dim3(ROWS * NF), 0, S->stream,
                           S->v, u, vals, rhs);
    else
        hipLaunchKernelGGL((k_stencil<DIM, NF, ROWS, stencil_direct<NF>(V), double>), grid, dim3(ROWS * NF), 0,
                           S->stream, S->v, u, vals, rhs);
}

template <int DIM, int NF>
void launch_stencil(mgpu_set* S, int variant, const double* u, double* vals, double* rhs) {
    switch (variant) {
        case 0: launch_stencil_variant<DIM, NF, 0>(S, u, vals, rhs); break;
        case 2: launch_stencil_variant<DIM, NF, 2>(S, u, vals, rhs); break;
        case 3: launch_stencil_variant<DIM, NF, 3>(S, u, vals, rhs); break;
        case 4: if constexpr (stencil_variants<NF>() > 4) { launch_stencil_variant<DIM, NF, 4>(S, u, vals, rhs); break; }
                [[fallthrough]];
        case 5: if constexpr (stencil_variants<NF>() > 5) { launch_stencil_variant<DIM, NF, 5>(S, u, vals, rhs); break; }
                [[fallthrough]];
        default: launch_stencil_variant<DIM, NF, 1>(S, u, vals, rhs); break;
    }
}

}  // namespace

namespace mgpu {
mgpu_status launch_assemble(mgpu_set* S, const double* u, double* vals, double* rhs, hipEvent_t* ev) {
    const SetHost& h = S->h;
    if (h.n_rows == 0) return MGPU_OK;
    if (h.generic) return launch_generic(S, u, vals, rhs, ev);
    const int dim = h.dim, nf = h.nf;
    if (ev) HIP_TRY(hipEventRecord(ev[0], S->stream));
    const bool learned = S->v.n_cut > 0 && S->quadrature == MGPU_QUAD_LEARNED;
    const bool mfit = S->v.n_cut > 0 && S->quadrature == MGPU_QUAD_MOMENT_FIT;
    if (learned) {
        mlp_forward_phi_launch(S->mlp, S->lv.vphi, S->lv.n_cut, S->lv.w_pred, S->stream);
        if (ev) HIP_TRY(hipEventRecord(ev[1], S->stream));
        lq_launch_fit(S->v, S->lv, S->stream);
        HIP_TRY(hipGetLastError());
    } else if (mfit) {
        if (ev) HIP_TRY(hipEventRecord(ev[1], S->stream));
        lq_launch_mfit8(S->v, S->lv, S->stream);
        HIP_TRY(hipGetLastError());
    } else if (ev) {
        HIP_TRY(hipEventRecord(ev[1], S->stream));
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2], S->stream));
    if (S->v.n_cut > 0) {
        if (learned) {
            lq_launch_element(S->v, S->lv, nf, S->precision == MGPU_PREC_MIXED, S->stream);
            launch_subdivision(S, S->lv.flag);
        } else if (mfit) {
            // every element takes the fitted rule: no subdivision launch
            lq_launch_element(S->v, S->lv, nf, S->precision == MGPU_PREC_MIXED, S->stream);
        } else {
            launch_subdivision(S, nullptr);
        }
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord(ev[3], S->stream));
    {
        const int64_t n_nodes = (int64_t)(S->v.nx + 1) * (S->v.ny + 1) * (dim == 3 ? S->v.nz + 1 : 1);
        dim3 gp((unsigned)((n_nodes + 255) / 256)), bp(256);
        if (dim == 3 && nf == 1) hipLaunchKernelGGL((k_prep<3, 1>), gp, bp, 0, S->stream, S->v, u);
        else if (dim == 3 && nf == 3) hipLaunchKernelGGL((k_prep<3, 3>), gp, bp, 0, S->stream, S->v, u);
        else if (dim == 2 && nf == 1) hipLaunchKernelGGL((k_prep<2, 1>), gp, bp, 0, S->stream, S->v, u);
        else hipLaunchKernelGGL((k_prep<2, 2>), gp, bp, 0, S->stream, S->v, u);
        HIP_TRY(hipGetLastError());
        if (ev) HIP_TRY(hipEventRecord(ev[4], S->stream));
        const int var = S->stencil_variant < 0 ? 1 : S->stencil_variant;
        if (dim == 3 && nf == 1) launch_stencil<3, 1>(S, var, u, vals, rhs);
        else if (dim == 3 && nf == 3) launch_stencil<3, 3>(S, var, u, vals, rhs);
        else if (dim == 2 && nf == 1) launch_stencil<2, 1>(S, var, u, vals, rhs);
        else launch_stencil<2, 2>(S, var, u, vals, rhs);
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord(ev[5], S->stream));
    return MGPU_OK;
}

}  // namespace mgpu

namespace {
// Buffers and constant tables of the learned quadrature (lattice hex8 sets); with need_phi the
// cut cells' nodal level-set values too (the network's inputs; the moment fit does not use them).
mgpu_status lq_prepare(mgpu_set* S, bool need_phi) {
    const SetHost& h = S->h;
    const int nc = S->v.n_cut;
    const char* what = need_phi ? "learned quadrature" : "moment-fitted quadrature";
    if (h.dim != 3) return set_error(MGPU_ERR_UNSUPPORTED, std::string(what) + " is implemented for 3D hex8 sets");
    if (h.generic)
        return set_error(MGPU_ERR_UNSUPPORTED, std::string(what) + " is implemented for lattice sets with nodal T rows");
    if (need_phi && nc > 0 && S->vphi_cut.empty())
        return set_error(MGPU_ERR_INVALID_ARG, "learned quadrature needs vertex_phi in the set descriptor");
    LqView& lv = S->lv;
    mgpu_status rc;
    if (need_phi && !lv.vphi && nc > 0)
        if ((rc = upload(S, S->vphi_cut.data(), S->vphi_cut.size(), const_cast<double**>(&lv.vphi))) != MGPU_OK) return rc;
    if (S->lq_ready) return MGPU_OK;
    lv.n_cut = nc;
    if (lv.tol == 0.0) lv.tol = 1e-6;
    if (lv.kappa == 0.0) lv.kappa = 4.0;
    lv.detJ_ip = 0.125 * h.lat.h[0] * h.lat.h[1] * h.lat.h[2];
    auto alloc = [&](size_t bytes, void** p) -> mgpu_status {
        HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 8)));
        S->allocs.push_back(*p);
        return MGPU_OK;
    };
    void* p;
    if ((rc = alloc((size_t)nc * kLqPoints * sizeof(float), &p)) != MGPU_OK) return rc;
    lv.w_pred = static_cast<float*>(p);
    if ((rc = alloc((size_t)nc * kLqMomentsPad * sizeof(double), &p)) != MGPU_OK) return rc;
    lv.mom = static_cast<double*>(p);
    if ((rc = alloc((size_t)nc * kLqPoints * sizeof(float), &p)) != MGPU_OK) return rc;
    lv.w_fit = static_cast<float*>(p);
    if ((rc = alloc((size_t)nc * sizeof(int32_t), &p)) != MGPU_OK) return rc;
    lv.flag = static_cast<int32_t*>(p);
    // work list of the moment kernel: each cut element takes ceil(sub-tets / 48) 16-lane quarters
    // of a wavefront (at most 4; a lane loops over about three sub-tets, which amortises the
    // wavefront's fixed cost: measured 3.5 % faster per batch than one sub-tet per lane);
    // 4-quarter elements go alone, 3-quarter ones beside a 1-quarter one, 2-quarter ones in pairs,
    // 1-quarter ones in fours
    {
        constexpr int kSubTetsPerLane = 3;
        std::vector<int32_t> byq[5];
        for (int32_t c = 0; c < nc; ++c) {
            const int nsub = S->cut_nip[c] / kTet11Points;
            byq[std::min(4, std::max(1, (nsub + 16 * kSubTetsPerLane - 1) / (16 * kSubTetsPerLane)))].push_back(c);
        }
        std::vector<int32_t> owner;
        size_t i1 = 0;
        auto take1 = [&]() { return i1 < byq[1].size() ? byq[1][i1++] : -1; };
        for (const int32_t c : byq[4]) owner.insert(owner.end(), {c, c, c, c});
        for (const int32_t c : byq[3]) owner.insert(owner.end(), {c, c, c, take1()});
        for (size_t k = 0; k < byq[2].size(); k += 2) {
            const int32_t a = byq[2][k];
            if (k + 1 < byq[2].size()) {
                owner.insert(owner.end(), {a, a, byq[2][k + 1], byq[2][k + 1]});
            } else {
                const int32_t x = take1(), y = take1();
                owner.insert(owner.end(), {a, a, x, y});
            }
        }
        while (i1 < byq[1].size()) {
            const int32_t a = take1(), b = take1(), c = take1(), d = take1();
            owner.insert(owner.end(), {a, b, c, d});
        }
        // records of 12: the 4 owners, then each quarter's first sub-cell and sub-cell count (so
        // that the kernel's first dependent load is the sub-cell coordinates)
        lv.n_mom_waves = (int32_t)(owner.size() / 4);
        std::vector<int32_t> rec((size_t)lv.n_mom_waves * 12, 0);
        for (int32_t w = 0; w < lv.n_mom_waves; ++w)
            for (int k = 0; k < 4; ++k) {
                const int32_t c = owner[(size_t)w * 4 + k];
                rec[(size_t)w * 12 + k] = c;
                if (c >= 0) {
                    rec[(size_t)w * 12 + 4 + k] = S->cut_ig0[(size_t)c];
                    rec[(size_t)w * 12 + 8 + k] = S->cut_nip[(size_t)c] / kTet11Points;
                }
            }
        if ((rc = upload(S, rec.data(), rec.size(), const_cast<int32_t**>(&lv.mom_owner))) != MGPU_OK) return rc;
    }
    LqOperators op;
    lq_build_operators(&op);
    std::vector<double> Vt(kLqPoints * kLqMoments), Pt(kLqMoments * kLqPoints);
    for (int q = 0; q < kLqPoints; ++q)
        for (int k = 0; k < kLqMoments; ++k) {
            Vt[q * kLqMoments + k] = op.V[k][q];
            Pt[k * kLqPoints + q] = op.P[q][k];
        }
    if ((rc = upload(S, Vt.data(), Vt.size(), const_cast<double**>(&lv.Vt))) != MGPU_OK) return rc;
    if ((rc = upload(S, Pt.data(), Pt.size(), const_cast<double**>(&lv.Pt))) != MGPU_OK) return rc;
    // physical gradients of the hex8 basis at the 27 points of the lattice cell (built in a host
    // translation unit: the tables are device constants here)
    std::vector<float> g27(kLqPoints * 24);
    lq_point_gradients(h.lat.h, g27.data());
    if ((rc = upload(S, g27.data(), g27.size(), const_cast<float**>(&lv.g27))) != MGPU_OK) return rc;
    S->lq_ready = true;
    return MGPU_OK;
}

}  // namespace

namespace mgpu {
int stencil_shape(int nf, int variant, int* n_variants) {
    if (nf == 1) {
        *n_variants = stencil_variants<1>();
        return stencil_rows<1>(variant);
    }
    if (nf == 3) {
        *n_variants = stencil_variants<3>();
        return stencil_rows<3>(variant);
    }
    *n_variants = stencil_variants<2>();
    return stencil_rows<2>(variant);
}

mgpu_status check_err_flag(mgpu_set* S) {
    if (!S->v.err_flag) {   // generic path: no row-length check
        HIP_TRY(hipStreamSynchronize(S->stream));
        return MGPU_OK;
    }
    int32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, S->v.err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, S->stream));
    HIP_TRY(hipStreamSynchronize(S->stream));
    if (flag != 0) {
        // reported once: the flag is cleared so that the next clean call succeeds
        static const int32_t zero = 0;
        HIP_TRY(hipMemcpy(S->v.err_flag, &zero, sizeof(int32_t), hipMemcpyHostToDevice));
        return mgpu::set_error(MGPU_ERR_STATE, "kernel row length disagrees with the CSR pattern (rows of that call "
                                               "were left unwritten)");
    }
    return MGPU_OK;
}

mgpu_status set_error(mgpu_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace mgpu

extern "C" const char* mgpu_last_error_string(void) { return g_last_error.c_str(); }

namespace mgpu {
// The per-set buffers of a lattice set after its inputs are on the device: the cut elements'
// matrices, their sub-cell counts and longest-first order (from the host copy of ig_ptr), the
// error flag, the cell records and the planar u.
mgpu_status finish_lattice_set(mgpu_set* S, const int32_t* ig_ptr) {
    SetHost& h = S->h;
    DevView& v = S->v;
    mgpu_status rc;
    if (v.n_cut > 0) {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, (size_t)v.n_cut * h.na * h.na * sizeof(double)));
        S->allocs.push_back(p);
        v.ke_cut = static_cast<double*>(p);
        HIP_TRY(hipMalloc(&p, (size_t)v.n_cut * h.na * h.na * sizeof(float)));
        S->allocs.push_back(p);
        v.ke_cut_f = static_cast<float*>(p);
        S->cut_nip.resize(v.n_cut);
        S->cut_ig0.resize(v.n_cut);
        for (int32_t c = 0; c < v.n_cut; ++c) {
            const int32_t ce = h.cut_elems[c];
            S->cut_nip[c] = (ig_ptr[ce + 1] - ig_ptr[ce]) * (h.dim == 3 ? kTet11Points : kTri7Points);
            S->cut_ig0[c] = ig_ptr[ce];
        }
        // longest-first processing order of the cut elements (sub-cell count descending, stable)
        std::vector<int32_t> order(v.n_cut);
        for (int32_t c = 0; c < v.n_cut; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            const int32_t ex = h.cut_elems[x], ey = h.cut_elems[y];
            return ig_ptr[ex + 1] - ig_ptr[ex] > ig_ptr[ey + 1] - ig_ptr[ey];
        });
        if ((rc = upload(S, order.data(), order.size(), const_cast<int32_t**>(&S->d_cut_order))) != MGPU_OK)
            return rc;
    }
    {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, sizeof(int32_t)));
        S->allocs.push_back(p);
        HIP_TRY(hipMemset(p, 0, sizeof(int32_t)));
        v.err_flag = static_cast<int32_t*>(p);
        const size_t n_cells = (size_t)v.nx * v.ny * v.nz;
        HIP_TRY(hipMalloc(&p, std::max<size_t>(n_cells, 1) * (h.nf == 1 ? 1 : 4) * sizeof(double)));
        S->allocs.push_back(p);
        v.cell_rec = static_cast<double*>(p);
        HIP_TRY(hipMalloc(&p, std::max<size_t>((size_t)h.n_nodes, 1) * h.nf * sizeof(double)));
        S->allocs.push_back(p);
        v.u_node = static_cast<double*>(p);
    }
    return MGPU_OK;
}
}  // namespace mgpu

extern "C" mgpu_status mgpu_set_create(const mgpu_set_desc* desc, mgpu_set** out) {
    MGPU_GUARD_BEGIN
    if (!desc || !out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    // the descriptor is validated before any device is touched
    auto S = std::make_unique<mgpu_set>();
    std::string err;
    mgpu_status rc = build_set_host(*desc, S->h, &err);
    if (rc != MGPU_OK) return set_error(rc, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(MGPU_ERR_NO_DEVICE, "no HIP device visible");
    SetHost& h = S->h;
    DevView& v = S->v;
    if (h.generic) {
        v.n_el = h.n_el;
        v.n_cut = (int32_t)h.cut_elems.size();
        v.N = h.n_adofs_type;
        v.n_rows = h.n_rows;
        S->n_ip_cut = (int64_t)desc->ig_ptr[h.n_el] * (h.nb == 27 ? kTet35Points : (h.dim == 3 ? kTet11Points : kTri7Points));
        if ((rc = generic_upload(S.get(), *desc)) != MGPU_OK) return rc;
        *out = S.release();
        return MGPU_OK;
    }
    v.n_el = h.n_el;
    v.n_cut = (int32_t)h.cut_elems.size();
    v.N = h.n_adofs_type;
    v.n_rows = h.n_rows;
    v.nx = h.lat.n[0];
    v.ny = h.lat.n[1];
    v.nz = h.dim == 3 ? h.lat.n[2] : 1;
    for (int d = 0; d < 3; ++d) v.h[d] = h.lat.h[d];
    const int nv_ig = h.dim + 1;
    const size_t n_ig = (size_t)desc->ig_ptr[h.n_el];
    S->n_ip_cut = (int64_t)n_ig * (h.dim == 3 ? kTet11Points : kTri7Points);
    if ((rc = upload(S.get(), desc->material, (size_t)h.n_el * h.nm, const_cast<double**>(&v.material))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.cell_elem.data(), h.cell_elem.size(), const_cast<int32_t**>(&v.cell_elem))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.elem_cut.data(), h.elem_cut.size(), const_cast<int32_t**>(&v.elem_cut))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.node_adof.data(), h.node_adof.size(), const_cast<int32_t**>(&v.node_adof))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.row_node.data(), h.row_node.size(), const_cast<int32_t**>(&v.row_node))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.row_ptr.data(), h.row_ptr.size(), const_cast<int32_t**>(&v.row_ptr))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), h.cut_elems.data(), h.cut_elems.size(), const_cast<int32_t**>(&v.cut_elems))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), desc->ig_ptr, (size_t)h.n_el + 1, const_cast<int32_t**>(&v.ig_ptr))) != MGPU_OK) return rc;
    if ((rc = upload(S.get(), desc->ig_xi, n_ig * nv_ig * h.dim, const_cast<double**>(&v.ig_xi))) != MGPU_OK) return rc;
    if (desc->vertex_phi && h.dim == 3) {
        S->vphi_cut.resize(h.cut_elems.size() * 8);
        for (size_t c = 0; c < h.cut_elems.size(); ++c)
            for (int k = 0; k < 8; ++k) S->vphi_cut[c * 8 + k] = desc->vertex_phi[(size_t)h.cut_elems[c] * 8 + k];
    }
    if ((rc = finish_lattice_set(S.get(), desc->ig_ptr)) != MGPU_OK) return rc;
    *out = S.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_destroy(mgpu_set* set) {
    delete set;   // ~mgpu_set frees the device buffers
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_options(mgpu_set* set, int32_t precision, int32_t quadrature) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (precision != MGPU_PREC_FP64 && precision != MGPU_PREC_MIXED) return set_error(MGPU_ERR_INVALID_ARG, "bad precision");
    if (quadrature != MGPU_QUAD_SUBDIVISION && quadrature != MGPU_QUAD_LEARNED && quadrature != MGPU_QUAD_MOMENT_FIT)
        return set_error(MGPU_ERR_INVALID_ARG, "bad quadrature");
    if (quadrature == MGPU_QUAD_LEARNED && !set->mlp)
        return set_error(MGPU_ERR_STATE, "learned quadrature needs a model (mgpu_set_quadrature_model)");
    if (quadrature == MGPU_QUAD_MOMENT_FIT) {
        if (!set->h.generic)
            if (mgpu_status rc = lq_prepare(set, false); rc != MGPU_OK) return rc;
    }
    set->precision = precision;
    set->quadrature = quadrature;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_stream(mgpu_set* set, void* stream) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    set->stream = static_cast<hipStream_t>(stream);
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_pattern_get(const mgpu_set* set, int32_t* n_rows, int64_t* nnz,
                                        const int32_t** row_ptr, const int32_t** col_idx) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (row_ptr || col_idx)
        if (mgpu_status rc = ensure_host(const_cast<mgpu_set*>(set)); rc != MGPU_OK) return rc;
    if (n_rows) *n_rows = set->h.n_rows;
    if (nnz) *nnz = set->h.nnz;
    if (row_ptr) *row_ptr = set->h.row_ptr.data();
    if (col_idx) *col_idx = set->h.col_idx.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_element_topology(const mgpu_set* set, int32_t* n_per_element, const int32_t** ids) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (mgpu_status rc = ensure_host(const_cast<mgpu_set*>(set)); rc != MGPU_OK) return rc;
    if (set->h.generic) {
        const auto& na_t = set->h.na_t;
        for (size_t e = 1; e < na_t.size(); ++e)
            if (na_t[e] != na_t[0])
                return set_error(MGPU_ERR_UNSUPPORTED, "elements of the set have topologies of different sizes");
        if (n_per_element) *n_per_element = na_t.empty() ? 0 : set->h.nf * na_t[0];
        if (ids) *ids = set->h.topo.data();
        return MGPU_OK;
    }
    if (n_per_element) *n_per_element = set->h.na;
    if (ids) *ids = set->h.topo.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_element_topology_csr(const mgpu_set* set, const int64_t** offsets, const int32_t** ids) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (mgpu_status rc = ensure_host(const_cast<mgpu_set*>(set)); rc != MGPU_OK) return rc;
    if (offsets) *offsets = set->h.topo_ptr.data();
    if (ids) *ids = set->h.topo.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_assemble(mgpu_set* set, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (set->h.n_rows > 0 && (!u || !vals || !rhs)) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    if (set->quadrature == MGPU_QUAD_LEARNED && !set->mlp) return set_error(MGPU_ERR_STATE, "no quadrature model");
    mgpu_status rc = launch_assemble(set, u, vals, rhs, nullptr);
    if (rc != MGPU_OK) return rc;
    set->stats[0] = set->h.n_el;
    set->stats[1] = set->v.n_cut;
    set->stats[2] = 0;
    set->stats[3] = 0;
    const int uncut_ip = set->h.nb == 27 ? kHex3Points : (set->h.dim == 3 ? kHex2Points : kQuad2Points);
    set->stats[4] = (int64_t)(set->h.n_el - set->v.n_cut) * uncut_ip + set->n_ip_cut;
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_assemble_host(mgpu_set* set, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    const size_t nr = (size_t)set->h.n_rows, nnz = (size_t)set->h.nnz;
    if (nr == 0) return MGPU_OK;
    if (!u || !vals || !rhs) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    if (!set->d_u_tmp) {
        HIP_TRY(hipMalloc(&set->d_u_tmp, nr * sizeof(double)));
        HIP_TRY(hipMalloc(&set->d_rhs_tmp, nr * sizeof(double)));
        HIP_TRY(hipMalloc(&set->d_vals_tmp, std::max<size_t>(nnz, 1) * sizeof(double)));
    }
    HIP_TRY(hipMemcpyAsync(set->d_u_tmp, u, nr * sizeof(double), hipMemcpyHostToDevice, set->stream));
    mgpu_status rc = mgpu_set_assemble(set, set->d_u_tmp, set->d_vals_tmp, set->d_rhs_tmp);
    if (rc != MGPU_OK) return rc;
    HIP_TRY(hipMemcpyAsync(vals, set->d_vals_tmp, nnz * sizeof(double), hipMemcpyDeviceToHost, set->stream));
    HIP_TRY(hipMemcpyAsync(rhs, set->d_rhs_tmp, nr * sizeof(double), hipMemcpyDeviceToHost, set->stream));
    HIP_TRY(hipStreamSynchronize(set->stream));
    return check_err_flag(set);
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_stats(const mgpu_set* set, int64_t stats[5]) {
    if (!set || !stats) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < 5; ++i) stats[i] = set->stats[i];
    if ((set->quadrature == MGPU_QUAD_LEARNED || (set->quadrature == MGPU_QUAD_MOMENT_FIT && !set->h.generic)) &&
        set->v.n_cut > 0) {
        std::vector<int32_t> f((size_t)set->v.n_cut);
        HIP_TRY(hipStreamSynchronize(set->stream));
        HIP_TRY(hipMemcpy(f.data(), set->lv.flag, f.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        int64_t acc = 0;
        for (int32_t x : f) acc += x != 0;
        stats[1] = set->v.n_cut - acc;
        stats[2] = acc;
        stats[3] = set->v.n_cut - acc;
        int64_t n_ip = (int64_t)(set->h.n_el - set->v.n_cut) * 8 + acc * kLqPoints;
        for (size_t c = 0; c < f.size(); ++c)
            if (f[c] == 0) n_ip += set->cut_nip[c];
        stats[4] = n_ip;
    }
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_statistics(const mgpu_set* set, int64_t* stats, int32_t n_stats) {
    if (!set || (!stats && n_stats > 0) || n_stats < 0) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    int64_t all[7];
    mgpu_status rc = mgpu_set_stats(set, all);
    if (rc != MGPU_OK) return rc;
    // points of sub-cells with detJ < -1e-6 are skipped in every mode; the learned path skips
    // them only on its fallback elements (its accepted elements use 27 points of the IP cell)
    all[5] = set->h.n_skip_points;
    all[6] = set->h.n_skip_subcells;
    for (int32_t i = 0; i < n_stats && i < 7; ++i) stats[i] = all[i];
    for (int32_t i = 7; i < n_stats; ++i) stats[i] = 0;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_quadrature_model(mgpu_set* set, const mgpu_mlp* mlp, double tol, int32_t iters,
                                                 double kappa) {
    MGPU_GUARD_BEGIN
    if (!set || !mlp) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    if (!(tol > 0.0) || iters < 0 || !(kappa >= 1.0))
        return set_error(MGPU_ERR_INVALID_ARG, "bad tolerance / iteration count / cancellation bound");
    mgpu_status rc = lq_prepare(set, true);
    if (rc != MGPU_OK) return rc;
    set->lv.tol = tol;
    set->lv.iters = iters;
    set->lv.kappa = kappa;
    set->mlp = mlp;
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_lq_result(const mgpu_set* set, float* w_pred, float* w_fit, int32_t* flags) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (!set->lq_ready) return set_error(MGPU_ERR_STATE, "no learned quadrature attached");
    const size_t nc = (size_t)set->v.n_cut;
    HIP_TRY(hipStreamSynchronize(set->stream));
    if (w_pred && nc) HIP_TRY(hipMemcpy(w_pred, set->lv.w_pred, nc * kLqPoints * sizeof(float), hipMemcpyDeviceToHost));
    if (w_fit && nc) HIP_TRY(hipMemcpy(w_fit, set->lv.w_fit, nc * kLqPoints * sizeof(float), hipMemcpyDeviceToHost));
    if (flags && nc) HIP_TRY(hipMemcpy(flags, set->lv.flag, nc * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_time_assemble(mgpu_set* set, const double* u, double* vals, double* rhs,
                                              int32_t iters, double* per_call_ms, double kernel_ms[5]) {
    MGPU_GUARD_BEGIN
    if (!set || iters < 1) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    if (set->quadrature == MGPU_QUAD_LEARNED && !set->mlp) return set_error(MGPU_ERR_STATE, "no quadrature model");
    // events released on every return path
    struct Events {
        hipEvent_t e[8] = {};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    for (auto& e : evs.e) HIP_TRY(hipEventCreate(&e));
    hipEvent_t* ev = evs.e;
    hipEvent_t t0 = evs.e[6], t1 = evs.e[7];
    double t[5] = {0, 0, 0, 0, 0};
    // one untimed call first: the first launch of each kernel in a process loads its code object
    {
        mgpu_status rc = launch_assemble(set, u, vals, rhs, nullptr);
        if (rc != MGPU_OK) return rc;
        HIP_TRY(hipStreamSynchronize(set->stream));
    }
    // per-stage segments, one call at a time
    for (int it = 0; it < iters; ++it) {
        mgpu_status rc = launch_assemble(set, u, vals, rhs, ev);
        if (rc != MGPU_OK) return rc;
        HIP_TRY(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; ++k) {
            float a = 0.f;
            HIP_TRY(hipEventElapsedTime(&a, ev[k], ev[k + 1]));
            t[k] += a;
        }
    }
    // back-to-back calls for the per-call mean
    HIP_TRY(hipEventRecord(t0, set->stream));
    for (int it = 0; it < iters; ++it) {
        mgpu_status rc = launch_assemble(set, u, vals, rhs, nullptr);
        if (rc != MGPU_OK) return rc;
    }
    HIP_TRY(hipEventRecord(t1, set->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, t0, t1));
    if (per_call_ms) *per_call_ms = total / iters;
    if (kernel_ms)
        for (int k = 0; k < 5; ++k) kernel_ms[k] = t[k] / iters;
    return check_err_flag(set);
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_device_count(int32_t* n) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (n) *n = c;
    return MGPU_OK;
}
extern "C" mgpu_status mgpu_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(MGPU_ERR_INVALID_ARG, "null ptr");
    HIP_TRY(hipMalloc(ptr, bytes));
    return MGPU_OK;
}
extern "C" mgpu_status mgpu_free(void* ptr) {
    HIP_TRY(hipFree(ptr));
    return MGPU_OK;
}
extern "C" mgpu_status mgpu_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MGPU_OK;
}
extern "C" mgpu_status mgpu_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MGPU_OK;
}
extern "C" mgpu_status mgpu_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return MGPU_OK;
}
