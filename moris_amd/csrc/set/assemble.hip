// assemble.hip -- HIP kernels of the FP64 element K+R assembly and the mgpu_set entry points that
// drive them.  Target: gfx950 (MI355X), 64-wide wavefronts.
//
// Path of one mgpu_set_assemble call on a lattice set:
//   k_cut_moments   one wavefront per cut element.  For every IG sub-cell and every point of its
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
//   k_stencil       owner-computes scatter: one thread per CSR row.  The row sums the K_e rows of
//                   the (up to 8 / 4) lattice cells around its node -- uncut cells from the
//                   block-shared reference moments (HEX_2x2x2 / QUAD_2x2 on the IP cell), cut cells
//                   from k_cut_moments -- and writes every value of the row exactly once; the row
//                   of R is the same row applied to u (R = sum_e K_e u_e, linear IWGs).  Replaces
//                   Epetra_FECrsMatrix::SumIntoGlobalValues (cl_Sparse_Matrix_EpetraFECrs.cpp
//                   :103-133); deterministic, no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../common/fem_tables.hpp"
#include "../setup/internal.hpp"
#include "set_internal.hpp"

namespace mgpu {

// ------------------------------------------------------------------------------------------------
// device-side view of a set
// ------------------------------------------------------------------------------------------------
struct DevView {
    int32_t n_el, n_cut, N, n_rows;
    int32_t nx, ny, nz;
    double h[3];
    const double* material;     // [n_el][nm]
    const int32_t* cell_elem;   // [n_cells]
    const int32_t* elem_cut;    // [n_el]
    const int32_t* node_adof;   // [n_nodes]
    const int32_t* row_node;    // [N]
    const int32_t* row_ptr;     // [n_rows + 1]
    const int32_t* cut_elems;   // [n_cut]
    const int32_t* ig_ptr;      // [n_el + 1]
    const double* ig_xi;        // [n_ig][nv][dim]
    const uint8_t* vert_pos;    // [n_el][nb]
    double* ke_cut;             // [n_cut][na][na]
    int32_t* err_flag;          // [1]
};

// Voigt index of the symmetric pair (i, j): 3D [xx,yy,zz,yz,xz,xy]
// (CM_Struc_Linear::eval_strain_3d, cl_FEM_CM_Struc_Linear.cpp:448-453); 2D [xx,yy,xy].
template <int DIM>
__host__ __device__ constexpr int voigt(int i, int j) {
    return DIM == 3 ? (i == j ? i : (i + j == 3 ? 3 : (i + j == 2 ? 4 : 5))) : (i == j ? i : 2);
}

// Structural non-zeros of the isotropic D: the normal block and the shear diagonal.
template <int DIM>
__host__ __device__ constexpr bool d_nonzero(int a, int b) {
    return DIM == 3 ? ((a < 3 && b < 3) || a == b) : ((a < 2 && b < 2) || a == b);
}

// Constitutive matrix D (Voigt, engineering shear):
//   3D  CM_Struc_Linear_Isotropic::full_3d, cl_FEM_CM_Struc_Linear_Isotropic.cpp:374-390
//   2D  CM_Struc_Linear_Isotropic::full_plane_stress, :248-259
template <int DIM>
__device__ inline void elastic_D(double E, double nu, double (*D)[DIM == 3 ? 6 : 3]) {
    constexpr int NS = DIM == 3 ? 6 : 3;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) D[i][j] = 0.0;
    if constexpr (DIM == 3) {
        const double pre = E / (1.0 + nu) / (1.0 - 2.0 * nu);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i][j] = i == j ? pre * (1.0 - nu) : pre * nu;
        for (int i = 3; i < 6; ++i) D[i][i] = pre * (1.0 - 2.0 * nu) / 2.0;
    } else {
        const double pre = E / (1.0 - nu * nu);
        D[0][0] = pre;
        D[1][1] = pre;
        D[0][1] = pre * nu;
        D[1][0] = pre * nu;
        D[2][2] = pre * 0.5 * (1.0 - nu);
    }
}

// dN/dxi of the linear Lagrange IP cell at xi: out[i][a].
template <int DIM>
__device__ inline void lagrange_dNdxi(const double* xi, double (*out)[DIM == 3 ? 8 : 4]) {
    if constexpr (DIM == 3) {
        for (int a = 0; a < 8; ++a) {
            const double sx = kHex8XiHat[a][0], sy = kHex8XiHat[a][1], sz = kHex8XiHat[a][2];
            const double fx = 1.0 + xi[0] * sx, fy = 1.0 + xi[1] * sy, fz = 1.0 + xi[2] * sz;
            out[0][a] = 0.125 * sx * fy * fz;
            out[1][a] = 0.125 * fx * sy * fz;
            out[2][a] = 0.125 * fx * fy * sz;
        }
    } else {
        for (int a = 0; a < 4; ++a) {
            const double sx = kQuad4XiHat[a][0], sy = kQuad4XiHat[a][1];
            out[0][a] = 0.25 * sx * (1.0 + xi[1] * sy);
            out[1][a] = 0.25 * (1.0 + xi[0] * sx) * sy;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// k_cut_moments: one wavefront per cut element
// ------------------------------------------------------------------------------------------------
template <int DIM, int NF>
__global__ __launch_bounds__(64) void k_cut_moments(DevView s) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;             // gradient vector length
    constexpr int NA = NF * NB;              // element adofs
    constexpr int NV = DIM + 1;              // IG simplex vertices
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    constexpr bool DIFF = NF == 1;
    // diffusion needs S[a][b] = sum_i M[(i,a),(i,b)] only
    constexpr int NACC_ENT = DIFF ? NB * NB : NG * NG;
    constexpr int NACC = (NACC_ENT + 63) / 64;
    constexpr int GPAD = NG + 1;

    __shared__ double sg[64 * GPAD];
    __shared__ double sw[64];
    __shared__ double sM[NACC_ENT];

    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int e = s.cut_elems[c];
    const int ig0 = s.ig_ptr[e], ig1 = s.ig_ptr[e + 1];
    const int nip = (ig1 - ig0) * NQ;
    double hh[DIM];
    for (int d = 0; d < DIM; ++d) hh[d] = s.h[d];

    double acc[NACC];
#pragma unroll
    for (int t = 0; t < NACC; ++t) acc[t] = 0.0;

    for (int base = 0; base < nip; base += 64) {
        const int ip = base + lane;
        double* g = sg + lane * GPAD;
        if (ip < nip) {
            const int sc = ig0 + ip / NQ;
            const int q = ip % NQ;
            const double* X = s.ig_xi + (size_t)sc * NV * DIM;
            double N[NV];
            double xq[DIM];
            if constexpr (DIM == 3) {
                xq[0] = kTet11Xi[q][0];
                xq[1] = kTet11Xi[q][1];
                xq[2] = kTet11Xi[q][2];
                N[0] = xq[0];
                N[1] = xq[1];
                N[2] = 1.0 - xq[0] - xq[1] - xq[2];
                N[3] = xq[2];
            } else {
                xq[0] = kTri7Xi[q][0];
                xq[1] = kTri7Xi[q][1];
                N[0] = xq[0];
                N[1] = xq[1];
                N[2] = 1.0 - xq[0] - xq[1];
            }
            double xip[DIM];
            for (int d = 0; d < DIM; ++d) {
                double v = 0.0;
                for (int k = 0; k < NV; ++k) v += N[k] * X[k * DIM + d];
                xip[d] = v;
            }
            // detJ of the simplex in physical space: rows X_r - X_2 (r = 0, 1 [, 3]) scaled by h/2
            double detJ;
            if constexpr (DIM == 3) {
                double J[3][3];
                const int rr[3] = {0, 1, 3};
                for (int r = 0; r < 3; ++r)
                    for (int d = 0; d < 3; ++d) J[r][d] = (X[rr[r] * 3 + d] - X[2 * 3 + d]) * (0.5 * hh[d]);
                detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
            } else {
                double J[2][2];
                for (int r = 0; r < 2; ++r)
                    for (int d = 0; d < 2; ++d) J[r][d] = (X[r * 2 + d] - X[2 * 2 + d]) * (0.5 * hh[d]);
                detJ = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / 2.0;
            }
            const double w = detJ < kDetJLowerLimit ? 0.0 : (DIM == 3 ? kTet11W[q] : kTri7W[q]) * detJ;
            double dN[DIM][NB];
            lagrange_dNdxi<DIM>(xip, dN);
            for (int i = 0; i < DIM; ++i)
                for (int a = 0; a < NB; ++a) g[i * NB + a] = dN[i][a] * (2.0 / hh[i]);
            sw[lane] = w;
        } else {
            for (int k = 0; k < NG; ++k) g[k] = 0.0;
            sw[lane] = 0.0;
        }
        __syncthreads();
        const int cnt = min(64, nip - base);
        for (int p = 0; p < cnt; ++p) {
            const double w = sw[p];
            const double* gp = sg + p * GPAD;
#pragma unroll
            for (int t = 0; t < NACC; ++t) {
                const int idx = lane + 64 * t;
                if (idx < NACC_ENT) {
                    if (DIFF) {
                        const int a = idx / NB, b = idx % NB;
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < DIM; ++i) v += gp[i * NB + a] * gp[i * NB + b];
                        acc[t] += w * v;
                    } else {
                        const int r = idx / NG, col = idx % NG;
                        acc[t] += w * gp[r] * gp[col];
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < NACC; ++t) {
        const int idx = lane + 64 * t;
        if (idx < NACC_ENT) sM[idx] = acc[t];
    }
    __syncthreads();

    // contraction with the constitutive matrix + projection to sorted adof order
    const uint8_t* vp = s.vert_pos + (size_t)e * NB;
    double* ke = s.ke_cut + (size_t)c * NA * NA;
    if (DIFF) {
        const double k = s.material[e];
        for (int idx = lane; idx < NB * NB; idx += 64) {
            const int a = idx / NB, b = idx % NB;
            ke[vp[a] * NA + vp[b]] = k * sM[idx];
        }
    } else {
        constexpr int NS = DIM == 3 ? 6 : 3;
        double D[NS][NS];
        elastic_D<DIM>(s.material[2 * e], s.material[2 * e + 1], D);
        for (int idx = lane; idx < NA * NA; idx += 64) {
            const int A = idx / NA, B = idx % NA;
            const int al = A / NB, a = A % NB, be = B / NB, b = B % NB;
            double v = 0.0;
            for (int i = 0; i < DIM; ++i)
                for (int j = 0; j < DIM; ++j)
                    if (d_nonzero<DIM>(voigt<DIM>(al, i), voigt<DIM>(be, j)))
                        v += D[voigt<DIM>(al, i)][voigt<DIM>(be, j)] * sM[(i * NB + a) * NG + (j * NB + b)];
            ke[(al * NB + vp[a]) * NA + (be * NB + vp[b])] = v;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// k_stencil: one thread per CSR row (dof type AL given by blockIdx.y)
// ------------------------------------------------------------------------------------------------
// vertex order -> sorted position for lattice cells (adof ids ascend with lattice node index)
template <int DIM>
__device__ constexpr int lattice_sorted_pos(int b) {
    return DIM == 3 ? (b == 2 ? 3 : (b == 3 ? 2 : (b == 6 ? 7 : (b == 7 ? 6 : b)))) : (b == 2 ? 3 : (b == 3 ? 2 : b));
}

template <int DIM, int NF, int AL>
__device__ inline void stencil_row(const DevView& s, const double* sM, int r, const double* u,
                                   double* vals, double* rhs) {
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

    double acc[NF][NS27];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int q = 0; q < NS27; ++q) acc[f][q] = 0.0;
    uint32_t mask = 0;

#pragma unroll
    for (int cc = 0; cc < (DIM == 3 ? 8 : 4); ++cc) {
        const int cx = (cc & 1) ? 0 : -1;
        const int cy = (cc & 2) ? 0 : -1;
        const int cz = DIM == 3 ? ((cc & 4) ? 0 : -1) : 0;
        const int ci = i + cx, cj = j + cy, ck = k + cz;
        if (ci < 0 || ci >= nx || cj < 0 || cj >= ny || (DIM == 3 && (ck < 0 || ck >= nz))) continue;
        const int cell = ci + nx * (cj + ny * ck);
        const int e = s.cell_elem[cell];
        if (e < 0) continue;
        // local vertex of the row node in this cell
        int a_loc = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
            if (off[0] == -cx && off[1] == -cy && (DIM == 2 || off[2] == -cz)) a_loc = b;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
            const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
            mask |= 1u << q;
        }
        const int cut = s.elem_cut[e];
        if (cut >= 0) {
            const double* ke = s.ke_cut + (size_t)cut * NA * NA + (size_t)(AL * NB + lattice_sorted_pos<DIM>(a_loc)) * NA;
#pragma unroll
            for (int be = 0; be < NF; ++be)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                    const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                    acc[be][q] += ke[be * NB + lattice_sorted_pos<DIM>(b)];
                }
        } else if (DIFF) {
            const double kk = s.material[e];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                acc[0][q] += kk * sM[a_loc * NB + b];
            }
        } else {
            constexpr int NS = DIM == 3 ? 6 : 3;
            double D[NS][NS];
            elastic_D<DIM>(s.material[2 * e], s.material[2 * e + 1], D);
#pragma unroll
            for (int be = 0; be < NF; ++be)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int* off = DIM == 3 ? kHex8Off[b] : kQuad4Off[b];
                    const int q = (cx + off[0] + 1) + 3 * (cy + off[1] + 1) + (DIM == 3 ? 9 * (cz + off[2] + 1) : 0);
                    double v = 0.0;
#pragma unroll
                    for (int ii = 0; ii < DIM; ++ii)
#pragma unroll
                        for (int jj = 0; jj < DIM; ++jj)
                            if (d_nonzero<DIM>(voigt<DIM>(AL, ii), voigt<DIM>(be, jj)))
                                v += D[voigt<DIM>(AL, ii)][voigt<DIM>(be, jj)] * sM[(ii * NB + a_loc) * NG + (jj * NB + b)];
                    acc[be][q] += v;
                }
        }
    }

    const int row = AL * s.N + r;
    const int start = s.row_ptr[row];
    const int cnt = __popc(mask);
    if (s.row_ptr[row + 1] - start != NF * cnt) {
        atomicOr(s.err_flag, 1);
        return;
    }
    double R = 0.0;
#pragma unroll
    for (int be = 0; be < NF; ++be) {
        int idx = start + be * cnt;
#pragma unroll
        for (int q = 0; q < NS27; ++q) {
            if (mask & (1u << q)) {
                const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = DIM == 3 ? q / 9 - 1 : 0;
                const int nbr = node + dx + (nx + 1) * (dy + (ny + 1) * dz);
                vals[idx++] = acc[be][q];
                R += acc[be][q] * u[(size_t)be * s.N + s.node_adof[nbr]];
            }
        }
    }
    rhs[row] = R;
}

template <int DIM, int NF>
__global__ __launch_bounds__(256) void k_stencil(DevView s, const double* u, double* vals, double* rhs) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr bool DIFF = NF == 1;
    constexpr int NM = DIFF ? NB * NB : NG * NG;
    constexpr int NQ = DIM == 3 ? kHex2Points : kQuad2Points;
    __shared__ double sM[NM];
    // reference moments of an uncut lattice cell: IG cell = IP cell, HEX_2x2x2 / QUAD_2x2,
    // detJ = prod(h/2) (box), dN/dx = dN/dxi * 2/h
    double detJ = 1.0;
    for (int d = 0; d < DIM; ++d) detJ *= 0.5 * s.h[d];
    for (int idx = threadIdx.x; idx < NM; idx += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < NQ; ++q) {
            double dN[DIM][NB];
            double xq[DIM];
            for (int d = 0; d < DIM; ++d) xq[d] = DIM == 3 ? kHex2Xi[q][d] : kQuad2Xi[q][d];
            lagrange_dNdxi<DIM>(xq, dN);
            const double w = (DIM == 3 ? kHex2W[q] : kQuad2W[q]) * detJ;
            if (DIFF) {
                const int a = idx / NB, b = idx % NB;
                double t = 0.0;
                for (int i = 0; i < DIM; ++i) t += (dN[i][a] * (2.0 / s.h[i])) * (dN[i][b] * (2.0 / s.h[i]));
                v += w * t;
            } else {
                const int ra = idx / NG, cb = idx % NG;
                const int i = ra / NB, a = ra % NB, jj = cb / NB, b = cb % NB;
                v += w * (dN[i][a] * (2.0 / s.h[i])) * (dN[jj][b] * (2.0 / s.h[jj]));
            }
        }
        sM[idx] = v;
    }
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.N) return;
    switch (blockIdx.y) {
        case 0: stencil_row<DIM, NF, 0>(s, sM, r, u, vals, rhs); break;
        case 1: if (NF > 1) stencil_row<DIM, NF, (NF > 1 ? 1 : 0)>(s, sM, r, u, vals, rhs); break;
        case 2: if (NF > 2) stencil_row<DIM, NF, (NF > 2 ? 2 : 0)>(s, sM, r, u, vals, rhs); break;
        default: break;
    }
}

}  // namespace mgpu

// ================================================================================================
// host side of the set
// ================================================================================================
using namespace mgpu;

struct mgpu_set {
    SetHost h;
    DevView v{};
    hipStream_t stream = nullptr;
    int precision = MGPU_PREC_FP64;
    int quadrature = MGPU_QUAD_SUBDIVISION;
    int64_t stats[5] = {0, 0, 0, 0, 0};
    int64_t n_ip_cut = 0;
    std::vector<void*> allocs;
    double* d_u_tmp = nullptr;
    double* d_vals_tmp = nullptr;
    double* d_rhs_tmp = nullptr;
};

namespace {

thread_local std::string g_last_error;

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

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

mgpu_status launch_assemble(mgpu_set* S, const double* u, double* vals, double* rhs, float* kms,
                            hipEvent_t* ev) {
    const SetHost& h = S->h;
    if (h.n_rows == 0) return MGPU_OK;
    const int dim = h.dim, nf = h.nf;
    if (ev) HIP_TRY(hipEventRecord(ev[0], S->stream));
    if (S->v.n_cut > 0) {
        dim3 g((unsigned)S->v.n_cut), b(64);
        if (dim == 3 && nf == 1) hipLaunchKernelGGL((k_cut_moments<3, 1>), g, b, 0, S->stream, S->v);
        else if (dim == 3 && nf == 3) hipLaunchKernelGGL((k_cut_moments<3, 3>), g, b, 0, S->stream, S->v);
        else if (dim == 2 && nf == 1) hipLaunchKernelGGL((k_cut_moments<2, 1>), g, b, 0, S->stream, S->v);
        else hipLaunchKernelGGL((k_cut_moments<2, 2>), g, b, 0, S->stream, S->v);
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord(ev[1], S->stream));
    {
        dim3 g((unsigned)((S->v.N + 255) / 256), (unsigned)nf), b(256);
        if (dim == 3 && nf == 1) hipLaunchKernelGGL((k_stencil<3, 1>), g, b, 0, S->stream, S->v, u, vals, rhs);
        else if (dim == 3 && nf == 3) hipLaunchKernelGGL((k_stencil<3, 3>), g, b, 0, S->stream, S->v, u, vals, rhs);
        else if (dim == 2 && nf == 1) hipLaunchKernelGGL((k_stencil<2, 1>), g, b, 0, S->stream, S->v, u, vals, rhs);
        else hipLaunchKernelGGL((k_stencil<2, 2>), g, b, 0, S->stream, S->v, u, vals, rhs);
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord(ev[2], S->stream));
    (void)kms;
    return MGPU_OK;
}

mgpu_status check_err_flag(mgpu_set* S) {
    int32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, S->v.err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, S->stream));
    HIP_TRY(hipStreamSynchronize(S->stream));
    if (flag != 0) return mgpu::set_error(MGPU_ERR_STATE, "kernel row length disagrees with the CSR pattern");
    return MGPU_OK;
}

}  // namespace

namespace mgpu {
mgpu_status set_error(mgpu_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace mgpu

extern "C" const char* mgpu_last_error_string(void) { return g_last_error.c_str(); }

extern "C" mgpu_status mgpu_set_create(const mgpu_set_desc* desc, mgpu_set** out) {
    MGPU_GUARD_BEGIN
    if (!desc || !out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(MGPU_ERR_NO_DEVICE, "no HIP device visible");
    auto S = std::make_unique<mgpu_set>();
    std::string err;
    mgpu_status rc = build_set_host(*desc, S->h, &err);
    if (rc != MGPU_OK) return set_error(rc, err);
    SetHost& h = S->h;
    DevView& v = S->v;
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
    if ((rc = upload(S.get(), h.vert_pos.data(), h.vert_pos.size(), const_cast<uint8_t**>(&v.vert_pos))) != MGPU_OK) return rc;
    if (v.n_cut > 0) {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, (size_t)v.n_cut * h.na * h.na * sizeof(double)));
        S->allocs.push_back(p);
        v.ke_cut = static_cast<double*>(p);
    }
    {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, sizeof(int32_t)));
        S->allocs.push_back(p);
        HIP_TRY(hipMemset(p, 0, sizeof(int32_t)));
        v.err_flag = static_cast<int32_t*>(p);
    }
    *out = S.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_destroy(mgpu_set* set) {
    if (!set) return MGPU_OK;
    for (void* p : set->allocs) (void)hipFree(p);
    if (set->d_u_tmp) (void)hipFree(set->d_u_tmp);
    if (set->d_vals_tmp) (void)hipFree(set->d_vals_tmp);
    if (set->d_rhs_tmp) (void)hipFree(set->d_rhs_tmp);
    delete set;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_options(mgpu_set* set, int32_t precision, int32_t quadrature) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (precision != MGPU_PREC_FP64 && precision != MGPU_PREC_MIXED) return set_error(MGPU_ERR_INVALID_ARG, "bad precision");
    if (quadrature != MGPU_QUAD_SUBDIVISION && quadrature != MGPU_QUAD_LEARNED) return set_error(MGPU_ERR_INVALID_ARG, "bad quadrature");
    if (precision != MGPU_PREC_FP64 || quadrature != MGPU_QUAD_SUBDIVISION)
        return set_error(MGPU_ERR_UNSUPPORTED, "only the FP64 subdivision path is built in this library");
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
    if (n_rows) *n_rows = set->h.n_rows;
    if (nnz) *nnz = (int64_t)set->h.col_idx.size();
    if (row_ptr) *row_ptr = set->h.row_ptr.data();
    if (col_idx) *col_idx = set->h.col_idx.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_element_topology(const mgpu_set* set, int32_t* n_per_element, const int32_t** ids) {
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (n_per_element) *n_per_element = set->h.na;
    if (ids) *ids = set->h.topo.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_assemble(mgpu_set* set, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    if (set->h.n_rows > 0 && (!u || !vals || !rhs)) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    mgpu_status rc = launch_assemble(set, u, vals, rhs, nullptr, nullptr);
    if (rc != MGPU_OK) return rc;
    set->stats[0] = set->h.n_el;
    set->stats[1] = set->v.n_cut;
    set->stats[2] = 0;
    set->stats[3] = 0;
    set->stats[4] = (int64_t)(set->h.n_el - set->v.n_cut) * (set->h.dim == 3 ? 8 : 4) + set->n_ip_cut;
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_set_assemble_host(mgpu_set* set, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!set) return set_error(MGPU_ERR_INVALID_ARG, "null set");
    const size_t nr = (size_t)set->h.n_rows, nnz = set->h.col_idx.size();
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
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_set_time_assemble(mgpu_set* set, const double* u, double* vals, double* rhs,
                                              int32_t iters, double* per_call_ms, double kernel_ms[4]) {
    MGPU_GUARD_BEGIN
    if (!set || iters < 1) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    hipEvent_t ev[3];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    double t_cut = 0.0, t_st = 0.0;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, set->stream));
    for (int it = 0; it < iters; ++it) {
        mgpu_status rc = launch_assemble(set, u, vals, rhs, nullptr, ev);
        if (rc != MGPU_OK) return rc;
        HIP_TRY(hipEventSynchronize(ev[2]));
        float a = 0.f, b = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[1], ev[2]));
        t_cut += a;
        t_st += b;
    }
    HIP_TRY(hipEventRecord(t1, set->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, t0, t1));
    if (per_call_ms) *per_call_ms = total / iters;
    if (kernel_ms) {
        kernel_ms[0] = t_cut / iters;
        kernel_ms[1] = t_st / iters;
        kernel_ms[2] = 0.0;
        kernel_ms[3] = 0.0;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
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
