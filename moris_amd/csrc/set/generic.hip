// generic.hip -- the generic element path of an mgpu_set: any T rows (B-spline p = 2, THB-like
// rows with several weighted entries) and sets without a lattice.
//
//   k_generic_element   one wavefront per element.  The moments
//                         M[(i,a),(j,b)] = sum_ip w* dN_a/dx_i dN_b/dx_j
//                       over the element's integration points (HEX_2x2x2 / QUAD_2x2 on the IP box
//                       when uncut, TET_11 / TRI_7 on the XTK sub-cells when cut; w* = w detJ,
//                       points with detJ < -1e-6 skipped, cl_FEM_Element_Bulk.cpp:258-355) are
//                       staged 64 points at a time in LDS; lane (a, b) owns the DIM x DIM block.
//                       K_p = B^T D B from M (IWG compute_jacobian,
//                       cl_FEM_IWG_Isotropic_Struc_Linear_Bulk.cpp:200, cl_FEM_IWG_Diffusion_Bulk.cpp
//                       :280), then K_e = T^T K_p T with the element's dense T built from its rows
//                       (Equation_Object::build_PADofMap / get_egn_obj_jacobian_and_residual,
//                       cl_MSI_Equation_Object.cpp:426-493, :789-842), in FP64 through LDS, written
//                       to the element's slot of the K_e pool.
//   k_generic_hex27     HEX27 IP cells (HEX_3x3x3 / TET_35): one block of four wavefronts per
//                       element, the moments on MFMA (FP64, or FP32 per 64-point chunk summed in
//                       FP64), K_p and K_e = T^T K_p T in FP64 through LDS (blocks h >= f,
//                       mirrored).  With MGPU_QUAD_MOMENT_FIT a cut element's TET_35 points only
//                       fit the 125 weights of a 5 x 5 x 5 Gauss-Legendre rule (FP64 MFMA), and M
//                       comes from those weights by sum factorisation (the same M up to rounding).
//                       In scatter mode (sets too large for a K_e pool) one launch per colour
//                       class, each element adding its K_e into the CSR values after k_generic_zero
//                       (FP64 atomics without return; one element per value within a colour).
//   k_generic_gather    one thread per CSR value: the sum of its K_e entries in ascending element
//                       order (deterministic SumIntoGlobalValues, cl_Sparse_Matrix_EpetraFECrs.cpp
//                       :103-133).
//   k_generic_rhs       one wavefront per row: R = K u (R_e = K_e u_e for these linear IWGs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../common/fem_tables.hpp"
#include "set_state.hpp"

namespace mgpu {

namespace {

template <class T>
mgpu_status gupload(mgpu_set* S, const T* src, size_t n, T** dst);

// the gather sources in 32 bits when every K_e entry index fits, else in 64 bits
mgpu_status upload_gather_src(mgpu_set* S, const std::vector<int64_t>& src, GenView& g) {
    g.gather_src = nullptr;
    g.gather_src32 = nullptr;
    const size_t n_ke = S->h.ke_off.empty() ? 0 : (size_t)S->h.ke_off.back();
    if (n_ke < ((size_t)1 << 31)) {
        std::vector<int32_t> s32(src.begin(), src.end());
        return gupload(S, s32.data(), s32.size(), const_cast<int32_t**>(&g.gather_src32));
    }
    return gupload(S, src.data(), src.size(), const_cast<int64_t**>(&g.gather_src));
}

template <class T>
mgpu_status gupload(mgpu_set* S, const T* src, size_t n, T** dst) {
    *dst = nullptr;
    if (n == 0) return MGPU_OK;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    S->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<T*>(p);
    return MGPU_OK;
}

// Integration point ip of element e: w* and the physical gradients g[i][a] of the IP cell.
template <int DIM>
__device__ inline double generic_point(const GenView& g, int e, bool cut, int ip, const double* hh,
                                       double (*grad)[DIM == 3 ? 8 : 4]) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    double xi[DIM];
    double w;
    if (!cut) {
        // IG cell = IP cell: HEX_2x2x2 / QUAD_2x2, detJ = prod(h / 2)
        for (int d = 0; d < DIM; ++d) xi[d] = DIM == 3 ? kHex2Xi[ip][d] : kQuad2Xi[ip][d];
        double detJ = 1.0;
        for (int d = 0; d < DIM; ++d) detJ *= 0.5 * hh[d];
        w = (DIM == 3 ? kHex2W[ip] : kQuad2W[ip]) * detJ;
    } else {
        constexpr int NV = DIM + 1;
        constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
        const int sc = g.ig_ptr[e] + ip / NQ;
        const int q = ip % NQ;
        const double* X = g.ig_xi + (size_t)sc * NV * DIM;
        double N[NV];
        if constexpr (DIM == 3) {
            N[0] = kTet11Xi[q][0];
            N[1] = kTet11Xi[q][1];
            N[2] = 1.0 - kTet11Xi[q][0] - kTet11Xi[q][1] - kTet11Xi[q][2];
            N[3] = kTet11Xi[q][2];
        } else {
            N[0] = kTri7Xi[q][0];
            N[1] = kTri7Xi[q][1];
            N[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
        }
        for (int d = 0; d < DIM; ++d) {
            double v = 0.0;
            for (int k = 0; k < NV; ++k) v += N[k] * X[k * DIM + d];
            xi[d] = v;
        }
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
        w = detJ < kDetJLowerLimit ? 0.0 : (DIM == 3 ? kTet11W[q] : kTri7W[q]) * detJ;
    }
    lagrange_dNdxi<DIM>(xi, grad);
    for (int i = 0; i < DIM; ++i)
        for (int a = 0; a < NB; ++a) grad[i][a] *= 2.0 / hh[i];
    return w;
}

constexpr int kGenLds = 24 * 24 + 8 * kGenericMaxAdofsPerType + 24 * 3 * kGenericMaxAdofsPerType;

// Moment-fitted rule of a cut HEX8 / QUAD4 cell (MGPU_QUAD_MOMENT_FIT on the generic path): the
// 3^DIM tensor points of the 3-point Gauss rule with weights W_g = sum_q w*_q L_i(x_q) L_j(y_q)
// (L_k(z_q)) over the cell's TET_11 / TRI_7 points, L the quadratics through the nodes -s, 0, s.
// The (bi)trilinear moment integrand has degree <= 2 per axis, so the rule gives the subdivision's
// sum.  One wavefront: lanes over the points, then a reduction in lane order into Wg.
constexpr int kGenFitW = kGenLds - 32;   // LDS offset of the fitted weights (doubles)
template <int DIM>
__device__ inline void gen_mfit_weights(const GenView& g, int e, const double* hh, double* lds, int lane) {
    constexpr int NV = DIM + 1;
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    constexpr int NW = DIM == 3 ? 27 : 9;
    static_assert(NW * 65 <= kGenFitW, "reduction rows and the fitted weights fit the LDS");
    const double sg = kHex3Xi[1][0];   // the 3-point Gauss node (the rule's literal)
    const int npts = (g.ig_ptr[e + 1] - g.ig_ptr[e]) * NQ;
    double acc[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) acc[k] = 0.0;
    for (int ip = lane; ip < npts; ip += 64) {
        const int sc = g.ig_ptr[e] + ip / NQ, q = ip % NQ;
        const double* X = g.ig_xi + (size_t)sc * NV * DIM;
        double N[NV];
        if constexpr (DIM == 3) {
            N[0] = kTet11Xi[q][0];
            N[1] = kTet11Xi[q][1];
            N[2] = 1.0 - kTet11Xi[q][0] - kTet11Xi[q][1] - kTet11Xi[q][2];
            N[3] = kTet11Xi[q][2];
        } else {
            N[0] = kTri7Xi[q][0];
            N[1] = kTri7Xi[q][1];
            N[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
        }
        double L[DIM][3];
        for (int d = 0; d < DIM; ++d) {
            double x = 0.0;
            for (int k = 0; k < NV; ++k) x += N[k] * X[k * DIM + d];
            L[d][0] = x * (x - sg) / (2.0 * sg * sg);
            L[d][1] = 1.0 - x * x / (sg * sg);
            L[d][2] = x * (x + sg) / (2.0 * sg * sg);
        }
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
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const int ix = k % 3, iy = (k / 3) % 3, iz = k / 9;
            double v = w * L[0][ix] * L[1][iy];
            if constexpr (DIM == 3) v *= L[2][iz];
            acc[k] += v;
        }
    }
#pragma unroll
    for (int k = 0; k < NW; ++k) lds[k * 65 + lane] = acc[k];
    __syncthreads();
    if (lane < NW) {
        double v = 0.0;
        for (int l = 0; l < 64; ++l) v += lds[lane * 65 + l];
        lds[kGenFitW + lane] = v;
    }
    __syncthreads();
}

template <int DIM, int NF, typename T, bool MFIT>
__global__ __launch_bounds__(64) void k_generic_element(GenView g) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NG = DIM * NB;
    constexpr int NP = NF * NB;
    constexpr int GP = NG + 1;
    static_assert(2 * 64 * GP * sizeof(T) <= kGenLds * sizeof(double), "LDS phases");
    __shared__ double lds[kGenLds];
    const int e = g.order ? g.order[blockIdx.x] : (int)blockIdx.x;
    const int lane = threadIdx.x;
    double hh[DIM];
    for (int d = 0; d < DIM; ++d) hh[d] = g.ip_box[(size_t)e * 2 * DIM + DIM + d];
    const bool cut = g.ig_ptr[e + 1] > g.ig_ptr[e];
    const bool fit = MFIT && cut;
    const int npts = fit ? (DIM == 3 ? 27 : 9)
                         : (cut ? (g.ig_ptr[e + 1] - g.ig_ptr[e]) * (DIM == 3 ? kTet11Points : kTri7Points)
                                : (DIM == 3 ? kHex2Points : kQuad2Points));
    if constexpr (MFIT)
        if (fit) gen_mfit_weights<DIM>(g, e, hh, lds, lane);

    // phase 1: moments, lane (a, b) owns M[(i,a),(j,b)], i, j < DIM
    T* sgw = reinterpret_cast<T*>(lds);
    T* sg = sgw + 64 * GP;
    const int a = (lane / NB) % NB, b = lane % NB;
    const bool owner = lane < NB * NB;
    T M[DIM][DIM];
    for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) M[i][j] = T(0);
    for (int p0 = 0; p0 < npts; p0 += 64) {
        const int ip = p0 + lane;
        if (ip < npts) {
            double grad[DIM][NB];
            double w;
            if (fit) {
                // tensor point ip of the 3-point Gauss rule, x fastest, and its fitted weight
                const double nd[3] = {-kHex3Xi[1][0], 0.0, kHex3Xi[1][0]};
                double xi[DIM];
                xi[0] = nd[ip % 3];
                xi[1] = nd[(ip / 3) % 3];
                if constexpr (DIM == 3) xi[2] = nd[ip / 9];
                lagrange_dNdxi<DIM>(xi, grad);
                for (int i = 0; i < DIM; ++i)
                    for (int c = 0; c < NB; ++c) grad[i][c] *= 2.0 / hh[i];
                w = lds[kGenFitW + ip];
            } else {
                w = generic_point<DIM>(g, e, cut, ip, hh, grad);
            }
            for (int i = 0; i < DIM; ++i)
                for (int c = 0; c < NB; ++c) {
                    sgw[lane * GP + i * NB + c] = (T)(w * grad[i][c]);
                    sg[lane * GP + i * NB + c] = (T)grad[i][c];
                }
        }
        __syncthreads();
        const int np = min(64, npts - p0);
        if (owner)
            for (int p = 0; p < np; ++p)
                for (int i = 0; i < DIM; ++i)
                    for (int j = 0; j < DIM; ++j) M[i][j] += sgw[p * GP + i * NB + a] * sg[p * GP + j * NB + b];
        __syncthreads();
    }

    // phase 2: K_p (NP x NP) from M and the material
    double* Kp = lds;
    double* Tt = Kp + NP * NP;                        // [NB][na_t]
    double* W = Tt + NB * kGenericMaxAdofsPerType;     // [NP][NF * na_t]
    const int na_t = g.na_t[e];
    const int NA = NF * na_t;
    if (owner) {
        if (NF == 1) {
            double v = 0.0;
            for (int i = 0; i < DIM; ++i) v += (double)M[i][i];
            Kp[a * NP + b] = g.material[e] * v;
        } else {
            constexpr int NS = DIM == 3 ? 6 : 3;
            double D[NS][NS];
            elastic_D<DIM>(g.material[2 * e], g.material[2 * e + 1], D);
            for (int al = 0; al < DIM; ++al)
                for (int be = 0; be < DIM; ++be) {
                    double v = 0.0;
                    for (int i = 0; i < DIM; ++i)
                        for (int j = 0; j < DIM; ++j)
                            if (d_nonzero<DIM>(voigt<DIM>(al, i), voigt<DIM>(be, j)))
                                v += D[voigt<DIM>(al, i)][voigt<DIM>(be, j)] * (double)M[i][j];
                    Kp[(al * NB + a) * NP + be * NB + b] = v;
                }
        }
    }
    // phase 3: dense element-local T (per dof type; the same rows serve every type)
    for (int idx = lane; idx < NB * na_t; idx += 64) Tt[idx] = 0.0;
    __syncthreads();
    if (lane < NB)
        for (int q = g.t_ptr[e * NB + lane]; q < g.t_ptr[e * NB + lane + 1]; ++q) Tt[lane * na_t + g.t_col[q]] += g.t_w[q];
    __syncthreads();
    // phase 4: W = K_p T (block diagonal T over dof types)
    for (int idx = lane; idx < NP * NA; idx += 64) {
        const int r = idx / NA, cy = idx % NA;
        const int gg = cy / na_t, y = cy % na_t;
        double v = 0.0;
        for (int b2 = 0; b2 < NB; ++b2) v += Kp[r * NP + gg * NB + b2] * Tt[b2 * na_t + y];
        W[idx] = v;
    }
    __syncthreads();
    // phase 5: K_e = T^T W
    double* ke = g.ke + g.ke_off[e];
    for (int idx = lane; idx < NA * NA; idx += 64) {
        const int X = idx / NA, Y = idx % NA;
        const int f = X / na_t, x = X % na_t;
        double v = 0.0;
        for (int b1 = 0; b1 < NB; ++b1) v += Tt[b1 * na_t + x] * W[(f * NB + b1) * NA + Y];
        ke[idx] = v;
    }
}

// HEX27 IP cells: one block of 4 wavefronts per element.  Points (HEX_3x3x3 uncut, TET_35 on the
// sub-tets, w* = w detJ, detJ < -1e-6 skipped) are staged 64 at a time as records of w* and the
// 1D quadratic Lagrange values and derivatives per axis, in the MFMA operand type (two LDS buffers,
// one barrier per chunk);
// each lane forms its MFMA operands, the physical gradients dN_a/dx_i (times w* for A), from them
// in registers, and the moments
//   M[(i,a),(j,b)] = sum_ip w* dN_a/dx_i dN_b/dx_j      (81 x 81, padded to 96 x 96)
// accumulate on MFMA: FP64 (v_mfma_f64_16x16x4f64) on the plain path; on the mixed path FP32
// products (v_mfma_f32_16x16x4f32) summed per 64-point chunk and added into FP64 registers.  M
// being symmetric, the four wavefronts compute its 21 upper tiles of 16 x 16 (6, 5, 5, 5) and
// mirror them.  Then K_p from
// M and the material in LDS, and K_e = T^T K_p T per (dof type, dof type) block in FP64.  The LDS
// (80 KB: K_p, T and one block of K_p T for up to 64 adofs per type, e.g. truncated hierarchical
// rows) uses gfx950's 160 KB per workgroup.
constexpr int kH27Chunk = 64;                    // points per staging round
static_assert((kH27Chunk & (kH27Chunk - 1)) == 0 && kH27Chunk >= 64, "4 slots per point, whole wavefronts");
constexpr int kH27PtStride = 20;                 // w*, then l[3][3], dl[3][3] of the point
constexpr int kH27Lds = 81 * 81 + 2 * 27 * kHex27MaxAdofsPerType;   // doubles: K_p, T, W of one block (80 KB)
static_assert(2 * kH27Chunk * kH27PtStride <= 81 * 81, "two point-record buffers fit the K_p region");

typedef double doublex4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// Moment-fitted rule of a cut HEX27 cell (MGPU_QUAD_MOMENT_FIT): the 5 x 5 x 5 Gauss-Legendre
// nodes xi_ijk with the weights
//   W_ijk = sum_q w*_q L_i(xi_q,x) L_j(xi_q,y) L_k(xi_q,z)
// over the cell's TET_35 points q, L the degree-4 Lagrange polynomials through the 5 nodes.  A
// polynomial f of degree <= 4 in each axis equals its interpolant sum_ijk f(xi_ijk) L_i L_j L_k, so
// sum_ijk W_ijk f(xi_ijk) = sum_q w*_q f(xi_q), the subdivision rule's sum.  The moment integrand
// dN_a/dx_i dN_b/dx_j of a box HEX27 cell (two triquadratic factors, one differentiated each) has
// degree <= 4 in each axis, so M and K equal the subdivision's up to rounding, from 125 points
// instead of 35 per sub-tet.
static constexpr double kGl5Host[5] = {-0.90617984593866399279762687829939, -0.53846931010568309103631442070021,
                                       0.0, 0.53846931010568309103631442070021, 0.90617984593866399279762687829939};
constexpr double gl5_inv_den(int i) {
    double d = 1.0;
    for (int m = 0; m < 5; ++m)
        if (m != i) d *= kGl5Host[i] - kGl5Host[m];
    return 1.0 / d;
}
static __constant__ const double kGl5[5] = {kGl5Host[0], kGl5Host[1], kGl5Host[2], kGl5Host[3], kGl5Host[4]};
static __constant__ const double kGl5InvDen[5] = {gl5_inv_den(0), gl5_inv_den(1), gl5_inv_den(2), gl5_inv_den(3),
                                                  gl5_inv_den(4)};
constexpr int kMfitPoints = 125;
constexpr int kMfitStride = 17;   // point record of the fit: w*, then L[3][5] (odd: the lanes of a staging wavefront, one point each, write distinct LDS banks)
static_assert(81 * 81 + 5 * 128 <= kH27Lds, "fitted weights and four partial sums after the K_p region");

// L_i(x), i < 5, through the Gauss-Legendre nodes (products of the other nodes' factors)
__device__ inline void lagrange5(double x, double* L) {
    double d[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) d[m] = x - kGl5[m];
    double pre = 1.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        L[i] = pre * kGl5InvDen[i];
        pre *= d[i];
    }
    double suf = 1.0;
#pragma unroll
    for (int i = 4; i >= 0; --i) {
        L[i] *= suf;
        suf *= d[i];
    }
}

// The fitted weights of cut element (ig0, nsub) into Wfit[125] (index (i 5 + j) 5 + k): the TET_35
// table in LDS once, the element's sub-tet vertices and their w-scale detJ (0 when detJ < -1e-6)
// copied to LDS up to kMfitTetBatch sub-tets at a time; the TET_35 points staged from them 64 at a
// time as FP64 records of w* and L[axis][5]; then
//   W[(i,j)][k] += (w* L_i(x) L_j(y)) L_k(z)
// on FP64 MFMA (rows (i, j) in two 16-row tiles, column k < 5; each wavefront over 16 points of a
// chunk), and the wavefronts' partial sums added in wavefront order (the same sum on every call).
constexpr int kMfitChunk = 128;                           // points per staging round of the fit
constexpr int kMfitTetBatch = 128;
constexpr int kMfitXs = 2 * kMfitChunk * kMfitStride;    // LDS offsets (doubles): staged vertices,
constexpr int kMfitDj = kMfitXs + 12 * kMfitTetBatch;     // detJ per staged sub-tet,
constexpr int kMfitT35 = kMfitDj + kMfitTetBatch;         // TET_35 [q][xi0, xi1, xi2, w]
static_assert(kMfitT35 + 4 * kTet35Points <= 81 * 81, "records, staged sub-tets and the rule fit the K_p region");
__device__ inline void mfit_weights(const GenView& g, int ig0, int nsub, const double* hh, double* lds, double* Wfit,
                                    int tid, int wave, int col16, int kq) {
    doublex4_t acc[2] = {doublex4_t{0.0, 0.0, 0.0, 0.0}, doublex4_t{0.0, 0.0, 0.0, 0.0}};
    // operand roles of the lane, as record offsets and 0 / 1 masks, so that every LDS read is
    // unconditional (a branch around a read keeps the compiler from batching the reads of a chunk)
    int ao[2], bo;
    double am[2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        const int r = rt * 16 + col16, rc = r < 25 ? r : 0;
        ao[rt] = (1 + rc / 5) | ((6 + rc % 5) << 8);
        am[rt] = r < 25 ? 1.0 : 0.0;
    }
    bo = 11 + (col16 < 5 ? col16 : 0);
    const double bm = col16 < 5 ? 1.0 : 0.0;
    double* Xs = lds + kMfitXs;
    double* Dj = lds + kMfitDj;
    double* T35 = lds + kMfitT35;
    if (tid < 4 * kTet35Points) {
        const int q = tid >> 2, c = tid & 3;
        T35[tid] = c < 3 ? kTet35Xi[q][c] : kTet35W[q];
    }
    for (int t0 = 0; t0 < nsub; t0 += kMfitTetBatch) {
        const int nt = min(kMfitTetBatch, nsub - t0), npts = nt * kTet35Points;
        if (t0 > 0) __syncthreads();   // the previous batch's records and vertices are read
        const double* src = g.ig_xi + (size_t)(ig0 + t0) * 12;
        for (int idx = tid; idx < nt * 12; idx += 256) Xs[idx] = src[idx];
        if (tid < nt) {
            const double* X = src + (size_t)tid * 12;
            double J[3][3];
            const int rr[3] = {0, 1, 3};
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) J[r][d] = (X[rr[r] * 3 + d] - X[2 * 3 + d]) * (0.5 * hh[d]);
            const double detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
            Dj[tid] = detJ < kDetJLowerLimit ? 0.0 : detJ;
        }
        __syncthreads();
        // software-pipelined: the MFMA work of chunk c is issued before chunk c + 1 is staged into
        // the other buffer, so the staging arithmetic overlaps the matrix work; one barrier per chunk
        const int nch = (npts + kMfitChunk - 1) / kMfitChunk;
        auto stage = [&](int c) {
            double* pts = lds + (c & 1) * kMfitChunk * kMfitStride;
            for (int slot = tid; slot < 4 * kMfitChunk; slot += 256) {
                const int pl = slot & (kMfitChunk - 1), role = slot / kMfitChunk;
                const int ip = c * kMfitChunk + pl;
                double* pp = pts + pl * kMfitStride;
                // points past the end read point 0's data and get weight 0
                const bool in = ip < npts;
                const int ipc = in ? ip : 0;
                const int t = ipc / kTet35Points, q = ipc - kTet35Points * t;
                const double* X = Xs + t * 12;
                const double* tq = T35 + 4 * q;
                if (role < 3) {
                    const int d = role;
                    const double xd = tq[0] * X[d] + tq[1] * X[3 + d] + (1.0 - tq[0] - tq[1] - tq[2]) * X[6 + d] +
                                      tq[2] * X[9 + d];
                    double L[5];
                    lagrange5(xd, L);
#pragma unroll
                    for (int k = 0; k < 5; ++k) pp[1 + 5 * d + k] = L[k];
                } else {
                    pp[0] = in ? tq[3] * Dj[t] : 0.0;
                }
            }
        };
        stage(0);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            const double* pts = lds + (c & 1) * kMfitChunk * kMfitStride;
#pragma unroll
            for (int s = 0; s < kMfitChunk / 16; ++s) {
                const double* pp = pts + (wave * (kMfitChunk / 4) + 4 * s + kq) * kMfitStride;
                const double b = bm * pp[bo];
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const double a = am[rt] * ((pp[0] * pp[ao[rt] & 255]) * pp[ao[rt] >> 8]);
                    acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rt], 0, 0, 0);
                }
            }
            if (c + 1 < nch) stage(c + 1);
            __syncthreads();
        }
    }
    // partial sums of the wavefronts after the 128 fitted weights (outside the record buffers)
    double* Wp = Wfit + 128 + wave * 128;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = rt * 16 + kq + 4 * j;
            if (r < 25 && col16 < 5) Wp[r * 5 + col16] = acc[rt][j];
        }
    __syncthreads();
    if (tid < kMfitPoints) Wfit[tid] = ((Wfit[128 + tid] + Wfit[256 + tid]) + Wfit[384 + tid]) + Wfit[512 + tid];
    __syncthreads();
}

// M is symmetric: only the 21 tiles (mt <= nt) of the 6 x 6 tile grid are computed, the rest
// mirrored.  Wavefront W owns H27Tiles<W> (rows sharing an A operand, columns a B operand):
// W0 row 0; W1 row 1; W2 row 2 and (5,5); W3 rows 3 and 4.
template <int W> struct H27Tiles;
template <> struct H27Tiles<0> {
    static constexpr int n = 6;
    static constexpr int mt[6] = {0, 0, 0, 0, 0, 0}, nt[6] = {0, 1, 2, 3, 4, 5};
};
template <> struct H27Tiles<1> {
    static constexpr int n = 5;
    static constexpr int mt[6] = {1, 1, 1, 1, 1, 0}, nt[6] = {1, 2, 3, 4, 5, 0};
};
template <> struct H27Tiles<2> {
    static constexpr int n = 5;
    static constexpr int mt[6] = {2, 2, 2, 2, 5, 0}, nt[6] = {2, 3, 4, 5, 5, 0};
};
template <> struct H27Tiles<3> {
    static constexpr int n = 5;
    static constexpr int mt[6] = {3, 3, 3, 4, 4, 0}, nt[6] = {3, 4, 5, 4, 5, 0};
};

// M tiles += Gw^T G over one chunk of points (A: row lane % 16 of the tile, k = lane / 16; B: k =
// lane / 16, column lane % 16), FP64 MFMA or FP32 MFMA summed into the FP64 accumulators, straight
// from the chunk's point records: the lane forms the operands of its tiles in registers (column (i, a) of point p is ((2/h_i dl_i) l_j) l_k, times w* for A), so that no
// staged gradient table and no barrier sit between the point records and the MFMA work.  go / gs:
// the lane's record offsets (packed 3 x 8 bits) and scale for column group cg (columns cg*16 + lane
// % 16; zero beyond column 80).
template <int W, typename T>
__device__ inline void h27_chunk_direct(const T* pts, const int* go, const T* gs, int kq, doublex4_t* acc) {
    using TL = H27Tiles<W>;
    constexpr int lo = W < 3 ? W : 3;   // the lowest column group of the wavefront's tiles
    [[maybe_unused]] floatx4_t accf[6];
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int t = 0; t < 6; ++t) accf[t] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll 2
    for (int ks = 0; ks < kH27Chunk / 4; ++ks) {
        const T* pp = pts + (4 * ks + kq) * kH27PtStride;
        const T w = pp[0];
        T v[6];
#pragma unroll
        for (int cg = lo; cg < 6; ++cg) {
            const int o = go[cg];
            v[cg] = ((gs[cg] * pp[o & 255]) * pp[(o >> 8) & 255]) * pp[o >> 16];
        }
#pragma unroll
        for (int t = 0; t < TL::n; ++t) {
            const T a = v[TL::mt[t]] * w, b = v[TL::nt[t]];
            if constexpr (sizeof(T) == 8) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
            else accf[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, accf[t], 0, 0, 0);
        }
    }
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int t = 0; t < TL::n; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[t][j] += (double)accf[t][j];
    }
}

// The wavefront's tiles to M in LDS (rows, columns < 81) and their mirror images.  FP64 MFMA:
// register j of lane l holds row l / 16 + 4 j; FP32 MFMA: row 4 (l / 16) + j; column l % 16.
template <int W, typename T>
__device__ inline void h27_store(double* Ms, const doublex4_t* acc, int col16, int kq) {
    using TL = H27Tiles<W>;
#pragma unroll
    for (int t = 0; t < TL::n; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = TL::mt[t] * 16 + (sizeof(T) == 8 ? kq + 4 * j : 4 * kq + j);
            const int c = TL::nt[t] * 16 + col16;
            if (r < 81 && c < 81) {
                Ms[r * 81 + c] = acc[t][j];
                if (TL::mt[t] != TL::nt[t]) Ms[c * 81 + r] = acc[t][j];
            }
        }
}

// M of a cut element from its fitted weights, by sum factorisation over the tensor grid of nodes:
//   M[(i,a),(j,b)] = s_i s_j sum_{x,y,z} W_xyz P_0(a_0,b_0; x) P_1(a_1,b_1; y) P_2(a_2,b_2; z),
// (a_0, a_1, a_2) the 1D indices of HEX27 basis a (kHex27Half), s_i = 2 / h_i, and
// P_d(alpha, beta; m) = F_alpha(xi_m) F'_beta(xi_m), F the quadratic Lagrange value, differentiated
// on the gradient's axis (d == i for alpha, d == j for beta): four factor types tau = 2 [d == i] +
// [d == j] per axis.  Contracted over z, then y, then x in FP64 (about 38k multiply-adds where 125
// points x 81 x 81 would take 820k); the six blocks i <= j are computed, the other three mirrored.
// LDS after the K_p region (doubles): the fitted weights, then the tables below.
constexpr int kSfP = 128;                   // P[tau][alpha][beta][m], 4 x 3 x 3 x 5
constexpr int kSfZ = kSfP + 180;            // Z[zt][x 5 + y][alpha_2 beta_2], tau_2 = {0, 1, 3}[zt]
constexpr int kSfY = kSfZ + 3 * 25 * 9;     // Y[block][x][alpha_1 beta_1][alpha_2 beta_2]
constexpr int kSfInv = kSfY + 6 * 5 * 81;   // int [27]: HEX27 basis of the 1D indices (a_0 9 + a_1 3 + a_2)
static_assert(81 * 81 + kSfInv + 14 <= kH27Lds, "sum factorisation tables fit after the K_p region");
static_assert(kSfP >= kMfitPoints, "the tables follow the fitted weights");
__device__ inline void mfit_moments(double* lds, const double* hh, int tid) {
    double* base = lds + 81 * 81;
    const double* W = base;
    double* P = base + kSfP;
    double* Z = base + kSfZ;
    double* Y = base + kSfY;
    int* inv = reinterpret_cast<int*>(base + kSfInv);
    if (tid < 180) {
        const int tau = tid / 45, r = tid - 45 * tau, al = r / 15, be = (r / 5) % 3, m = r % 5;
        double l[3], dl[3];
        lagrange_quadratic_1d(kGl5[m], l, dl);
        const double fa = (tau & 2) ? (al == 0 ? dl[0] : (al == 1 ? dl[1] : dl[2])) : (al == 0 ? l[0] : (al == 1 ? l[1] : l[2]));
        const double fb = (tau & 1) ? (be == 0 ? dl[0] : (be == 1 ? dl[1] : dl[2])) : (be == 0 ? l[0] : (be == 1 ? l[1] : l[2]));
        P[tid] = fa * fb;
    } else if (tid >= 192 && tid < 192 + 27) {
        const int a = tid - 192;
        inv[(kHex27Half[a][0] * 3 + kHex27Half[a][1]) * 3 + kHex27Half[a][2]] = a;
    }
    __syncthreads();
    for (int idx = tid; idx < 3 * 225; idx += 256) {
        const int zt = idx / 225, r = idx - 225 * zt, xy = r / 9, ab = r - 9 * xy;
        const double* Pp = P + ((zt == 2 ? 3 : zt) * 9 + ab) * 5;
        const double* Wp = W + xy * 5;
        double v = 0.0;
        for (int m = 0; m < 5; ++m) v += Wp[m] * Pp[m];
        Z[idx] = v;
    }
    __syncthreads();
    // block p: (i, j) = (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
    for (int idx = tid; idx < 6 * 405; idx += 256) {
        const int p = idx / 405, r = idx - 405 * p, x = r / 81, ab12 = r - 81 * x, ab1 = ab12 / 9, ab2 = ab12 - 9 * ab1;
        const int i = (p >= 3) + (p >= 5), j = p < 3 ? p : (p < 5 ? p - 2 : 2);
        const int tau1 = 2 * (i == 1) + (j == 1), tau2 = 2 * (i == 2) + (j == 2);
        const double* Zp = Z + (tau2 == 3 ? 2 : tau2) * 225 + x * 45 + ab2;
        const double* Pp = P + (tau1 * 9 + ab1) * 5;
        double v = 0.0;
        for (int y = 0; y < 5; ++y) v += Zp[9 * y] * Pp[y];
        Y[idx] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < 6 * 729; idx += 256) {
        const int p = idx / 729, r = idx - 729 * p, ab0 = r / 81, ab12 = r - 81 * ab0;
        const int i = (p >= 3) + (p >= 5), j = p < 3 ? p : (p < 5 ? p - 2 : 2);
        const int tau0 = 2 * (i == 0) + (j == 0);
        const double* Yp = Y + p * 405 + ab12;
        const double* Pp = P + (tau0 * 9 + ab0) * 5;
        double v = 0.0;
        for (int x = 0; x < 5; ++x) v += Yp[81 * x] * Pp[x];
        v *= (2.0 / hh[i]) * (2.0 / hh[j]);
        const int ab1 = ab12 / 9, ab2 = ab12 - 9 * ab1;
        const int a = inv[((ab0 / 3) * 3 + ab1 / 3) * 3 + ab2 / 3], b = inv[((ab0 % 3) * 3 + ab1 % 3) * 3 + ab2 % 3];
        lds[(i * 27 + a) * 81 + j * 27 + b] = v;
        if (i != j) lds[(j * 27 + b) * 81 + i * 27 + a] = v;
    }
}

template <int NF, typename T, bool SCATTER, bool MFIT>
__global__ __launch_bounds__(256) void k_generic_hex27(GenView g, int c0, double* vals) {
    constexpr int NB = 27, NG = 81, NP = NF * NB;
    __shared__ double lds[kH27Lds];
    // scatter mode: block b of a launch takes element b of the colour starting at c0
    const int e = SCATTER ? g.colour_el[c0 + blockIdx.x] : (g.order ? g.order[blockIdx.x] : (int)blockIdx.x);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    double hh[3];
    for (int d = 0; d < 3; ++d) hh[d] = g.ip_box[(size_t)e * 6 + 3 + d];
    const int ig0 = g.ig_ptr[e], nsub = g.ig_ptr[e + 1] - ig0;
    const bool cut = nsub > 0;
    // moment-fitted cut elements take M from the fitted weights (mfit_moments), not from points
    const bool fit = MFIT && cut;
    const int npts = fit ? 0 : (cut ? nsub * kTet35Points : kHex3Points);

    const int col16 = lane & 15, kq = lane >> 4;
    // moment-fitted rule (cut elements): the 125 weights after the K_p region
    if constexpr (MFIT)
        if (fit) mfit_weights(g, ig0, nsub, hh, lds, lds + NG * NG, tid, wave, col16, kq);
    // the lane's operand columns cg * 16 + lane % 16, cg < 6: (i, a) = (col / 27, col % 27), the
    // record offsets of dl_i(xi_i), l(xi_j), l(xi_k) and 2 / h_i (zero beyond column 80)
    int go[6];
    T gs[6];   // the records and operands are in T (FP64, or FP32 on the mixed path)
#pragma unroll
    for (int cg = 0; cg < 6; ++cg) {
        const int col = cg * 16 + col16;
        go[cg] = 0;
        gs[cg] = T(0);
        if (col < NG) {
            const int i = col / NB, a = col - NB * (col / NB);
            go[cg] = ((i == 0 ? 10 : 1) + kHex27Half[a][0]) | (((i == 1 ? 10 : 1) + 3 + kHex27Half[a][1]) << 8) |
                     (((i == 2 ? 10 : 1) + 6 + kHex27Half[a][2]) << 16);
            gs[cg] = (T)(2.0 / hh[i]);
        }
    }
    doublex4_t acc[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) acc[t] = doublex4_t{0.0, 0.0, 0.0, 0.0};

    // point records in two buffers: a chunk's records are written while the other wavefronts may
    // still read the previous chunk's, and the barrier of a chunk orders the one before it
    for (int p0 = 0, buf = 0; p0 < npts; p0 += kH27Chunk, buf ^= 1) {
        T* pts = reinterpret_cast<T*>(lds) + buf * kH27Chunk * kH27PtStride;
        // the chunk's points: w* and the 1D quadratic Lagrange values / derivatives per axis, by 4
        // threads per point (one per axis, one for the weight)
        for (int slot = tid; slot < 4 * kH27Chunk; slot += 256) {
            const int pl = slot & (kH27Chunk - 1), role = slot / kH27Chunk;
            const int ip = p0 + pl;
            T* pp = pts + pl * kH27PtStride;
            const int t = ip / kTet35Points, q = ip - kTet35Points * t;
            if (role < 3) {
                const int d = role;
                double xd = 0.0;
                if (ip < npts) {
                    if (!cut) {
                        xd = kHex3Xi[ip][d];
                    } else {
                        const double* X = g.ig_xi + (size_t)(ig0 + t) * 12;
                        const double N[4] = {kTet35Xi[q][0], kTet35Xi[q][1],
                                             1.0 - kTet35Xi[q][0] - kTet35Xi[q][1] - kTet35Xi[q][2], kTet35Xi[q][2]};
                        xd = N[0] * X[d] + N[1] * X[3 + d] + N[2] * X[6 + d] + N[3] * X[9 + d];
                    }
                }
                double l[3], dl[3];
                lagrange_quadratic_1d(xd, l, dl);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    pp[1 + 3 * d + k] = (T)l[k];
                    pp[10 + 3 * d + k] = (T)dl[k];
                }
            } else {
                double w = 0.0;
                if (ip < npts) {
                    if (!cut) {
                        double detJ = 1.0;
                        for (int d = 0; d < 3; ++d) detJ *= 0.5 * hh[d];
                        w = kHex3W[ip] * detJ;
                    } else {
                        const double* X = g.ig_xi + (size_t)(ig0 + t) * 12;
                        double J[3][3];
                        const int rr[3] = {0, 1, 3};
                        for (int r = 0; r < 3; ++r)
                            for (int d = 0; d < 3; ++d) J[r][d] = (X[rr[r] * 3 + d] - X[2 * 3 + d]) * (0.5 * hh[d]);
                        const double detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
                        w = detJ < kDetJLowerLimit ? 0.0 : kTet35W[q] * detJ;
                    }
                }
                pp[0] = (T)w;
            }
        }
        __syncthreads();
        switch (wave) {
            case 0: h27_chunk_direct<0, T>(pts, go, gs, kq, acc); break;
            case 1: h27_chunk_direct<1, T>(pts, go, gs, kq, acc); break;
            case 2: h27_chunk_direct<2, T>(pts, go, gs, kq, acc); break;
            default: h27_chunk_direct<3, T>(pts, go, gs, kq, acc); break;
        }
    }
    __syncthreads();   // the last chunk's records are read before M overwrites them

    // M to LDS: the wavefront's tiles and their mirror images, or (moment-fitted cut elements) the
    // sum factorisation of the fitted weights
    double* Ms = lds;
    if (fit) {
        mfit_moments(lds, hh, tid);
    } else {
        switch (wave) {
            case 0: h27_store<0, T>(Ms, acc, col16, kq); break;
            case 1: h27_store<1, T>(Ms, acc, col16, kq); break;
            case 2: h27_store<2, T>(Ms, acc, col16, kq); break;
            default: h27_store<3, T>(Ms, acc, col16, kq); break;
        }
    }
    __syncthreads();

    // K_p from M: thread pairs (a, b) read their 9 moments, then write their K_p entries
    double kv[3][9];
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) {
        const int pr = tid + 256 * s2;
        if (pr < NB * NB) {
            const int a = pr / NB, b = pr - NB * (pr / NB);
            double Mv[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Mv[i][j] = Ms[(i * NB + a) * NG + j * NB + b];
            if (NF == 1) {
                kv[s2][0] = g.material[e] * (Mv[0][0] + Mv[1][1] + Mv[2][2]);
            } else {
                double D[6][6];
                elastic_D<3>(g.material[2 * e], g.material[2 * e + 1], D);
                for (int al = 0; al < 3; ++al)
                    for (int be = 0; be < 3; ++be) {
                        double v = 0.0;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                if (d_nonzero<3>(voigt<3>(al, i), voigt<3>(be, j)))
                                    v += D[voigt<3>(al, i)][voigt<3>(be, j)] * Mv[i][j];
                        kv[s2][al * 3 + be] = v;
                    }
            }
        }
    }
    __syncthreads();
    double* Kp = lds;
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) {
        const int pr = tid + 256 * s2;
        if (pr < NB * NB) {
            const int a = pr / NB, b = pr - NB * (pr / NB);
            if (NF == 1) {
                Kp[a * NP + b] = kv[s2][0];
            } else {
                for (int al = 0; al < 3; ++al)
                    for (int be = 0; be < 3; ++be) Kp[(al * NB + a) * NP + be * NB + b] = kv[s2][al * 3 + be];
            }
        }
    }
    // dense element-local T (27 x na_t, the same for every dof type)
    double* Tt = lds + NG * NG;
    double* W = Tt + NB * kHex27MaxAdofsPerType;
    const int na_t = g.na_t[e];
    const int NA = NF * na_t;
    for (int idx = tid; idx < NB * na_t; idx += 256) Tt[idx] = 0.0;
    __syncthreads();
    if (tid < NB)
        for (int q = g.t_ptr[e * NB + tid]; q < g.t_ptr[e * NB + tid + 1]; ++q) Tt[tid * na_t + g.t_col[q]] += g.t_w[q];
    __syncthreads();
    // K_e = T^T K_p T per (f, h) block on FP64 MFMA (16 x 16 tiles, k in steps of 4, operands
    // from LDS; A: row lane % 16, k = lane / 16; B: k = lane / 16, column lane % 16; result
    // register j of a lane: row lane / 16 + 4 j, column lane % 16): W = K_p[f][h] T (27 x na_t),
    // then T^T W.  The wavefronts take the tiles in turn.  K_p being symmetric, so is K_e: only the
    // blocks h >= f are computed, block (h, f) is their mirror image.  K_e is stored to the
    // element's slot of the pool, or (scatter mode) added into the CSR values: entry (x, y) of block
    // (f, h) sits in row topo[f na_t + x], field block h of that row, at the rank of adof y among the
    // row adof's neighbours.  No other element of the colour touches the row, so the adds are
    // hardware FP64 atomics without return (global_atomic_add_f64, done at the L2: no old value
    // travels to the CU) that still reach every value in colour order: the same sums as a
    // read-modify-write, deterministic.
    double* ke = SCATTER ? nullptr : g.ke + g.ke_off[e];
    const int32_t* rows = SCATTER ? g.topo + g.topo_ptr[e] : nullptr;
    const uint16_t* rk = SCATTER ? g.rank + g.rank_off[e] : nullptr;
    const int n_nt = (na_t + 15) / 16;   // 16-column tiles of T, W and K_e
    for (int f = 0; f < NF; ++f)
        for (int hb = f; hb < NF; ++hb) {
            for (int tile = wave; tile < 2 * n_nt; tile += 4) {
                const int mt = tile / n_nt, nt = tile - n_nt * (tile / n_nt);
                const int arow = mt * 16 + col16, bcol = nt * 16 + col16;   // a; y
                doublex4_t c4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < (NB + 3) / 4; ++ks) {
                    const int k = 4 * ks + kq;   // b
                    const double av = (arow < NB && k < NB) ? Kp[(f * NB + arow) * NP + hb * NB + k] : 0.0;
                    const double bv = (k < NB && bcol < na_t) ? Tt[k * na_t + bcol] : 0.0;
                    c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c4, 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = mt * 16 + kq + 4 * j;
                    if (r < NB && bcol < na_t) W[r * na_t + bcol] = c4[j];
                }
            }
            __syncthreads();
            for (int tile = wave; tile < n_nt * n_nt; tile += 4) {
                const int mt = tile / n_nt, nt = tile - n_nt * (tile / n_nt);
                const int xrow = mt * 16 + col16, ycol = nt * 16 + col16;   // x; y
                // scatter mode: the tile's CSR positions (and those of its mirror image in block
                // (h, f)) are requested before its MFMA work, which hides their chain of dependent
                // loads
                [[maybe_unused]] int pos[4], pos2[4];
                if constexpr (SCATTER) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int x = mt * 16 + kq + 4 * j;
                        pos[j] = pos2[j] = -1;
                        if (x < na_t && ycol < na_t) {
                            const int32_t row = rows[f * na_t + x];
                            const int32_t r0 = g.row_ptr[row], len = (g.row_ptr[row + 1] - r0) / NF;
                            pos[j] = r0 + hb * len + rk[x * na_t + ycol];
                            if (hb != f) {
                                const int32_t row2 = rows[hb * na_t + ycol];
                                const int32_t r2 = g.row_ptr[row2], len2 = (g.row_ptr[row2 + 1] - r2) / NF;
                                pos2[j] = r2 + f * len2 + rk[ycol * na_t + x];
                            }
                        }
                    }
                }
                doublex4_t c4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < (NB + 3) / 4; ++ks) {
                    const int k = 4 * ks + kq;   // a
                    const double av = (k < NB && xrow < na_t) ? Tt[k * na_t + xrow] : 0.0;
                    const double bv = (k < NB && ycol < na_t) ? W[k * na_t + ycol] : 0.0;
                    c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, c4, 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = mt * 16 + kq + 4 * j;
                    if constexpr (SCATTER) {
                        if (pos[j] >= 0) unsafeAtomicAdd(vals + pos[j], c4[j]);
                        if (pos2[j] >= 0) unsafeAtomicAdd(vals + pos2[j], c4[j]);
                    } else if (x < na_t && ycol < na_t) {
                        ke[(size_t)(f * na_t + x) * NA + hb * na_t + ycol] = c4[j];
                        if (hb != f) ke[(size_t)(hb * na_t + ycol) * NA + f * na_t + x] = c4[j];
                    }
                }
            }
            __syncthreads();
        }
}

__global__ void k_generic_zero(double* vals, int64_t n) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(0.0, vals + p);   // an output: nothing reads it back here
}

template <typename I>
__global__ void k_generic_gather(GenView g, const I* src, double* vals) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nnz) return;
    double v = 0.0;
    for (int64_t k = g.gather_ptr[p]; k < g.gather_ptr[p + 1]; ++k) v += g.ke[src[k]];
    __builtin_nontemporal_store(v, vals + p);   // an output: nothing reads it back here
}

// one wavefront per row: lanes stride over the row's values, then a butterfly reduction (a fixed
// order, so R is the same on every call)
__global__ __launch_bounds__(256) void k_generic_rhs(GenView g, const double* vals, const double* u, double* rhs) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= g.n_rows) return;
    double v = 0.0;
    for (int p = g.row_ptr[r] + lane; p < g.row_ptr[r + 1]; p += 64) v += vals[p] * u[g.col_idx[p]];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) rhs[r] = v;
}

}  // namespace

mgpu_status generic_upload(mgpu_set* S, const mgpu_set_desc& d) {
    const SetHost& h = S->h;
    GenView& g = S->g;
    g.n_el = h.n_el;
    g.N = h.n_adofs_type;
    g.n_rows = h.n_rows;
    g.nnz = (int64_t)h.col_idx.size();
    const size_t n_ig = (size_t)d.ig_ptr[h.n_el];
    const size_t n_t = (size_t)d.t_ptr[(size_t)h.n_el * h.nb];
    mgpu_status rc;
    if ((rc = gupload(S, d.ip_box, (size_t)h.n_el * 2 * h.dim, const_cast<double**>(&g.ip_box))) ||
        (rc = gupload(S, d.material, (size_t)h.n_el * h.nm, const_cast<double**>(&g.material))) ||
        (rc = gupload(S, d.ig_ptr, (size_t)h.n_el + 1, const_cast<int32_t**>(&g.ig_ptr))) ||
        (rc = gupload(S, d.ig_xi, n_ig * (h.dim + 1) * h.dim, const_cast<double**>(&g.ig_xi))) ||
        (rc = gupload(S, d.t_ptr, (size_t)h.n_el * h.nb + 1, const_cast<int32_t**>(&g.t_ptr))) ||
        (rc = gupload(S, h.t_col.data(), n_t, const_cast<int32_t**>(&g.t_col))) ||
        (rc = gupload(S, d.t_weight, n_t, const_cast<double**>(&g.t_w))) ||
        (rc = gupload(S, h.na_t.data(), h.na_t.size(), const_cast<int32_t**>(&g.na_t))) ||
        (rc = gupload(S, h.ke_off.data(), h.ke_off.size(), const_cast<int64_t**>(&g.ke_off))) ||
        (rc = gupload(S, h.gather_ptr.data(), h.gather_ptr.size(), const_cast<int64_t**>(&g.gather_ptr))) ||
        (rc = upload_gather_src(S, h.gather_src, g)) ||
        (rc = gupload(S, h.row_ptr.data(), h.row_ptr.size(), const_cast<int32_t**>(&g.row_ptr))) ||
        (rc = gupload(S, h.col_idx.data(), h.col_idx.size(), const_cast<int32_t**>(&g.col_idx))))
        return rc;
    if (h.scatter) {
        if ((rc = gupload(S, h.topo.data(), h.topo.size(), const_cast<int32_t**>(&g.topo))) ||
            (rc = gupload(S, h.topo_ptr.data(), h.topo_ptr.size(), const_cast<int64_t**>(&g.topo_ptr))) ||
            (rc = gupload(S, h.colour_el.data(), h.colour_el.size(), const_cast<int32_t**>(&g.colour_el))) ||
            (rc = gupload(S, h.rank_off.data(), h.rank_off.size(), const_cast<int64_t**>(&g.rank_off))) ||
            (rc = gupload(S, h.rank.data(), h.rank.size(), const_cast<uint16_t**>(&g.rank))))
            return rc;
        return MGPU_OK;   // no K_e pool
    }
    // the elements with most sub-cells first (stable), so that a launch does not end on one long
    // element; used when the tuning selects it
    {
        std::vector<int32_t> order((size_t)h.n_el);
        for (int32_t e = 0; e < h.n_el; ++e) order[(size_t)e] = e;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return d.ig_ptr[x + 1] - d.ig_ptr[x] > d.ig_ptr[y + 1] - d.ig_ptr[y];
        });
        if (h.n_el > 0 && (rc = gupload(S, order.data(), order.size(), const_cast<int32_t**>(&S->d_gen_order))))
            return rc;
    }
    const size_t n_ke = h.ke_off.empty() ? 0 : (size_t)h.ke_off.back();
    if (n_ke > 0) {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, n_ke * sizeof(double)));
        S->allocs.push_back(p);
        g.ke = static_cast<double*>(p);
    }
    return MGPU_OK;
}

namespace {
template <bool SCATTER, bool MFIT>
void launch_hex27_q(mgpu_set* S, bool f32, dim3 gr, int c0, double* vals) {
    const dim3 bl(256);
    if (S->h.nf == 1) {
        if (f32) hipLaunchKernelGGL((k_generic_hex27<1, float, SCATTER, MFIT>), gr, bl, 0, S->stream, S->g, c0, vals);
        else hipLaunchKernelGGL((k_generic_hex27<1, double, SCATTER, MFIT>), gr, bl, 0, S->stream, S->g, c0, vals);
    } else {
        if (f32) hipLaunchKernelGGL((k_generic_hex27<3, float, SCATTER, MFIT>), gr, bl, 0, S->stream, S->g, c0, vals);
        else hipLaunchKernelGGL((k_generic_hex27<3, double, SCATTER, MFIT>), gr, bl, 0, S->stream, S->g, c0, vals);
    }
}

template <bool MFIT>
void launch_generic_element(mgpu_set* S, bool f32, dim3 gr, dim3 bl) {
    const SetHost& h = S->h;
    if (h.dim == 3 && h.nf == 1) {
        if (f32) hipLaunchKernelGGL((k_generic_element<3, 1, float, MFIT>), gr, bl, 0, S->stream, S->g);
        else hipLaunchKernelGGL((k_generic_element<3, 1, double, MFIT>), gr, bl, 0, S->stream, S->g);
    } else if (h.dim == 3) {
        if (f32) hipLaunchKernelGGL((k_generic_element<3, 3, float, MFIT>), gr, bl, 0, S->stream, S->g);
        else hipLaunchKernelGGL((k_generic_element<3, 3, double, MFIT>), gr, bl, 0, S->stream, S->g);
    } else if (h.nf == 1) {
        if (f32) hipLaunchKernelGGL((k_generic_element<2, 1, float, MFIT>), gr, bl, 0, S->stream, S->g);
        else hipLaunchKernelGGL((k_generic_element<2, 1, double, MFIT>), gr, bl, 0, S->stream, S->g);
    } else {
        if (f32) hipLaunchKernelGGL((k_generic_element<2, 2, float, MFIT>), gr, bl, 0, S->stream, S->g);
        else hipLaunchKernelGGL((k_generic_element<2, 2, double, MFIT>), gr, bl, 0, S->stream, S->g);
    }
}

template <bool SCATTER>
void launch_hex27(mgpu_set* S, bool f32, dim3 gr, int c0, double* vals) {
    if (S->quadrature == MGPU_QUAD_MOMENT_FIT) launch_hex27_q<SCATTER, true>(S, f32, gr, c0, vals);
    else launch_hex27_q<SCATTER, false>(S, f32, gr, c0, vals);
}
}  // namespace

mgpu_status launch_generic(mgpu_set* S, const double* u, double* vals, double* rhs, hipEvent_t* ev) {
    const SetHost& h = S->h;
    if (ev)
        for (int k = 0; k < 3; ++k) HIP_TRY(hipEventRecord(ev[k], S->stream));
    if (h.n_el > 0) {
        const bool f32 = S->precision == MGPU_PREC_MIXED;
        const dim3 gr((unsigned)h.n_el), bl(64);
        if (h.nb == 27 && h.scatter) {
            // zero the values, then one launch per colour
            if (S->g.nnz > 0)
                hipLaunchKernelGGL(k_generic_zero, dim3((unsigned)std::min<int64_t>((S->g.nnz + 255) / 256, 65536)),
                                   dim3(256), 0, S->stream, vals, S->g.nnz);
            for (size_t c = 0; c + 1 < h.colour_ptr.size(); ++c) {
                const int32_t c0 = h.colour_ptr[c], cnt = h.colour_ptr[c + 1] - c0;
                if (cnt > 0) launch_hex27<true>(S, f32, dim3((unsigned)cnt), c0, vals);
            }
        } else if (h.nb == 27) {
            launch_hex27<false>(S, f32, gr, 0, vals);
        } else if (S->quadrature == MGPU_QUAD_MOMENT_FIT) {
            launch_generic_element<true>(S, f32, gr, bl);
        } else {
            launch_generic_element<false>(S, f32, gr, bl);
        }
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord(ev[3], S->stream));
    if (S->g.nnz > 0 && !h.scatter) {
        const dim3 gg((unsigned)((S->g.nnz + 255) / 256));
        if (S->g.gather_src32)
            hipLaunchKernelGGL(k_generic_gather<int32_t>, gg, dim3(256), 0, S->stream, S->g, S->g.gather_src32, vals);
        else
            hipLaunchKernelGGL(k_generic_gather<int64_t>, gg, dim3(256), 0, S->stream, S->g, S->g.gather_src, vals);
    }
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[4], S->stream));
    if (h.n_rows > 0)
        hipLaunchKernelGGL(k_generic_rhs, dim3((unsigned)((h.n_rows + 3) / 4)), dim3(256), 0, S->stream, S->g,
                           vals, u, rhs);
    HIP_TRY(hipGetLastError());
    if (ev) HIP_TRY(hipEventRecord(ev[5], S->stream));
    return MGPU_OK;
}

}  // namespace mgpu
