// lq_assemble.hip -- learned cut-cell quadrature inside the assembly.
//
// Per cut element (hex8 IP cell):
//   (MLP forward)    8 -> 64 -> 64 -> 27 on BF16 MFMA (mlp.hip), its inputs the 8 nodal level-set
//                    values over their largest magnitude: a prior for the 27 weights
//   k_lq_moments     one lane per XTK sub-tet, elements packed into wavefronts by 16-lane
//                    quarters (a 48-lane element beside a 16-lane one, two 32-lane ones, ...):
//                    the 23 moments of the material sub-cells in FP64 (TET_11 on the sub-tet,
//                    its FP64 Jacobian taken once for its 11 points, reduction through LDS)
//   k_lq_project     one wavefront per 16 elements, FP64 on MFMA (v_mfma_f64_16x16x4f64): the
//                    moment fit started from the MLP prior: the minimum-norm correction onto the moment
//                    constraints (iters = 0), optionally alternated `iters` times with the
//                    nonnegative orthant; the prior is first scaled to the element's volume.
//                    The final weights are rounded to FP32 as the element kernel uses them; the
//                    element stays learned if their residual ||V w - m|| / ||m|| <= tol and
//                    sum |w| <= kappa * volume (bounded cancellation), else it falls back to
//                    subdivision
//   k_lq_element     blocks of 4 wavefronts sharing the 27-point table, a wavefront per accepted
//                    element at a time: M = sum_q w_q g_q g_q^T on FP32 MFMA
//                    (v_mfma_f32_16x16x4f32, FP32 accumulation) over the 27 fixed points, then
//                    B^T D B in FP64, stored in vertex order
// Elements over tolerance run the subdivision kernels (assemble.hip) unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "../set/set_device.hpp"
#include "lq_view.hpp"
#include "quad.hpp"

namespace mgpu {

__global__ __launch_bounds__(64) void k_lq_moments(DevView s, LqView lv) {
    // the wavefront's elements: each owns a run of 16-lane quarters (a slot) and loops over its
    // sub-tets with the slot's width as stride
    const int lane = threadIdx.x;
    // the wavefront's record: cut element, first sub-cell and sub-cell count of each quarter
    int own[4], sub0[4], nsubq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        own[k] = lv.mom_owner[12 * blockIdx.x + k];
        sub0[k] = lv.mom_owner[12 * blockIdx.x + 4 + k];
        nsubq[k] = lv.mom_owner[12 * blockIdx.x + 8 + k];
    }
    // (selects, not indexing, so that own[] stays in registers)
    auto own_at = [&](int k) { return k == 0 ? own[0] : (k == 1 ? own[1] : (k == 2 ? own[2] : own[3])); };
    const int quarter = lane >> 4;
    const int c = own_at(quarter);
    int q0 = 0, q1 = 3;   // the slot of the lane's quarter: last start <= quarter, first end >= quarter
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (k <= quarter && own[k] != own[k - 1]) q0 = k;
#pragma unroll
    for (int k = 2; k >= 0; --k)
        if (k >= quarter && own[k] != own[k + 1]) q1 = k;
    const int width = 16 * (q1 - q0 + 1);
    double acc[kLqMoments];
#pragma unroll
    for (int k = 0; k < kLqMoments; ++k) acc[k] = 0.0;
    const int ig0 = (quarter == 0 ? sub0[0] : (quarter == 1 ? sub0[1] : (quarter == 2 ? sub0[2] : sub0[3])));
    const int nsub = c < 0 ? 0 : (quarter == 0 ? nsubq[0] : (quarter == 1 ? nsubq[1] : (quarter == 2 ? nsubq[2] : nsubq[3])));
    for (int t = lane - 16 * q0; t < nsub; t += width) {
        const double2* src2 = reinterpret_cast<const double2*>(s.ig_xi + (size_t)(ig0 + t) * 12);
        double X[12];
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const double2 v = src2[h];
            X[2 * h] = v.x;
            X[2 * h + 1] = v.y;
        }
        double J[3][3];
        const int rr[3] = {0, 1, 3};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int d = 0; d < 3; ++d) J[r][d] = X[rr[r] * 3 + d] - X[6 + d];
        const double detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
        if (detJ < kDetJLowerLimit) continue;
#pragma unroll
        for (int q = 0; q < kTet11Points; ++q) {
            // the point: vertex 2 plus the rule's coordinates along the edges J (the same affine
            // map as the vertex weights N = (xi0, xi1, 1 - xi0 - xi1 - xi2, xi2))
            double xip[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                xip[d] = X[6 + d] + kTet11Xi[q][0] * J[0][d] + kTet11Xi[q][1] * J[1][d] + kTet11Xi[q][2] * J[2][d];
            // the 23 monomials x^a y^b z^c (a + b + c <= 4, a fastest) times w*, as
            // (w* x^a) (y^b z^c): six products per point before the 23 multiply-adds
            const double wq = kTet11W[q] * detJ;
            const double wx[3] = {wq, wq * xip[0], wq * (xip[0] * xip[0])};
            const double yv[3] = {1.0, xip[1], xip[1] * xip[1]};
            const double zv[3] = {1.0, xip[2], xip[2] * xip[2]};
            int k = 0;
#pragma unroll
            for (int cc = 0; cc <= 2; ++cc)
#pragma unroll
                for (int bb = 0; bb <= 2; ++bb) {
                    const double yz = yv[bb] * zv[cc];
#pragma unroll
                    for (int aa = 0; aa <= 2; ++aa)
                        if (aa + bb + cc <= 4) acc[k++] += wx[aa] * yz;
                }
        }
    }
    // two-level reduction through LDS in three chunks of 8 moments (4 KB, so that LDS does not
    // bound the occupancy): 8 x 4 partial sums of 16 lanes, then 4 per moment
    __shared__ double red[8 * 65];
    __shared__ double part[kLqMomentsPad * 4];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        if (ch > 0) __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) red[k * 65 + lane] = (8 * ch + k < kLqMoments) ? acc[8 * ch + k] : 0.0;
        __syncthreads();
        if (lane < 32) {
            const int k = lane / 4, q = lane % 4;
            double v = 0.0;
#pragma unroll
            for (int l = 0; l < 16; ++l) v += red[k * 65 + 16 * q + l];
            part[(8 * ch + k) * 4 + q] = v;
        }
    }
    __syncthreads();
    // each slot sums its quarters in order: (slot start quarter, moment) pairs over the lanes
    for (int t = lane; t < 4 * kLqMomentsPad; t += 64) {
        const int qs = t / kLqMomentsPad, k = t % kLqMomentsPad;
        const int cs = own_at(qs);
        if (cs < 0 || (qs > 0 && own_at(qs - 1) == cs)) continue;
        double v = 0.0;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq)
            if (qq >= qs && own[qq] == cs) v += part[k * 4 + qq];
        lv.mom[(size_t)cs * kLqMomentsPad + k] = k < kLqMoments ? v : 0.0;
    }
}

// k_mfit8_weights (MGPU_QUAD_MOMENT_FIT on lattice sets): the same work list and sub-tet loop as
// k_lq_moments, and each element's 27 weights at the HEX_3x3x3 points:
//   W_g = sum_q w*_q L_i(xi_q) L_j(eta_q) L_k(zeta_q),   (i, j, k) the 1D nodes of point g,
// L the quadratic Lagrange polynomials through the rule's 1D nodes -s, 0, s (s its literal).  A
// polynomial of degree <= 2 in each axis equals its interpolant at the 27 points, so the rule
// sum_g W_g f(xi_g) is the subdivision's TET_11 sum for every such f; the trilinear stiffness
// dN_a/dx_i dN_b/dx_j is one.  The points accumulate the 27 monomial moments x^a y^b z^c (a, b,
// c <= 2) as (w* x^a)(y^b z^c); per element they become the weights through the coefficients of
// L_i = sum_a C_ia x^a: W_ijk = sum_abc C_ia C_jb C_kc m_abc.  No fit, no network, no fallback:
// flag = 1 for every element.
constexpr int kMf8 = 27;
__global__ __launch_bounds__(64) void k_mfit8_weights(DevView s, LqView lv) {
    const int lane = threadIdx.x;
    int own[4], sub0[4], nsubq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        own[k] = lv.mom_owner[12 * blockIdx.x + k];
        sub0[k] = lv.mom_owner[12 * blockIdx.x + 4 + k];
        nsubq[k] = lv.mom_owner[12 * blockIdx.x + 8 + k];
    }
    auto own_at = [&](int k) { return k == 0 ? own[0] : (k == 1 ? own[1] : (k == 2 ? own[2] : own[3])); };
    const int quarter = lane >> 4;
    const int c = own_at(quarter);
    int q0 = 0, q1 = 3;
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (k <= quarter && own[k] != own[k - 1]) q0 = k;
#pragma unroll
    for (int k = 2; k >= 0; --k)
        if (k >= quarter && own[k] != own[k + 1]) q1 = k;
    const int width = 16 * (q1 - q0 + 1);
    // 1D nodes -s, 0, s: L_0 = x (x - s) / (2 s^2), L_1 = 1 - x^2 / s^2, L_2 = x (x + s) / (2 s^2)
    const double sg = kHex3Xi[1][0];
    double acc[kMf8];
#pragma unroll
    for (int k = 0; k < kMf8; ++k) acc[k] = 0.0;
    const int ig0 = (quarter == 0 ? sub0[0] : (quarter == 1 ? sub0[1] : (quarter == 2 ? sub0[2] : sub0[3])));
    const int nsub = c < 0 ? 0 : (quarter == 0 ? nsubq[0] : (quarter == 1 ? nsubq[1] : (quarter == 2 ? nsubq[2] : nsubq[3])));
    for (int t = lane - 16 * q0; t < nsub; t += width) {
        const double2* src2 = reinterpret_cast<const double2*>(s.ig_xi + (size_t)(ig0 + t) * 12);
        double X[12];
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const double2 v = src2[h];
            X[2 * h] = v.x;
            X[2 * h + 1] = v.y;
        }
        double J[3][3];
        const int rr[3] = {0, 1, 3};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int d = 0; d < 3; ++d) J[r][d] = X[rr[r] * 3 + d] - X[6 + d];
        const double detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
        if (detJ < kDetJLowerLimit) continue;
#pragma unroll
        for (int q = 0; q < kTet11Points; ++q) {
            double xip[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                xip[d] = X[6 + d] + kTet11Xi[q][0] * J[0][d] + kTet11Xi[q][1] * J[1][d] + kTet11Xi[q][2] * J[2][d];
            const double wq = kTet11W[q] * detJ;
            const double wx[3] = {wq, wq * xip[0], wq * (xip[0] * xip[0])};
            const double yv[3] = {1.0, xip[1], xip[1] * xip[1]};
            const double zv[3] = {1.0, xip[2], xip[2] * xip[2]};
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const double yz = yv[b] * zv[c];
#pragma unroll
                    for (int a = 0; a < 3; ++a) acc[(c * 3 + b) * 3 + a] += wx[a] * yz;
                }
        }
    }
    // reduction through LDS in four chunks of 8 moments (the last padded), then per slot
    __shared__ double red[8 * 65];
    __shared__ double part[32 * 4];
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
        if (ch > 0) __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) red[k * 65 + lane] = (8 * ch + k < kMf8) ? acc[8 * ch + k] : 0.0;
        __syncthreads();
        if (lane < 32) {
            const int k = lane / 4, q = lane % 4;
            double v = 0.0;
#pragma unroll
            for (int l = 0; l < 16; ++l) v += red[k * 65 + 16 * q + l];
            part[(8 * ch + k) * 4 + q] = v;
        }
    }
    __syncthreads();
    // each slot's moments (its quarters summed in order), then its weights
    __shared__ double mtot[4 * 32];
    for (int t = lane; t < 4 * 32; t += 64) {
        const int qs = t / 32, k = t % 32;
        const int cs = own_at(qs);
        double v = 0.0;
        if (cs >= 0 && k < kMf8 && !(qs > 0 && own_at(qs - 1) == cs)) {
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
                if (qq >= qs && own[qq] == cs) v += part[k * 4 + qq];
        }
        mtot[t] = v;
    }
    __syncthreads();
    const double Cf[3][3] = {{0.0, -0.5 / sg, 0.5 / (sg * sg)}, {1.0, 0.0, -1.0 / (sg * sg)}, {0.0, 0.5 / sg, 0.5 / (sg * sg)}};
    for (int t = lane; t < 4 * 32; t += 64) {
        const int qs = t / 32, g = t % 32;
        const int cs = own_at(qs);
        if (g >= kMf8 || cs < 0 || (qs > 0 && own_at(qs - 1) == cs)) continue;
        // point g of the HEX_3x3x3 table: its 1D node per axis (-s, 0, s -> 0, 1, 2)
        int nd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double x = kHex3Xi[g][d];
            nd[d] = x < -0.5 * sg ? 0 : (x > 0.5 * sg ? 2 : 1);
        }
        double ci[3], cj[3], ck[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ci[a] = nd[0] == 0 ? Cf[0][a] : (nd[0] == 1 ? Cf[1][a] : Cf[2][a]);
            cj[a] = nd[1] == 0 ? Cf[0][a] : (nd[1] == 1 ? Cf[1][a] : Cf[2][a]);
            ck[a] = nd[2] == 0 ? Cf[0][a] : (nd[2] == 1 ? Cf[1][a] : Cf[2][a]);
        }
        const double* m = mtot + qs * 32;
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int a = 0; a < 3; ++a) v += (ci[a] * cj[b]) * ck[c] * m[(c * 3 + b) * 3 + a];
        lv.w_fit[(size_t)cs * kLqPoints + g] = (float)v;
        if (g == 0) lv.flag[cs] = 1;
    }
}

typedef double doublex4_t __attribute__((ext_vector_type(4)));

// The lane's fragments of the fit operators, loaded once per wavefront: V (A operand of the
// residual, rows = moments, k = points) and P (A operand of the correction, rows = points, k =
// moments), zero-padded (A: row lane % 16 of tile mt, k = 4 ks + lane / 16).
struct LqFrags {
    double v[2][8];
    double p[2][kLqMomentsPad / 4];
};
__device__ inline void lq_load_frags(const LqView& lv, LqFrags& fr, int lane) {
    const int col = lane % 16, grp = lane / 16;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const int row = 16 * mt + col;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int k = 4 * ks + grp;
            fr.v[mt][ks] = (row < kLqMoments && k < kLqPoints) ? lv.Vt[k * kLqMoments + row] : 0.0;
        }
#pragma unroll
        for (int ks = 0; ks < kLqMomentsPad / 4; ++ks) {
            const int k = 4 * ks + grp;
            fr.p[mt][ks] = (row < kLqPoints && k < kLqMoments) ? lv.Pt[k * kLqPoints + row] : 0.0;
        }
    }
}

// Rs = Ms - V Ws for the 16 columns of the block (rows 0..23, V zero-padded to 24 x 32), on
// v_mfma_f64_16x16x4f64: A = V (row lane % 16 of the tile, k = lane / 16), B = Ws (k = lane / 16,
// column lane % 16); result register j of a lane holds row lane / 16 + 4 j of column lane % 16.
__device__ inline void lq_residual(const LqFrags& fr, const double* Ws, const double* Ms, double* Rs, int lane) {
    const int col = lane % 16, grp = lane / 16;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        doublex4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int k = 4 * ks + grp;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fr.v[mt][ks], Ws[k * 16 + col], acc, 0, 0, 0);
        }
        for (int j = 0; j < 4; ++j) {
            const int r = 16 * mt + grp + 4 * j;
            if (r < kLqMomentsPad) Rs[r * 16 + col] = Ms[r * 16 + col] - acc[j];
        }
    }
}

// The moment fit of 16 cut elements per wavefront (columns of Ws / Ms), in FP64 on MFMA:
// prior scaled to the volume moment, w += P (m - V w) (iters + 1 times, alternated with the
// nonnegative orthant when iters > 0), weights rounded to FP32, residual of those, decision.
__global__ __launch_bounds__(64) void k_lq_project(LqView lv) {
    __shared__ double Ws[32 * 16];
    __shared__ double Ms[kLqMomentsPad * 16];
    __shared__ double Rs[kLqMomentsPad * 16];
    const int lane = threadIdx.x;
    const int col = lane % 16, grp = lane / 16;
    const int c0 = blockIdx.x * 16;
    LqFrags fr;
    lq_load_frags(lv, fr, lane);   // requested together with the moments and weights below
    for (int idx = lane; idx < kLqMomentsPad * 16; idx += 64) {
        const int r = idx / 16, ce = c0 + idx % 16;
        Ms[idx] = ce < lv.n_cut ? lv.mom[(size_t)ce * kLqMomentsPad + r] : 0.0;
    }
    for (int idx = lane; idx < 32 * 16; idx += 64) {
        const int r = idx / 16, ce = c0 + idx % 16;
        Ws[idx] = (ce < lv.n_cut && r < kLqPoints) ? 8.0 * (double)lv.w_pred[(size_t)ce * kLqPoints + r] : 0.0;
    }
    __syncthreads();
    if (lane < 16) {
        double sw = 0.0;
        for (int r = 0; r < kLqPoints; ++r) sw += Ws[r * 16 + lane];
        const double m0 = Ms[lane];
        if (sw > 0.0 && m0 > 0.0) {
            const double f = m0 / sw;
            for (int r = 0; r < kLqPoints; ++r) Ws[r * 16 + lane] *= f;
        }
    }
    __syncthreads();
    for (int it = 0; it <= lv.iters; ++it) {
        if (it > 0) {
            for (int idx = lane; idx < 32 * 16; idx += 64) Ws[idx] = fmax(Ws[idx], 0.0);
            __syncthreads();
        }
        lq_residual(fr, Ws, Ms, Rs, lane);
        __syncthreads();
        // Ws += P Rs (P zero-padded to 32 x 24)
        for (int mt = 0; mt < 2; ++mt) {
            doublex4_t acc = {0.0, 0.0, 0.0, 0.0};
            const int row = 16 * mt + col;
#pragma unroll
            for (int ks = 0; ks < kLqMomentsPad / 4; ++ks) {
                const int k = 4 * ks + grp;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(fr.p[mt][ks], Rs[k * 16 + col], acc, 0, 0, 0);
            }
            for (int j = 0; j < 4; ++j) Ws[(16 * mt + grp + 4 * j) * 16 + col] += acc[j];
        }
        __syncthreads();
    }
    // the weights as the element kernel uses them (FP32), then the residual of exactly those
    for (int idx = lane; idx < 32 * 16; idx += 64) {
        const double v = lv.iters > 0 ? fmax(Ws[idx], 0.0) : Ws[idx];
        Ws[idx] = (double)(float)v;
    }
    __syncthreads();
    lq_residual(fr, Ws, Ms, Rs, lane);
    __syncthreads();
    const int c = c0 + lane;
    if (lane < 16 && c < lv.n_cut) {
        double rk = 0.0, mk = 0.0, sa = 0.0;
        for (int r = 0; r < kLqMoments; ++r) {
            rk += Rs[r * 16 + lane] * Rs[r * 16 + lane];
            mk += Ms[r * 16 + lane] * Ms[r * 16 + lane];
        }
        for (int r = 0; r < kLqPoints; ++r) sa += fabs(Ws[r * 16 + lane]);
        const double m0 = Ms[lane];
        // accepted: moments reproduced to tol, and little cancellation among the weights
        // (sum |w| <= kappa * volume bounds the FP32 rounding of the 27-point products relative to K_e)
        const bool ok = mk > 0.0 && m0 > 0.0 && sqrt(rk / mk) <= lv.tol && sa <= lv.kappa * m0;
        lv.flag[c] = ok ? 1 : 0;
    }
    // the block's 16 x 27 weights are one contiguous range: all lanes store it in order
    const int n_el = min(16, lv.n_cut - c0);
    for (int idx = lane; idx < n_el * kLqPoints; idx += 64) {
        const int ce = idx / kLqPoints, r = idx - kLqPoints * ce;
        lv.w_fit[(size_t)c0 * kLqPoints + idx] = (float)Ws[r * 16 + ce];
    }
}

typedef float floatx4_t __attribute__((ext_vector_type(4)));

// K_e of an accepted element from its 27 weights.  NF = 1 diffusion, 3 elasticity (3D hex8).
// Blocks of kLqElemWaves wavefronts share the 27-point gradient table; each wavefront walks its own
// elements (block base + wavefront, stepping by the grid's wavefront count) with the flag, element
// id and weights of its next element loaded under the current one's arithmetic.  The block's
// iteration count is uniform, so block barriers order the LDS phases; a wavefront whose element is
// rejected or past the end idles through an iteration.  Each wavefront's matrix buffer holds M,
// then the element matrix that leaves in 16-byte vectors.
constexpr int kLqElemWaves = 4;

template <int NF, typename KT>
__global__ __launch_bounds__(64 * kLqElemWaves) void k_lq_element(DevView s, LqView lv) {
    constexpr int NB = 8, NG = 24, NA = NF * NB, NW = kLqElemWaves;
    constexpr int MBUF = (32 * 33 > NA * NA * (int)sizeof(KT) / 4) ? 32 * 33 : NA * NA * (int)sizeof(KT) / 4;
    __shared__ float gq[kLqPoints * NG];
    __shared__ float wq[NW][32];
    __shared__ __align__(16) float mbuf[NW][MBUF];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    float* Ms = mbuf[wave];
    KT* Ks = reinterpret_cast<KT*>(mbuf[wave]);
    for (int idx = tid; idx < kLqPoints * NG; idx += 64 * NW) gq[idx] = lv.g27[idx];
    const int stride = (int)gridDim.x * NW;
    int c = (int)blockIdx.x * NW + wave, f = 0, e = 0;
    float w = 0.0f;
    if (c < lv.n_cut) {
        f = lv.flag[c];
        e = s.cut_elems[c];
        if (lane < kLqPoints) w = (float)(lv.w_fit[(size_t)c * kLqPoints + lane] * lv.detJ_ip);
    }
    for (int base = (int)blockIdx.x * NW; base < lv.n_cut; base += stride) {
        const bool act = c < lv.n_cut && f != 0;   // uniform across the wavefront
        double m0 = 0.0, m1 = 0.0;
        if (act) {
            m0 = s.material[NF == 1 ? e : 2 * e];
            if (NF != 1) m1 = s.material[2 * e + 1];
        }
        const int cn = c + stride;
        int fn = 0, en = 0;
        float wn = 0.0f;
        if (cn < lv.n_cut) {
            fn = lv.flag[cn];
            en = s.cut_elems[cn];
            if (lane < kLqPoints) wn = (float)(lv.w_fit[(size_t)cn * kLqPoints + lane] * lv.detJ_ip);
        }
        __syncthreads();
        if (act && lane < 32) wq[wave][lane] = w;
        __syncthreads();
        if (act) {
            // M (24 x 24, padded 32 x 32) = sum_q (w_q g_q) g_q^T, K = 28 points in steps of 4;
            // M is symmetric: the lower-left tile is not formed (read from the upper-right one)
            for (int mt = 0; mt < 2; ++mt)
                for (int nt = mt; nt < 2; ++nt) {
                    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
                    const int mrow = mt * 16 + lane % 16, ncol = nt * 16 + lane % 16;
                    for (int ks = 0; ks < 7; ++ks) {
                        const int q = ks * 4 + lane / 16;
                        const float a = (q < kLqPoints && mrow < NG) ? wq[wave][q] * gq[q * NG + mrow] : 0.0f;
                        const float b = (q < kLqPoints && ncol < NG) ? gq[q * NG + ncol] : 0.0f;
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
                    }
                    for (int j = 0; j < 4; ++j) Ms[(mt * 16 + 4 * (lane / 16) + j) * 33 + ncol] = acc[j];
                }
        }
        __syncthreads();
        // B^T D B in FP64 from M: lane (a, b) owns the NF x NF block (a, b)
        auto m_at = [&](int r, int c2) { return (r >= 16 && c2 < 16) ? Ms[c2 * 33 + r] : Ms[r * 33 + c2]; };
        const int a = lane / NB, b = lane % NB;
        KT kv[NF * NF];
        if (act) {
            if (NF == 1) {
                double v = 0.0;
                for (int i = 0; i < 3; ++i) v += (double)m_at(i * NB + a, i * NB + b);
                kv[0] = (KT)(m0 * v);
            } else {
                double D[6][6];
                elastic_D<3>(m0, m1, D);
                for (int al = 0; al < 3; ++al)
                    for (int be = 0; be < 3; ++be) {
                        double v = 0.0;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j)
                                if (d_nonzero<3>(voigt<3>(al, i), voigt<3>(be, j)))
                                    v += D[voigt<3>(al, i)][voigt<3>(be, j)] * (double)m_at(i * NB + a, j * NB + b);
                        kv[al * 3 + be] = (KT)v;
                    }
            }
        }
        __syncthreads();
        if (act) {
            if (NF == 1) {
                Ks[a * NB + b] = kv[0];
            } else {
                for (int al = 0; al < 3; ++al)
                    for (int be = 0; be < 3; ++be) Ks[(a * 3 + al) * NA + (be * NB + b)] = kv[al * 3 + be];
            }
        }
        __syncthreads();
        if (act) {
            // the element matrix leaves in 16-byte vectors (NA * NA * sizeof(KT) is a multiple of 16)
            using V4 = typename std::conditional<sizeof(KT) == 4, float4, double2>::type;
            constexpr int NV = NA * NA * (int)sizeof(KT) / 16;
            const V4* src = reinterpret_cast<const V4*>(Ks);
            V4* dst = reinterpret_cast<V4*>(ke_store<KT>(s) + (size_t)c * NA * NA);
            for (int t = lane; t < NV; t += 64) dst[t] = src[t];
        }
        c = cn;
        f = fn;
        e = en;
        w = wn;
    }
}

// Launch the learned part of an assembly on stream st (after the MLP forward has written w_pred).
void lq_launch_fit(const DevView& s, const LqView& lv, hipStream_t st) {
    if (lv.n_cut > 0) {
        hipLaunchKernelGGL(k_lq_moments, dim3(lv.n_mom_waves), dim3(64), 0, st, s, lv);
        hipLaunchKernelGGL(k_lq_project, dim3((lv.n_cut + 15) / 16), dim3(64), 0, st, lv);
    }
}
void lq_launch_mfit8(const DevView& s, const LqView& lv, hipStream_t st) {
    if (lv.n_cut > 0) hipLaunchKernelGGL(k_mfit8_weights, dim3(lv.n_mom_waves), dim3(64), 0, st, s, lv);
}
void lq_launch_element(const DevView& s, const LqView& lv, int nf, bool f32_store, hipStream_t st) {
    if (lv.n_cut == 0) return;
    const dim3 grid((unsigned)std::min((lv.n_cut + kLqElemWaves - 1) / kLqElemWaves, 2048)), block(64 * kLqElemWaves);
    if (f32_store) {
        if (nf == 1) hipLaunchKernelGGL((k_lq_element<1, float>), grid, block, 0, st, s, lv);
        else hipLaunchKernelGGL((k_lq_element<3, float>), grid, block, 0, st, s, lv);
    } else {
        if (nf == 1) hipLaunchKernelGGL((k_lq_element<1, double>), grid, block, 0, st, s, lv);
        else hipLaunchKernelGGL((k_lq_element<3, double>), grid, block, 0, st, s, lv);
    }
}

}  // namespace mgpu
