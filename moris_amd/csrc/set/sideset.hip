// sideset.hip -- side-set IWGs (Nitsche Dirichlet, Neumann) on the facets of a bulk set's IG cells,
// added into the bulk set's CSR values and residual.  Target: gfx950 (MI355X).
//
//   Element_Sideset::compute_jacobian_and_residual  projects/FEM/INT/src/ELEM/cl_FEM_Element_Sideset.cpp:245-330
//   IWG_Isotropic_Struc_Linear_Dirichlet            projects/FEM/INT/src/IWG/cl_FEM_IWG_Isotropic_Struc_Linear_Dirichlet.cpp:52-240
//   IWG_Diffusion_Dirichlet_Nitsche                 projects/FEM/INT/src/IWG/cl_FEM_IWG_Diffusion_Dirichlet_Nitsche.cpp
//   IWG_*_Neumann                                   projects/FEM/INT/src/IWG/cl_FEM_IWG_Isotropic_Struc_Linear_Neumann.cpp:38-85
//   SP_Dirichlet_Nitsche (gamma = alpha mat / h)     projects/FEM/INT/src/SP/cl_FEM_SP_Dirichlet_Nitsche.cpp:75-115
//   side cluster length measure                     projects/FEM/INT/src/ELEM/cl_FEM_Cluster.cpp:858-919
//
// k_side: one wavefront per facet.  For every point of the side rule (TRI_7 / QUAD_2x2 / BAR_3) the
// point is mapped into the IP cell, w* = w detJ_side, and lane (a, b) adds the (a, b) vertex block
// of the facet matrix
//     K[(f,a),(g,b)] += w* ( -N_a T_f(g,b) + beta T_g(f,a) N_b + gamma delta_fg N_a N_b )
// where T_f(g,b), the f-th traction component of the unit displacement g at vertex b, is the
// isotropic closed form  lambda n_f d_g N_b + mu n_g d_f N_b + mu delta_fg (n . grad N_b)  of
// n_flat D B (diffusion: k n . grad N_b).  R = K u - w* (beta T^T + gamma N^T) u_D for Dirichlet,
// R = -w* N^T t_N for Neumann.  The facet matrices go to a pool; k_side_gather adds each CSR value's
// (and each row's) contributions in ascending facet order, a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../../include/moris_gpu_sideset.h"
#include "../common/fem_tables.hpp"
#include "set_state.hpp"

namespace mgpu {

struct SideView {
    int32_t n_facets, N;
    double beta, alpha, value[3];
    int32_t neumann;
    const double* xi;        // [n][nvs][dim] facet vertices in the IP cell
    const double* normal;    // [n][dim]
    const double* hlen;      // [n] length measure of the facet's side cluster
    const double* box;       // [n][2 dim] IP box of the facet's element (lower corner, extent)
    const double* mat;       // [n][2] material of the facet's element (k | E, nu)
    const int32_t* vadof;    // [n][nb] type-local adof of each IP-cell vertex
    double* ke;              // [n][na][na] facet matrices, pdof order (field-major)
    double* re;              // [n][na]
};

template <int DIM, int NF, int GEOM>
__global__ __launch_bounds__(64) void k_side(SideView s, const double* u) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NA = NF * NB;
    constexpr int NVS = GEOM == MGPU_GEOM_TRI3 ? 3 : (GEOM == MGPU_GEOM_QUAD4 ? 4 : 2);
    constexpr int NQ = GEOM == MGPU_GEOM_TRI3 ? kTri7Points : (GEOM == MGPU_GEOM_QUAD4 ? kQuad2Points : kBar3Points);
    __shared__ double Ks[NA * NA];
    __shared__ double ue[NA], rd[NA];
    const int fct = blockIdx.x;
    const int lane = threadIdx.x;
    double X[NVS][DIM], n[DIM], lo[DIM], ext[DIM];
    for (int k = 0; k < NVS; ++k)
        for (int d = 0; d < DIM; ++d) X[k][d] = s.xi[((size_t)fct * NVS + k) * DIM + d];
    for (int d = 0; d < DIM; ++d) {
        n[d] = s.normal[(size_t)fct * DIM + d];
        lo[d] = s.box[(size_t)fct * 2 * DIM + d];
        ext[d] = s.box[(size_t)fct * 2 * DIM + DIM + d];
    }
    const double m0 = s.mat[2 * (size_t)fct], m1 = s.mat[2 * (size_t)fct + 1];
    double lam = 0.0, mu = 0.0;
    if constexpr (NF > 1) {
        constexpr int NS = DIM == 3 ? 6 : 3;
        double D[NS][NS];
        elastic_D<DIM>(m0, m1, D);
        lam = D[0][1];
        mu = D[NS - 1][NS - 1];
    }
    const double gamma = s.alpha * m0 / s.hlen[fct];
    const double beta = s.beta;
    const int a = lane / NB, b = lane % NB;
    const bool act = lane < NB * NB;
    double K[NF][NF];
    double r[NF];
    for (int f = 0; f < NF; ++f) {
        r[f] = 0.0;
        for (int g = 0; g < NF; ++g) K[f][g] = 0.0;
    }
    for (int q = 0; q < NQ; ++q) {
        // point in the IP cell and the side Jacobian (physical)
        double phi[NVS], xi[DIM], detJ;
        if constexpr (GEOM == MGPU_GEOM_TRI3) {
            phi[0] = kTri7Xi[q][0];
            phi[1] = kTri7Xi[q][1];
            phi[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
            double e1[3], e2[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = (X[0][d] - X[2][d]) * 0.5 * ext[d];
                e2[d] = (X[1][d] - X[2][d]) * 0.5 * ext[d];
            }
            const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2],
                         c2 = e1[0] * e2[1] - e1[1] * e2[0];
            detJ = sqrt(c0 * c0 + c1 * c1 + c2 * c2) / 2.0;
        } else if constexpr (GEOM == MGPU_GEOM_QUAD4) {
            const double ss = kQuad2Xi[q][0], tt = kQuad2Xi[q][1];
            phi[0] = 0.25 * (1.0 - ss) * (1.0 - tt);
            phi[1] = 0.25 * (1.0 + ss) * (1.0 - tt);
            phi[2] = 0.25 * (1.0 + ss) * (1.0 + tt);
            phi[3] = 0.25 * (1.0 - ss) * (1.0 + tt);
            double js[3], jt[3];
            for (int d = 0; d < 3; ++d) {
                js[d] = 0.25 * (-(1.0 - tt) * X[0][d] + (1.0 - tt) * X[1][d] + (1.0 + tt) * X[2][d] - (1.0 + tt) * X[3][d]) *
                        0.5 * ext[d];
                jt[d] = 0.25 * (-(1.0 - ss) * X[0][d] - (1.0 + ss) * X[1][d] + (1.0 + ss) * X[2][d] + (1.0 - ss) * X[3][d]) *
                        0.5 * ext[d];
            }
            const double c0 = js[1] * jt[2] - js[2] * jt[1], c1 = js[2] * jt[0] - js[0] * jt[2],
                         c2 = js[0] * jt[1] - js[1] * jt[0];
            detJ = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        } else {
            const double ss = kBar3Xi[q];
            phi[0] = 0.5 * (1.0 - ss);
            phi[1] = 0.5 * (1.0 + ss);
            double l2 = 0.0;
            for (int d = 0; d < DIM; ++d) {
                const double t = (X[1][d] - X[0][d]) * 0.5 * ext[d];
                l2 += t * t;
            }
            detJ = 0.5 * sqrt(l2);
        }
        const double w = (GEOM == MGPU_GEOM_TRI3 ? kTri7W[q] : (GEOM == MGPU_GEOM_QUAD4 ? kQuad2W[q] : kBar3W[q])) * detJ;
        for (int d = 0; d < DIM; ++d) {
            double v = 0.0;
            for (int k = 0; k < NVS; ++k) v += phi[k] * X[k][d];
            xi[d] = v;
        }
        // N and physical gradients of the IP cell's vertices a and b
        auto shape = [&](int v, double& N, double* g) {
            double f[DIM];
            for (int d = 0; d < DIM; ++d)
                f[d] = 1.0 + (DIM == 3 ? kHex8XiHat[v][d] : kQuad4XiHat[v][d]) * xi[d];
            N = 1.0;
            for (int d = 0; d < DIM; ++d) N *= f[d];
            N /= (double)(1 << DIM);
            for (int i = 0; i < DIM; ++i) {
                double p = (DIM == 3 ? kHex8XiHat[v][i] : kQuad4XiHat[v][i]) / (double)(1 << DIM);
                for (int d = 0; d < DIM; ++d)
                    if (d != i) p *= f[d];
                g[i] = p * 2.0 / ext[i];
            }
        };
        if (!act) continue;
        double Na, Nb, ga[DIM], gb[DIM];
        shape(a, Na, ga);
        shape(b, Nb, gb);
        double nga = 0.0, ngb = 0.0;
        for (int d = 0; d < DIM; ++d) {
            nga += n[d] * ga[d];
            ngb += n[d] * gb[d];
        }
        if (NF == 1) {
            const double k = m0;
            K[0][0] += w * (-Na * k * ngb + beta * k * nga * Nb + gamma * Na * Nb);
            if (b == 0) r[0] += s.neumann ? w * Na * s.value[0] : w * (beta * k * nga + gamma * Na) * s.value[0];
        } else {
            for (int f = 0; f < NF; ++f)
                for (int g = 0; g < NF; ++g) {
                    const double Tfgb = lam * n[f] * gb[g] + mu * n[g] * gb[f] + (f == g ? mu * ngb : 0.0);
                    const double Tgfa = lam * n[g] * ga[f] + mu * n[f] * ga[g] + (f == g ? mu * nga : 0.0);
                    K[f][g] += w * (-Na * Tfgb + beta * Tgfa * Nb + (f == g ? gamma * Na * Nb : 0.0));
                }
            if (b == 0)
                for (int f = 0; f < NF; ++f) {
                    if (s.neumann) {
                        r[f] += w * Na * s.value[f];
                    } else {
                        double t = gamma * Na * s.value[f];
                        for (int g = 0; g < NF; ++g) {
                            const double Tgfa = lam * n[g] * ga[f] + mu * n[f] * ga[g] + (f == g ? mu * nga : 0.0);
                            t += beta * Tgfa * s.value[g];
                        }
                        r[f] += w * t;
                    }
                }
        }
    }
    if (act) {
        for (int f = 0; f < NF; ++f) {
            for (int g = 0; g < NF; ++g) Ks[(f * NB + a) * NA + g * NB + b] = s.neumann ? 0.0 : K[f][g];
            if (b == 0) rd[f * NB + a] = r[f];
        }
    }
    if (lane < NA) {
        const int f = lane / NB, v = lane % NB;
        ue[lane] = u[(size_t)f * s.N + s.vadof[(size_t)fct * NB + v]];
    }
    __syncthreads();
    double* ke = s.ke + (size_t)fct * NA * NA;
    for (int t = lane; t < NA * NA; t += 64) ke[t] = Ks[t];
    if (lane < NA) {
        double v = 0.0;
        if (!s.neumann)
            for (int c = 0; c < NA; ++c) v += Ks[lane * NA + c] * ue[c];
        s.re[(size_t)fct * NA + lane] = v - rd[lane];
    }
}

// ------------------------------------------------------------------------------------------------
// Side sets on generic-path sets (T rows with several entries: p = 2 B-splines on HEX8 / QUAD4
// cells, HEX27 cells, THB rows).  One block per side cluster (the facets of one element): the
// facet matrix K_p over the IP cell's NB bases is summed over the cluster's facets and points,
// then projected, K_e = T^T K_p T and R_e = T^T (K_p T u_e - r_p), as
// Equation_Object::get_egn_obj_jacobian_and_residual (cl_MSI_Equation_Object.cpp:789-842) does for
// the reference's side clusters.  HEX27 bases are the tensor products of the 1D quadratic
// Lagrange polynomials (kHex27Half node order).
// ------------------------------------------------------------------------------------------------
constexpr int kSideMaxNat = 64;   // adofs per dof type of a side cluster's element

struct SideGenView {
    int32_t n_clusters, N, nm;
    double beta, alpha, value[3];
    int32_t neumann;
    const int32_t* cl_elem;    // [n_cl] element of each side cluster
    const int32_t* cl_fptr;    // [n_cl + 1] its facets
    const double* xi;          // [n facets][nvs][dim], in cluster order
    const double* normal;      // [n facets][dim]
    const double* hlen;        // [n_cl] length measure of the cluster
    const int64_t* cl_aptr;    // [n_cl + 1] into cl_adof
    const int32_t* cl_adof;    // the cluster element's adofs (global rows, types concatenated)
    const int64_t* ke_off;     // [n_cl + 1] pool offsets, (nf na_t)^2 per cluster
    const int64_t* re_off;     // [n_cl + 1] nf na_t per cluster
    double* ke;
    double* re;
};

template <int DIM, int NB>
__device__ inline void side_basis(int a, const double* xi, const double* ext, double& N, double* g) {
    if constexpr (NB == 27) {
        double l[3][3], dl[3][3];
        for (int d = 0; d < 3; ++d) lagrange_quadratic_1d(xi[d], l[d], dl[d]);
        const int i0 = kHex27Half[a][0], i1 = kHex27Half[a][1], i2 = kHex27Half[a][2];
        N = l[0][i0] * l[1][i1] * l[2][i2];
        g[0] = dl[0][i0] * l[1][i1] * l[2][i2] * 2.0 / ext[0];
        g[1] = l[0][i0] * dl[1][i1] * l[2][i2] * 2.0 / ext[1];
        g[2] = l[0][i0] * l[1][i1] * dl[2][i2] * 2.0 / ext[2];
    } else {
        double f[DIM];
        for (int d = 0; d < DIM; ++d) f[d] = 1.0 + (DIM == 3 ? kHex8XiHat[a][d] : kQuad4XiHat[a][d]) * xi[d];
        N = 1.0;
        for (int d = 0; d < DIM; ++d) N *= f[d];
        N /= (double)(1 << DIM);
        for (int i = 0; i < DIM; ++i) {
            double p = (DIM == 3 ? kHex8XiHat[a][i] : kQuad4XiHat[a][i]) / (double)(1 << DIM);
            for (int d = 0; d < DIM; ++d)
                if (d != i) p *= f[d];
            g[i] = p * 2.0 / ext[i];
        }
    }
}

template <int DIM, int NF, int GEOM, int NB>
__global__ __launch_bounds__(256) void k_side_gen(SideGenView s, GenView gv, const double* u) {
    constexpr int NA = NF * NB;
    constexpr int NVS = GEOM == MGPU_GEOM_TRI3 ? 3 : (GEOM == MGPU_GEOM_QUAD4 ? 4 : 2);
    constexpr int NQ = GEOM == MGPU_GEOM_TRI3 ? kTri7Points : (GEOM == MGPU_GEOM_QUAD4 ? kQuad2Points : kBar3Points);
    constexpr int M = kSideMaxNat;
    __shared__ double Kp[NA * NA];
    __shared__ double rd[NA];
    __shared__ double Nq[NB], Gq[NB * DIM];
    __shared__ double Tm[NB * M];
    __shared__ double W[NB * M];
    __shared__ double ue[NF * M], up[NA];
    __shared__ double pt[DIM + 1];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int e = s.cl_elem[c];
    double lo[DIM], ext[DIM];
    for (int d = 0; d < DIM; ++d) {
        lo[d] = gv.ip_box[(size_t)e * 2 * DIM + d];
        ext[d] = gv.ip_box[(size_t)e * 2 * DIM + DIM + d];
    }
    const double m0 = gv.material[(size_t)e * s.nm], m1 = s.nm > 1 ? gv.material[(size_t)e * s.nm + 1] : 0.0;
    double lam = 0.0, mu = 0.0;
    if constexpr (NF > 1) {
        constexpr int NS = DIM == 3 ? 6 : 3;
        double D[NS][NS];
        elastic_D<DIM>(m0, m1, D);
        lam = D[0][1];
        mu = D[NS - 1][NS - 1];
    }
    const double gamma = s.neumann ? 0.0 : s.alpha * m0 / s.hlen[c];
    const double beta = s.beta;
    for (int t = tid; t < NA * NA; t += 256) Kp[t] = 0.0;
    for (int t = tid; t < NA; t += 256) rd[t] = 0.0;
    for (int fct = s.cl_fptr[c]; fct < s.cl_fptr[c + 1]; ++fct) {
        double X[NVS][DIM], n[DIM];
        for (int k = 0; k < NVS; ++k)
            for (int d = 0; d < DIM; ++d) X[k][d] = s.xi[((size_t)fct * NVS + k) * DIM + d];
        for (int d = 0; d < DIM; ++d) n[d] = s.normal[(size_t)fct * DIM + d];
        for (int q = 0; q < NQ; ++q) {
            __syncthreads();   // the previous point's N and gradients are no longer read
            if (tid == 0) {
                double phi[NVS], detJ;
                if constexpr (GEOM == MGPU_GEOM_TRI3) {
                    phi[0] = kTri7Xi[q][0];
                    phi[1] = kTri7Xi[q][1];
                    phi[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
                    double e1[3], e2[3];
                    for (int d = 0; d < 3; ++d) {
                        e1[d] = (X[0][d] - X[2][d]) * 0.5 * ext[d];
                        e2[d] = (X[1][d] - X[2][d]) * 0.5 * ext[d];
                    }
                    const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2],
                                 c2 = e1[0] * e2[1] - e1[1] * e2[0];
                    detJ = sqrt(c0 * c0 + c1 * c1 + c2 * c2) / 2.0;
                } else if constexpr (GEOM == MGPU_GEOM_QUAD4) {
                    const double ss = kQuad2Xi[q][0], tt = kQuad2Xi[q][1];
                    phi[0] = 0.25 * (1.0 - ss) * (1.0 - tt);
                    phi[1] = 0.25 * (1.0 + ss) * (1.0 - tt);
                    phi[2] = 0.25 * (1.0 + ss) * (1.0 + tt);
                    phi[3] = 0.25 * (1.0 - ss) * (1.0 + tt);
                    double js[3], jt[3];
                    for (int d = 0; d < 3; ++d) {
                        js[d] = 0.25 * (-(1.0 - tt) * X[0][d] + (1.0 - tt) * X[1][d] + (1.0 + tt) * X[2][d] -
                                        (1.0 + tt) * X[3][d]) * 0.5 * ext[d];
                        jt[d] = 0.25 * (-(1.0 - ss) * X[0][d] - (1.0 + ss) * X[1][d] + (1.0 + ss) * X[2][d] +
                                        (1.0 - ss) * X[3][d]) * 0.5 * ext[d];
                    }
                    const double c0 = js[1] * jt[2] - js[2] * jt[1], c1 = js[2] * jt[0] - js[0] * jt[2],
                                 c2 = js[0] * jt[1] - js[1] * jt[0];
                    detJ = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
                } else {
                    const double ss = kBar3Xi[q];
                    phi[0] = 0.5 * (1.0 - ss);
                    phi[1] = 0.5 * (1.0 + ss);
                    double l2 = 0.0;
                    for (int d = 0; d < DIM; ++d) {
                        const double t = (X[1][d] - X[0][d]) * 0.5 * ext[d];
                        l2 += t * t;
                    }
                    detJ = 0.5 * sqrt(l2);
                }
                for (int d = 0; d < DIM; ++d) {
                    double v = 0.0;
                    for (int k = 0; k < NVS; ++k) v += phi[k] * X[k][d];
                    pt[d] = v;
                }
                pt[DIM] = (GEOM == MGPU_GEOM_TRI3 ? kTri7W[q] : (GEOM == MGPU_GEOM_QUAD4 ? kQuad2W[q] : kBar3W[q])) * detJ;
            }
            __syncthreads();
            if (tid < NB) {
                double xi[DIM], g[DIM], N;
                for (int d = 0; d < DIM; ++d) xi[d] = pt[d];
                side_basis<DIM, NB>(tid, xi, ext, N, g);
                Nq[tid] = N;
                for (int d = 0; d < DIM; ++d) Gq[tid * DIM + d] = g[d];
            }
            __syncthreads();
            const double w = pt[DIM];
            for (int pr = tid; pr < NB * NB; pr += 256) {
                const int a = pr / NB, b = pr % NB;
                const double Na = Nq[a], Nb = Nq[b];
                const double* ga = Gq + a * DIM;
                const double* gb = Gq + b * DIM;
                double nga = 0.0, ngb = 0.0;
                for (int d = 0; d < DIM; ++d) {
                    nga += n[d] * ga[d];
                    ngb += n[d] * gb[d];
                }
                if (NF == 1) {
                    const double k = m0;
                    if (!s.neumann) Kp[a * NA + b] += w * (-Na * k * ngb + beta * k * nga * Nb + gamma * Na * Nb);
                    if (b == 0) rd[a] += s.neumann ? w * Na * s.value[0] : w * (beta * k * nga + gamma * Na) * s.value[0];
                } else {
                    if (!s.neumann)
                        for (int f = 0; f < NF; ++f)
                            for (int g2 = 0; g2 < NF; ++g2) {
                                const double Tfgb = lam * n[f] * gb[g2] + mu * n[g2] * gb[f] + (f == g2 ? mu * ngb : 0.0);
                                const double Tgfa = lam * n[g2] * ga[f] + mu * n[f] * ga[g2] + (f == g2 ? mu * nga : 0.0);
                                Kp[(f * NB + a) * NA + g2 * NB + b] +=
                                    w * (-Na * Tfgb + beta * Tgfa * Nb + (f == g2 ? gamma * Na * Nb : 0.0));
                            }
                    if (b == 0)
                        for (int f = 0; f < NF; ++f) {
                            double t;
                            if (s.neumann) {
                                t = Na * s.value[f];
                            } else {
                                t = gamma * Na * s.value[f];
                                for (int g2 = 0; g2 < NF; ++g2) {
                                    const double Tgfa = lam * n[g2] * ga[f] + mu * n[f] * ga[g2] + (f == g2 ? mu * nga : 0.0);
                                    t += beta * Tgfa * s.value[g2];
                                }
                            }
                            rd[f * NB + a] += w * t;
                        }
                }
            }
        }
    }
    // the element's T (NB x na, element-local columns) and its adof values
    const int na = gv.na_t[e];
    for (int t = tid; t < NB * M; t += 256) Tm[t] = 0.0;
    __syncthreads();
    if (tid < NB)
        for (int k = gv.t_ptr[(size_t)e * NB + tid]; k < gv.t_ptr[(size_t)e * NB + tid + 1]; ++k)
            Tm[tid * M + gv.t_col[k]] += gv.t_w[k];
    const int64_t a0 = s.cl_aptr[c];
    for (int t = tid; t < NF * na; t += 256) ue[t] = u[s.cl_adof[a0 + t]];
    __syncthreads();
    // u_p = T u_e per field, then r_p = K_p u_p - rd (into up, after a barrier)
    double upv = 0.0;
    if (tid < NA) {
        const int f = tid / NB, b = tid % NB;
        for (int x = 0; x < na; ++x) upv += Tm[b * M + x] * ue[f * na + x];
    }
    __syncthreads();
    if (tid < NA) up[tid] = upv;
    __syncthreads();
    double rp = 0.0;
    if (tid < NA) {
        if (!s.neumann)
            for (int j = 0; j < NA; ++j) rp += Kp[tid * NA + j] * up[j];
        rp -= rd[tid];
    }
    __syncthreads();
    if (tid < NA) up[tid] = rp;
    __syncthreads();
    double* re = s.re + s.re_off[c];
    for (int t = tid; t < NF * na; t += 256) {
        const int f = t / na, x = t % na;
        double v = 0.0;
        for (int b = 0; b < NB; ++b) v += Tm[b * M + x] * up[f * NB + b];
        re[t] = v;
    }
    if (s.neumann) return;
    // K_e = T^T K_p T, one (f, g) block at a time through W = K_p(f, g) T
    double* ke = s.ke + s.ke_off[c];
    const int nat = NF * na;
    for (int f = 0; f < NF; ++f)
        for (int g2 = 0; g2 < NF; ++g2) {
            __syncthreads();
            for (int t = tid; t < NB * na; t += 256) {
                const int b = t / na, y = t % na;
                double v = 0.0;
                for (int b2 = 0; b2 < NB; ++b2) v += Kp[(f * NB + b) * NA + g2 * NB + b2] * Tm[b2 * M + y];
                W[b * M + y] = v;
            }
            __syncthreads();
            for (int t = tid; t < na * na; t += 256) {
                const int x = t / na, y = t % na;
                double v = 0.0;
                for (int b = 0; b < NB; ++b) v += Tm[b * M + x] * W[b * M + y];
                ke[(size_t)(f * na + x) * nat + g2 * na + y] = v;
            }
        }
}

// out[dst[t]] += sum of pool[src[k]], k in [ptr[t], ptr[t+1]) (one thread per target)
__global__ __launch_bounds__(256) void k_side_gather(int64_t n, const int64_t* ptr, const int64_t* dst,
                                                     const int64_t* src, const double* pool, double* out) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double v = 0.0;
    for (int64_t k = ptr[t]; k < ptr[t + 1]; ++k) v += pool[src[k]];
    out[dst[t]] += v;
}

// ------------------------------------------------------------------------------------------------
// Ghost facets (double-sided): one block per facet, the leader's and the follower's NB bases side by
// side (2 NB columns per dof type), K_p in LDS from the face points (QUAD_2x2 / BAR_3 on the shared
// face), then both elements' T rows against the union of their adofs.
//   K_p[(f,X,a),(g,Y,b)] += w* gamma s_X s_Y t_{X a f} . T_{Y b g},   s_L = 1, s_F = -1,
//   t_{X a f} = N_n eps(N_{X a} e_f)         (test traction, no material),
//   T_{Y b g} = N_n D_Y eps(N_{Y b} e_g)     (traction of side Y; diffusion: t = n . grad N,
//                                              T = k_Y n . grad N)
// ------------------------------------------------------------------------------------------------
struct GhostView {
    int32_t n_facets, N, nm;
    const int32_t* leader;
    const int32_t* follower;
    const int32_t* axis;       // [n] axis of the face normal
    const double* side;        // [n] sign of the leader's outward normal along it
    const double* gamma;       // [n] alpha h
    const int64_t* cl_aptr;    // [n + 1] into cl_adof
    const int32_t* cl_adof;    // union of the two elements' adofs (global rows, types concatenated)
    const int64_t* map_ptr;    // [n + 1] into map
    const int32_t* map;        // element-local column -> union column: leader's, then follower's
    const int64_t* ke_off;
    const int64_t* re_off;
    double* ke;
    double* re;
};

template <int DIM, int NF>
__global__ __launch_bounds__(256) void k_ghost(GhostView s, GenView gv, const double* u) {
    constexpr int NB = DIM == 3 ? 8 : 4;
    constexpr int NB2 = 2 * NB;
    constexpr int NA2 = NF * NB2;
    constexpr int NQ = DIM == 3 ? kQuad2Points : kBar3Points;
    constexpr int NS = DIM == 3 ? 6 : 3;
    constexpr int M = kSideMaxNat;
    __shared__ double Kp[NA2 * NA2];
    __shared__ double Gq[NB2 * DIM];
    __shared__ double Tm[NB2 * M];
    __shared__ double W[NB2 * M];
    __shared__ double ue[NF * M], up[NA2];
    __shared__ double Dm[2][NS * NS];
    __shared__ double km[2];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int eX[2] = {s.leader[c], s.follower[c]};
    const int ax = s.axis[c];
    const double sd = s.side[c];
    double ext[DIM];
    for (int d = 0; d < DIM; ++d) ext[d] = gv.ip_box[(size_t)eX[0] * 2 * DIM + DIM + d];
    double n[DIM];
    for (int d = 0; d < DIM; ++d) n[d] = d == ax ? sd : 0.0;
    if (tid < 2) {
        const int e = eX[tid];
        if constexpr (NF == 1) {
            km[tid] = gv.material[(size_t)e * s.nm];
        } else {
            double D[NS][NS];
            elastic_D<DIM>(gv.material[(size_t)e * s.nm], gv.material[(size_t)e * s.nm + 1], D);
            for (int i = 0; i < NS; ++i)
                for (int j = 0; j < NS; ++j) Dm[tid][i * NS + j] = D[i][j];
        }
    }
    for (int t = tid; t < NA2 * NA2; t += 256) Kp[t] = 0.0;
    // the face's other axes (ascending) and the point's weight w * detJ of the face
    int oa[2] = {0, 0}, no = 0;
    for (int d = 0; d < DIM; ++d)
        if (d != ax) oa[no++] = d;
    double face_det = 1.0;
    for (int k = 0; k < DIM - 1; ++k) face_det *= 0.5 * ext[oa[k]];
    const double gamma = s.gamma[c];
    for (int q = 0; q < NQ; ++q) {
        __syncthreads();   // the previous point's gradients are no longer read (and Dm / km are set)
        double w;
        double xi[DIM];
        if constexpr (DIM == 3) {
            xi[oa[0]] = kQuad2Xi[q][0];
            xi[oa[1]] = kQuad2Xi[q][1];
            w = kQuad2W[q] * face_det;
        } else {
            xi[oa[0]] = kBar3Xi[q];
            w = kBar3W[q] * face_det;
        }
        if (tid < NB2) {
            const int X = tid / NB, a = tid % NB;
            double x[DIM], g[DIM], N;
            for (int d = 0; d < DIM; ++d) x[d] = xi[d];
            x[ax] = X == 0 ? sd : -sd;   // the leader's face, and the same face seen from the follower
            side_basis<DIM, NB>(a, x, ext, N, g);
            for (int d = 0; d < DIM; ++d) Gq[tid * DIM + d] = g[d];
        }
        __syncthreads();
        for (int pr = tid; pr < NB2 * NB2; pr += 256) {
            const int i = pr / NB2, j = pr % NB2;   // (X, a), (Y, b)
            const int X = i / NB, Y = j / NB;
            const double sg = (X == Y ? 1.0 : -1.0) * w * gamma;
            const double* ga = Gq + i * DIM;
            const double* gb = Gq + j * DIM;
            if constexpr (NF == 1) {
                double na = 0.0, nb = 0.0;
                for (int d = 0; d < DIM; ++d) {
                    na += n[d] * ga[d];
                    nb += n[d] * gb[d];
                }
                Kp[i * NA2 + j] += sg * na * km[Y] * nb;
            } else {
                // t (no material) of (X, a, f) and T (side Y's material) of (Y, b, g)
                double t[NF][DIM], T[NF][DIM];
                for (int f = 0; f < NF; ++f) {
                    double ea[NS], eb[NS];
                    for (int k = 0; k < NS; ++k) ea[k] = eb[k] = 0.0;
                    // engineering strain of N e_f: [xx, yy, zz, yz, xz, xy] (2D: [xx, yy, xy])
                    if constexpr (DIM == 3) {
                        const int sh1 = f == 0 ? 5 : (f == 1 ? 5 : 4), sh2 = f == 0 ? 4 : 3;
                        const int o1 = f == 0 ? 1 : 0, o2 = f == 2 ? 1 : 2;
                        ea[f] = ga[f];
                        eb[f] = gb[f];
                        ea[sh1] = ga[o1];
                        eb[sh1] = gb[o1];
                        ea[sh2] = ga[o2];
                        eb[sh2] = gb[o2];
                    } else {
                        ea[f] = ga[f];
                        eb[f] = gb[f];
                        ea[2] = ga[1 - f];
                        eb[2] = gb[1 - f];
                    }
                    double sb[NS];
                    for (int r = 0; r < NS; ++r) {
                        double v = 0.0;
                        for (int k = 0; k < NS; ++k) v += Dm[Y][r * NS + k] * eb[k];
                        sb[r] = v;
                    }
                    // N_n v: the Voigt traction operator of n
                    if constexpr (DIM == 3) {
                        t[f][0] = n[0] * ea[0] + n[2] * ea[4] + n[1] * ea[5];
                        t[f][1] = n[1] * ea[1] + n[2] * ea[3] + n[0] * ea[5];
                        t[f][2] = n[2] * ea[2] + n[1] * ea[3] + n[0] * ea[4];
                        T[f][0] = n[0] * sb[0] + n[2] * sb[4] + n[1] * sb[5];
                        T[f][1] = n[1] * sb[1] + n[2] * sb[3] + n[0] * sb[5];
                        T[f][2] = n[2] * sb[2] + n[1] * sb[3] + n[0] * sb[4];
                    } else {
                        t[f][0] = n[0] * ea[0] + n[1] * ea[2];
                        t[f][1] = n[1] * ea[1] + n[0] * ea[2];
                        T[f][0] = n[0] * sb[0] + n[1] * sb[2];
                        T[f][1] = n[1] * sb[1] + n[0] * sb[2];
                    }
                }
                for (int f = 0; f < NF; ++f)
                    for (int g2 = 0; g2 < NF; ++g2) {
                        double v = 0.0;
                        for (int d = 0; d < DIM; ++d) v += t[f][d] * T[g2][d];
                        Kp[(f * NB2 + i) * NA2 + g2 * NB2 + j] += sg * v;
                    }
            }
        }
    }
    // both elements' T against the union of their adofs, and the union's adof values
    const int64_t m0 = s.map_ptr[c];
    const int nL = gv.na_t[eX[0]];
    const int64_t a0 = s.cl_aptr[c];
    const int nu = (int)((s.cl_aptr[c + 1] - a0) / NF);
    for (int t = tid; t < NB2 * M; t += 256) Tm[t] = 0.0;
    __syncthreads();
    if (tid < NB2) {
        const int X = tid / NB, b = tid % NB, e = eX[X];
        const int64_t mo = m0 + (X == 0 ? 0 : nL);
        for (int k = gv.t_ptr[(size_t)e * NB + b]; k < gv.t_ptr[(size_t)e * NB + b + 1]; ++k)
            Tm[tid * M + s.map[mo + gv.t_col[k]]] += gv.t_w[k];
    }
    for (int t = tid; t < NF * nu; t += 256) ue[t] = u[s.cl_adof[a0 + t]];
    __syncthreads();
    double upv = 0.0;
    if (tid < NA2) {
        const int f = tid / NB2, b = tid % NB2;
        for (int x = 0; x < nu; ++x) upv += Tm[b * M + x] * ue[f * nu + x];
    }
    __syncthreads();
    if (tid < NA2) up[tid] = upv;
    __syncthreads();
    double rp = 0.0;
    if (tid < NA2)
        for (int j = 0; j < NA2; ++j) rp += Kp[tid * NA2 + j] * up[j];
    __syncthreads();
    if (tid < NA2) up[tid] = rp;
    __syncthreads();
    double* re = s.re + s.re_off[c];
    for (int t = tid; t < NF * nu; t += 256) {
        const int f = t / nu, x = t % nu;
        double v = 0.0;
        for (int b = 0; b < NB2; ++b) v += Tm[b * M + x] * up[f * NB2 + b];
        re[t] = v;
    }
    double* ke = s.ke + s.ke_off[c];
    const int nat = NF * nu;
    for (int f = 0; f < NF; ++f)
        for (int g2 = 0; g2 < NF; ++g2) {
            __syncthreads();
            for (int t = tid; t < NB2 * nu; t += 256) {
                const int b = t / nu, y = t % nu;
                double v = 0.0;
                for (int b2 = 0; b2 < NB2; ++b2) v += Kp[(f * NB2 + b) * NA2 + g2 * NB2 + b2] * Tm[b2 * M + y];
                W[b * M + y] = v;
            }
            __syncthreads();
            for (int t = tid; t < nu * nu; t += 256) {
                const int x = t / nu, y = t % nu;
                double v = 0.0;
                for (int b = 0; b < NB2; ++b) v += Tm[b * M + x] * W[b * M + y];
                ke[(size_t)(f * nu + x) * nat + g2 * nu + y] = v;
            }
        }
}

}  // namespace mgpu

using namespace mgpu;

struct mgpu_sideset {
    mgpu_set* bulk = nullptr;
    int dim = 3, nf = 1, nb = 8, geom = MGPU_GEOM_TRI3;
    bool generic = false;      // side clusters of a generic-path set (k_side_gen)
    bool ghost = false;        // double-sided ghost facets (k_ghost)
    SideView v{};
    SideGenView gv{};
    GhostView hv{};
    int64_t n_kt = 0, n_rt = 0;
    const int64_t *kt_ptr = nullptr, *kt_dst = nullptr, *kt_src = nullptr;
    const int64_t *rt_ptr = nullptr, *rt_dst = nullptr, *rt_src = nullptr;
    std::vector<void*> allocs;
    mgpu_sideset() = default;
    mgpu_sideset(const mgpu_sideset&) = delete;
    mgpu_sideset& operator=(const mgpu_sideset&) = delete;
    ~mgpu_sideset() {
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {
template <class T>
mgpu_status side_upload(mgpu_sideset* S, const T* src, size_t n, const T** dst) {
    *dst = nullptr;
    if (n == 0) return MGPU_OK;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    S->allocs.push_back(p);
    HIP_TRY(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<const T*>(p);
    return MGPU_OK;
}

// sorted (target, pool index) pairs -> per-target ranges (stable: ascending pool index per target)
void build_gather(std::vector<std::pair<int64_t, int64_t>>& pairs, std::vector<int64_t>& ptr,
                  std::vector<int64_t>& dst, std::vector<int64_t>& src) {
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const std::pair<int64_t, int64_t>& x, const std::pair<int64_t, int64_t>& y) { return x.first < y.first; });
    ptr.assign(1, 0);
    for (size_t k = 0; k < pairs.size(); ++k) {
        if (k == 0 || pairs[k].first != pairs[k - 1].first) {
            if (k > 0) ptr.push_back((int64_t)k);
            dst.push_back(pairs[k].first);
        }
        src.push_back(pairs[k].second);
    }
    if (!pairs.empty()) ptr.push_back((int64_t)pairs.size());
}
}  // namespace

namespace {
// Side set of a generic-path bulk set: facets grouped into side clusters (one per element, in
// element order), the cluster's length measure, its element's adofs, pool offsets and the gather
// lists from each K_e / R_e entry to its CSR value / row.
mgpu_status create_generic_side(mgpu_set* bulk, const mgpu_sideset_desc* d, int kind, double beta, int nvs,
                                mgpu_sideset** out) {
    const SetHost& h = bulk->h;
    const int dim = h.dim, nf = h.nf, nb = h.nb;
    if (!((dim == 3 && (nb == 8 || nb == 27)) || (dim == 2 && nb == 4)))
        return set_error(MGPU_ERR_UNSUPPORTED, "side sets on generic sets need HEX8 / HEX27 / QUAD4 IP cells");
    const int32_t nfc = d->n_facets;
    for (int32_t k = 0; k < nfc; ++k)
        if (h.na_t[(size_t)d->facet_elem[k]] > kSideMaxNat)
            return set_error(MGPU_ERR_UNSUPPORTED, "side cluster element with more than 64 adofs per dof type");
    auto S = std::make_unique<mgpu_sideset>();
    S->bulk = bulk;
    S->dim = dim;
    S->nf = nf;
    S->nb = nb;
    S->geom = d->side_geometry;
    S->generic = true;
    SideGenView& v = S->gv;
    v.N = h.n_adofs_type;
    v.nm = h.nm;
    v.beta = beta;
    v.alpha = d->nitsche_alpha;
    v.neumann = kind == 1;
    for (int k = 0; k < 3; ++k) v.value[k] = d->value[k];
    // facets in element order (stable), grouped into clusters
    std::vector<int32_t> order((size_t)nfc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return d->facet_elem[x] < d->facet_elem[y]; });
    std::vector<double> box((size_t)h.n_el * 2 * dim);
    if (h.n_el > 0)
        HIP_TRY(hipMemcpy(box.data(), bulk->g.ip_box, box.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int32_t> cl_elem, cl_fptr(1, 0);
    std::vector<double> xi((size_t)nfc * nvs * dim), nrm((size_t)nfc * dim), hlen;
    double A = 0.0;
    for (int32_t k = 0; k < nfc; ++k) {
        const int32_t f = order[(size_t)k], e = d->facet_elem[f];
        if (k == 0 || e != cl_elem.back()) {
            if (k > 0) {
                cl_fptr.push_back(k);
                hlen.push_back(dim == 3 ? std::sqrt(4.0 * A / 3.141592653589793) : A);
            }
            cl_elem.push_back(e);
            A = 0.0;
        }
        std::copy(d->facet_xi + (size_t)f * nvs * dim, d->facet_xi + (size_t)(f + 1) * nvs * dim, xi.data() + (size_t)k * nvs * dim);
        std::copy(d->facet_normal + (size_t)f * dim, d->facet_normal + (size_t)(f + 1) * dim, nrm.data() + (size_t)k * dim);
        const double* X = d->facet_xi + (size_t)f * nvs * dim;
        const double* ext = box.data() + (size_t)e * 2 * dim + dim;
        auto px = [&](int vtx, int q) { return (X[vtx * dim + q] + 1.0) * 0.5 * ext[q]; };
        if (nvs == 2) {
            double l2 = 0.0;
            for (int q = 0; q < dim; ++q) l2 += (px(1, q) - px(0, q)) * (px(1, q) - px(0, q));
            A += std::sqrt(l2);
        } else {
            for (int t = 0; t + 2 < nvs; ++t) {
                double e1[3], e2[3];
                for (int q = 0; q < 3; ++q) {
                    e1[q] = px(t + 1, q) - px(0, q);
                    e2[q] = px(t + 2, q) - px(0, q);
                }
                const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2],
                             c2 = e1[0] * e2[1] - e1[1] * e2[0];
                A += 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            }
        }
    }
    if (nfc > 0) {
        cl_fptr.push_back(nfc);
        hlen.push_back(dim == 3 ? std::sqrt(4.0 * A / 3.141592653589793) : A);
    }
    const int32_t ncl = (int32_t)cl_elem.size();
    v.n_clusters = ncl;
    for (int32_t c = 0; c < ncl; ++c)
        if (kind == 0 && !(hlen[(size_t)c] > 0.0)) return set_error(MGPU_ERR_GEOMETRY, "a Dirichlet side cluster has zero measure");
    // adofs, pool offsets and gather lists per cluster
    std::vector<int64_t> aptr(1, 0), koff(1, 0), roff(1, 0);
    std::vector<int32_t> adof;
    std::vector<std::pair<int64_t, int64_t>> kp, rp;
    for (int32_t c = 0; c < ncl; ++c) {
        const int32_t e = cl_elem[(size_t)c];
        const int64_t t0 = h.topo_ptr[(size_t)e], nat = h.topo_ptr[(size_t)e + 1] - t0;
        for (int64_t r = 0; r < nat; ++r) {
            const int32_t row = h.topo[(size_t)(t0 + r)];
            adof.push_back(row);
            rp.push_back({row, roff.back() + r});
            if (kind == 1) continue;
            const int32_t* cb = h.col_idx.data() + h.row_ptr[(size_t)row];
            const int32_t* ce = h.col_idx.data() + h.row_ptr[(size_t)row + 1];
            for (int64_t cc = 0; cc < nat; ++cc) {
                const int32_t col = h.topo[(size_t)(t0 + cc)];
                const int32_t* it = std::lower_bound(cb, ce, col);
                if (it == ce || *it != col) return set_error(MGPU_ERR_STATE, "side entry outside the bulk set's pattern");
                kp.push_back({(int64_t)(it - h.col_idx.data()), koff.back() + r * nat + cc});
            }
        }
        aptr.push_back((int64_t)adof.size());
        koff.push_back(koff.back() + (kind == 1 ? 0 : nat * nat));
        roff.push_back(roff.back() + nat);
    }
    std::vector<int64_t> kptr, kdst, ksrc, rptr, rdst, rsrc;
    build_gather(kp, kptr, kdst, ksrc);
    build_gather(rp, rptr, rdst, rsrc);
    S->n_kt = (int64_t)kdst.size();
    S->n_rt = (int64_t)rdst.size();
    mgpu_status rc;
    if ((rc = side_upload(S.get(), cl_elem.data(), cl_elem.size(), &v.cl_elem)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), cl_fptr.data(), cl_fptr.size(), &v.cl_fptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), xi.data(), xi.size(), &v.xi)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), nrm.data(), nrm.size(), &v.normal)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), hlen.data(), hlen.size(), &v.hlen)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), aptr.data(), aptr.size(), &v.cl_aptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), adof.data(), adof.size(), &v.cl_adof)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), koff.data(), koff.size(), &v.ke_off)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), roff.data(), roff.size(), &v.re_off)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kptr.data(), kptr.size(), &S->kt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kdst.data(), kdst.size(), &S->kt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), ksrc.data(), ksrc.size(), &S->kt_src)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rptr.data(), rptr.size(), &S->rt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rdst.data(), rdst.size(), &S->rt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rsrc.data(), rsrc.size(), &S->rt_src)) != MGPU_OK) return rc;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, std::max<int64_t>(koff.back(), 1) * sizeof(double)));
    S->allocs.push_back(p);
    v.ke = static_cast<double*>(p);
    HIP_TRY(hipMalloc(&p, std::max<int64_t>(roff.back(), 1) * sizeof(double)));
    S->allocs.push_back(p);
    v.re = static_cast<double*>(p);
    S->v.n_facets = nfc;
    *out = S.release();
    return MGPU_OK;
}
}  // namespace

namespace {
// Ghost facets of a generic-path bulk set whose descriptor coupled every facet's two elements: per
// facet the union of the elements' adofs, the element-local -> union column maps, gamma = alpha h,
// pool offsets and the gather lists from each K_e / R_e entry to its CSR value / row.
mgpu_status create_ghost_side(mgpu_set* bulk, const mgpu_sideset_desc* d, mgpu_sideset** out) {
    const SetHost& h = bulk->h;
    const int dim = h.dim, nf = h.nf, nb = h.nb;
    if (!((dim == 3 && nb == 8) || (dim == 2 && nb == 4)))
        return set_error(MGPU_ERR_UNSUPPORTED, "ghost facets need HEX8 / QUAD4 IP cells");
    if (!h.generic) return set_error(MGPU_ERR_STATE, "the bulk set does not couple the ghost facets' elements");
    if (!(d->nitsche_alpha >= 0.0)) return set_error(MGPU_ERR_INVALID_ARG, "the ghost parameter alpha must be >= 0");
    const int32_t nfc = d->n_facets;
    std::vector<double> box((size_t)h.n_el * 2 * dim);
    if (h.n_el > 0)
        HIP_TRY(hipMemcpy(box.data(), bulk->g.ip_box, box.size() * sizeof(double), hipMemcpyDeviceToHost));
    auto S = std::make_unique<mgpu_sideset>();
    S->bulk = bulk;
    S->dim = dim;
    S->nf = nf;
    S->nb = nb;
    S->geom = d->side_geometry;
    S->ghost = true;
    GhostView& v = S->hv;
    v.n_facets = nfc;
    v.N = h.n_adofs_type;
    v.nm = h.nm;
    const int32_t N = h.n_adofs_type;
    std::vector<int32_t> axis((size_t)nfc), map;
    std::vector<double> side((size_t)nfc), gamma((size_t)nfc);
    std::vector<int64_t> aptr(1, 0), mptr(1, 0), koff(1, 0), roff(1, 0);
    std::vector<int32_t> adof, uni;
    std::vector<std::pair<int64_t, int64_t>> kp, rp;
    for (int32_t k = 0; k < nfc; ++k) {
        const int32_t eL = d->facet_elem[k], eF = d->facet_follower[k];
        if (eF < 0 || eF >= h.n_el || eF == eL) return set_error(MGPU_ERR_INVALID_ARG, "bad ghost follower element");
        const double* n = d->facet_normal + (size_t)k * dim;
        int ax = -1;
        for (int q = 0; q < dim; ++q)
            if (std::fabs(n[q]) == 1.0) ax = q;
        for (int q = 0; q < dim && ax >= 0; ++q)
            if (q != ax && n[q] != 0.0) ax = -1;
        if (ax < 0) return set_error(MGPU_ERR_INVALID_ARG, "a ghost facet normal must be a lattice axis");
        const double* bl = box.data() + (size_t)eL * 2 * dim;
        const double* bf = box.data() + (size_t)eF * 2 * dim;
        for (int q = 0; q < dim; ++q) {
            const double shift = q == ax ? n[ax] * bl[dim + q] : 0.0;
            if (bf[dim + q] != bl[dim + q] || std::fabs(bf[q] - (bl[q] + shift)) > 1e-12 * (1.0 + std::fabs(bl[q])))
                return set_error(MGPU_ERR_INVALID_ARG, "ghost facet elements are not equal neighbouring boxes");
        }
        axis[(size_t)k] = ax;
        side[(size_t)k] = n[ax];
        double A = 1.0;
        for (int q = 0; q < dim; ++q)
            if (q != ax) A *= bl[dim + q];
        gamma[(size_t)k] = d->nitsche_alpha * (dim == 3 ? std::sqrt(4.0 * A / 3.141592653589793) : A);
        // union of the two elements' type-local adofs (the type-0 rows of their topologies)
        const int32_t* tl = h.topo.data() + h.topo_ptr[(size_t)eL];
        const int32_t* tf = h.topo.data() + h.topo_ptr[(size_t)eF];
        const int nL = h.na_t[(size_t)eL], nF = h.na_t[(size_t)eF];
        uni.clear();
        std::set_union(tl, tl + nL, tf, tf + nF, std::back_inserter(uni));
        const int nu = (int)uni.size();
        if (nu > kSideMaxNat) return set_error(MGPU_ERR_UNSUPPORTED, "ghost facet with more than 64 adofs per dof type");
        for (int x = 0; x < nL; ++x) map.push_back((int32_t)(std::lower_bound(uni.begin(), uni.end(), tl[x]) - uni.begin()));
        for (int x = 0; x < nF; ++x) map.push_back((int32_t)(std::lower_bound(uni.begin(), uni.end(), tf[x]) - uni.begin()));
        mptr.push_back((int64_t)map.size());
        const int64_t nat = (int64_t)nf * nu;
        for (int64_t r = 0; r < nat; ++r) {
            const int32_t row = uni[(size_t)(r % nu)] + (int32_t)(r / nu) * N;
            adof.push_back(row);
            rp.push_back({row, roff.back() + r});
            const int32_t* cb = h.col_idx.data() + h.row_ptr[(size_t)row];
            const int32_t* ce = h.col_idx.data() + h.row_ptr[(size_t)row + 1];
            for (int64_t cc = 0; cc < nat; ++cc) {
                const int32_t col = uni[(size_t)(cc % nu)] + (int32_t)(cc / nu) * N;
                const int32_t* it = std::lower_bound(cb, ce, col);
                if (it == ce || *it != col) return set_error(MGPU_ERR_STATE, "ghost entry outside the bulk set's pattern");
                kp.push_back({(int64_t)(it - h.col_idx.data()), koff.back() + r * nat + cc});
            }
        }
        aptr.push_back((int64_t)adof.size());
        koff.push_back(koff.back() + nat * nat);
        roff.push_back(roff.back() + nat);
    }
    std::vector<int64_t> kptr, kdst, ksrc, rptr, rdst, rsrc;
    build_gather(kp, kptr, kdst, ksrc);
    build_gather(rp, rptr, rdst, rsrc);
    S->n_kt = (int64_t)kdst.size();
    S->n_rt = (int64_t)rdst.size();
    mgpu_status rc;
    if ((rc = side_upload(S.get(), d->facet_elem, (size_t)nfc, &v.leader)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), d->facet_follower, (size_t)nfc, &v.follower)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), axis.data(), axis.size(), &v.axis)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), side.data(), side.size(), &v.side)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), gamma.data(), gamma.size(), &v.gamma)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), aptr.data(), aptr.size(), &v.cl_aptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), adof.data(), adof.size(), &v.cl_adof)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), mptr.data(), mptr.size(), &v.map_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), map.data(), map.size(), &v.map)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), koff.data(), koff.size(), &v.ke_off)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), roff.data(), roff.size(), &v.re_off)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kptr.data(), kptr.size(), &S->kt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kdst.data(), kdst.size(), &S->kt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), ksrc.data(), ksrc.size(), &S->kt_src)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rptr.data(), rptr.size(), &S->rt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rdst.data(), rdst.size(), &S->rt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rsrc.data(), rsrc.size(), &S->rt_src)) != MGPU_OK) return rc;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, std::max<int64_t>(koff.back(), 1) * sizeof(double)));
    S->allocs.push_back(p);
    v.ke = static_cast<double*>(p);
    HIP_TRY(hipMalloc(&p, std::max<int64_t>(roff.back(), 1) * sizeof(double)));
    S->allocs.push_back(p);
    v.re = static_cast<double*>(p);
    S->v.n_facets = nfc;
    *out = S.release();
    return MGPU_OK;
}
}  // namespace

extern "C" mgpu_status mgpu_sideset_create(mgpu_set* bulk, const mgpu_sideset_desc* d, mgpu_sideset** out) {
    MGPU_GUARD_BEGIN
    if (!bulk || !d || !out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (mgpu_status rc = ensure_host(bulk); rc != MGPU_OK) return rc;
    const SetHost& h = bulk->h;
    const int dim = h.dim, nf = h.nf, nb = h.nb, na = nf * nb;
    const bool diff = h.iwg == MGPU_IWG_SPATIALDIFF_BULK;
    int kind = -1;
    double beta = 0.0;
    switch (d->iwg) {
        case MGPU_IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE: kind = diff ? 0 : -1; beta = -1.0; break;
        case MGPU_IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE: kind = diff ? 0 : -1; beta = 1.0; break;
        case MGPU_IWG_SPATIALDIFF_NEUMANN: kind = diff ? 1 : -1; break;
        case MGPU_IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE: kind = diff ? -1 : 0; beta = -1.0; break;
        case MGPU_IWG_STRUC_LINEAR_DIRICHLET_UNSYMMETRIC_NITSCHE: kind = diff ? -1 : 0; beta = 1.0; break;
        case MGPU_IWG_STRUC_LINEAR_NEUMANN: kind = diff ? -1 : 1; break;
        case MGPU_IWG_SPATIALDIFF_GHOST: kind = diff ? 2 : -1; break;
        case MGPU_IWG_STRUC_LINEAR_GHOST: kind = diff ? -1 : 2; break;
        default: return set_error(MGPU_ERR_UNSUPPORTED, "unknown side IWG");
    }
    if (kind < 0) return set_error(MGPU_ERR_INVALID_ARG, "side IWG does not match the bulk set's field");
    const bool geom_ok = dim == 3 ? (d->side_geometry == MGPU_GEOM_TRI3 || d->side_geometry == MGPU_GEOM_QUAD4)
                                  : d->side_geometry == MGPU_GEOM_LINE2;
    if (!geom_ok) return set_error(MGPU_ERR_INVALID_ARG, "side geometry must be TRI3 / QUAD4 in 3D, LINE2 in 2D");
    if (d->n_facets < 0 || (d->n_facets > 0 && (!d->facet_elem || !d->facet_xi || !d->facet_normal)))
        return set_error(MGPU_ERR_INVALID_ARG, "bad facet arrays");
    if (kind == 0 && !(d->nitsche_alpha >= 0.0)) return set_error(MGPU_ERR_INVALID_ARG, "nitsche_alpha must be >= 0");
    const int nvs = d->side_geometry == MGPU_GEOM_TRI3 ? 3 : (d->side_geometry == MGPU_GEOM_QUAD4 ? 4 : 2);
    const int32_t nfc = d->n_facets;
    for (int32_t k = 0; k < nfc; ++k)
        if (d->facet_elem[k] < 0 || d->facet_elem[k] >= h.n_el) return set_error(MGPU_ERR_INVALID_ARG, "facet element out of range");
    if (kind == 2) {
        if (d->side_geometry != (dim == 3 ? MGPU_GEOM_QUAD4 : MGPU_GEOM_LINE2))
            return set_error(MGPU_ERR_INVALID_ARG, "ghost facets are whole cell faces: QUAD4 in 3D, LINE2 in 2D");
        if (nfc > 0 && !d->facet_follower) return set_error(MGPU_ERR_INVALID_ARG, "ghost facets need facet_follower");
        return create_ghost_side(bulk, d, out);
    }
    if (h.generic) return create_generic_side(bulk, d, kind, beta, nvs, out);
    auto S = std::make_unique<mgpu_sideset>();
    S->bulk = bulk;
    S->dim = dim;
    S->nf = nf;
    S->nb = nb;
    S->geom = d->side_geometry;
    SideView& v = S->v;
    v.n_facets = nfc;
    v.N = h.n_adofs_type;
    v.beta = beta;
    v.alpha = d->nitsche_alpha;
    v.neumann = kind == 1;
    for (int k = 0; k < 3; ++k) v.value[k] = d->value[k];
    // element -> lattice cell (IP box) and material: from the bulk set's host maps and descriptor copy
    std::vector<int64_t> elem_cellidx((size_t)h.n_el, -1);
    for (int64_t c = 0; c < (int64_t)h.cell_elem.size(); ++c)
        if (h.cell_elem[(size_t)c] >= 0) elem_cellidx[(size_t)h.cell_elem[(size_t)c]] = c;
    std::vector<double> mat_host((size_t)h.n_el * h.nm);
    if (h.n_el > 0)
        HIP_TRY(hipMemcpy(mat_host.data(), bulk->v.material, mat_host.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> box((size_t)nfc * 2 * dim), mat((size_t)nfc * 2, 0.0), measure((size_t)nfc);
    std::vector<int32_t> vadof((size_t)nfc * nb);
    for (int32_t k = 0; k < nfc; ++k) {
        const int32_t e = d->facet_elem[k];
        const int64_t c = elem_cellidx[(size_t)e];
        const int64_t ijk[3] = {c % h.lat.n[0], (c / h.lat.n[0]) % h.lat.n[1], dim == 3 ? c / ((int64_t)h.lat.n[0] * h.lat.n[1]) : 0};
        for (int q = 0; q < dim; ++q) {
            box[(size_t)k * 2 * dim + q] = h.lat.lo[q] + (double)ijk[q] * h.lat.h[q];
            box[(size_t)k * 2 * dim + dim + q] = h.lat.h[q];
        }
        for (int q = 0; q < h.nm; ++q) mat[2 * (size_t)k + q] = mat_host[(size_t)e * h.nm + q];
        for (int b = 0; b < nb; ++b) vadof[(size_t)k * nb + b] = h.topo[(size_t)e * h.na + h.vert_pos[(size_t)e * nb + b]];
        // physical measure of the facet
        const double* X = d->facet_xi + (size_t)k * nvs * dim;
        auto px = [&](int vtx, int q) { return (X[vtx * dim + q] + 1.0) * 0.5 * h.lat.h[q]; };
        double m = 0.0;
        if (nvs == 2) {
            double l2 = 0.0;
            for (int q = 0; q < dim; ++q) l2 += (px(1, q) - px(0, q)) * (px(1, q) - px(0, q));
            m = std::sqrt(l2);
        } else {
            for (int t = 0; t + 2 < nvs; ++t) {   // fan of triangles (one for TRI3, two for QUAD4)
                double e1[3], e2[3];
                for (int q = 0; q < 3; ++q) {
                    e1[q] = px(t + 1, q) - px(0, q);
                    e2[q] = px(t + 2, q) - px(0, q);
                }
                const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2],
                             c2 = e1[0] * e2[1] - e1[1] * e2[0];
                m += 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            }
        }
        measure[(size_t)k] = m;
    }
    // side cluster (all facets of one element in this set): length measure from the total measure
    std::vector<double> cluster((size_t)h.n_el, 0.0), hlen((size_t)nfc);
    for (int32_t k = 0; k < nfc; ++k) cluster[(size_t)d->facet_elem[k]] += measure[(size_t)k];
    for (int32_t k = 0; k < nfc; ++k) {
        const double A = cluster[(size_t)d->facet_elem[k]];
        hlen[(size_t)k] = dim == 3 ? std::sqrt(4.0 * A / 3.141592653589793) : A;
        if (kind == 0 && !(hlen[(size_t)k] > 0.0)) return set_error(MGPU_ERR_GEOMETRY, "a Dirichlet side cluster has zero measure");
    }
    // gather lists: facet matrix entry (r, c) -> CSR value of (row adof(r), column adof(c))
    const int32_t N = h.n_adofs_type;
    std::vector<std::pair<int64_t, int64_t>> kp, rp;
    kp.reserve((size_t)nfc * na * na);
    rp.reserve((size_t)nfc * na);
    for (int32_t k = 0; k < nfc; ++k)
        for (int r = 0; r < na; ++r) {
            const int32_t row = (r / nb) * N + vadof[(size_t)k * nb + r % nb];
            rp.push_back({row, (int64_t)k * na + r});
            if (kind == 1) continue;
            const int32_t* cb = h.col_idx.data() + h.row_ptr[(size_t)row];
            const int32_t* ce = h.col_idx.data() + h.row_ptr[(size_t)row + 1];
            for (int c = 0; c < na; ++c) {
                const int32_t col = (c / nb) * N + vadof[(size_t)k * nb + c % nb];
                const int32_t* it = std::lower_bound(cb, ce, col);
                if (it == ce || *it != col) return set_error(MGPU_ERR_STATE, "facet entry outside the bulk set's pattern");
                kp.push_back({(int64_t)(it - h.col_idx.data()), ((int64_t)k * na + r) * na + c});
            }
        }
    std::vector<int64_t> kptr, kdst, ksrc, rptr, rdst, rsrc;
    build_gather(kp, kptr, kdst, ksrc);
    build_gather(rp, rptr, rdst, rsrc);
    S->n_kt = (int64_t)kdst.size();
    S->n_rt = (int64_t)rdst.size();
    mgpu_status rc;
    if ((rc = side_upload(S.get(), d->facet_xi, (size_t)nfc * nvs * dim, &v.xi)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), d->facet_normal, (size_t)nfc * dim, &v.normal)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), hlen.data(), hlen.size(), &v.hlen)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), box.data(), box.size(), &v.box)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), mat.data(), mat.size(), &v.mat)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), vadof.data(), vadof.size(), &v.vadof)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kptr.data(), kptr.size(), &S->kt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), kdst.data(), kdst.size(), &S->kt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), ksrc.data(), ksrc.size(), &S->kt_src)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rptr.data(), rptr.size(), &S->rt_ptr)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rdst.data(), rdst.size(), &S->rt_dst)) != MGPU_OK) return rc;
    if ((rc = side_upload(S.get(), rsrc.data(), rsrc.size(), &S->rt_src)) != MGPU_OK) return rc;
    if (nfc > 0) {
        void* p = nullptr;
        HIP_TRY(hipMalloc(&p, (size_t)nfc * na * na * sizeof(double)));
        S->allocs.push_back(p);
        v.ke = static_cast<double*>(p);
        HIP_TRY(hipMalloc(&p, (size_t)nfc * na * sizeof(double)));
        S->allocs.push_back(p);
        v.re = static_cast<double*>(p);
    }
    *out = S.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_sideset_destroy(mgpu_sideset* ss) {
    delete ss;
    return MGPU_OK;
}

namespace {
template <int DIM, int NF>
void launch_side(mgpu_sideset* S, const double* u, hipStream_t st) {
    const dim3 g((unsigned)S->v.n_facets), b(64);
    if constexpr (DIM == 3) {
        if (S->geom == MGPU_GEOM_TRI3) hipLaunchKernelGGL((k_side<3, NF, MGPU_GEOM_TRI3>), g, b, 0, st, S->v, u);
        else hipLaunchKernelGGL((k_side<3, NF, MGPU_GEOM_QUAD4>), g, b, 0, st, S->v, u);
    } else {
        hipLaunchKernelGGL((k_side<2, NF, MGPU_GEOM_LINE2>), g, b, 0, st, S->v, u);
    }
}
template <int DIM, int NF, int NB>
void launch_side_gen_nb(mgpu_sideset* S, const double* u, hipStream_t st) {
    const dim3 g((unsigned)S->gv.n_clusters), b(256);
    if constexpr (DIM == 3) {
        if (S->geom == MGPU_GEOM_TRI3)
            hipLaunchKernelGGL((k_side_gen<3, NF, MGPU_GEOM_TRI3, NB>), g, b, 0, st, S->gv, S->bulk->g, u);
        else
            hipLaunchKernelGGL((k_side_gen<3, NF, MGPU_GEOM_QUAD4, NB>), g, b, 0, st, S->gv, S->bulk->g, u);
    } else {
        hipLaunchKernelGGL((k_side_gen<2, NF, MGPU_GEOM_LINE2, NB>), g, b, 0, st, S->gv, S->bulk->g, u);
    }
}

void launch_side_gen(mgpu_sideset* S, const double* u, hipStream_t st) {
    if (S->gv.n_clusters == 0) return;
    if (S->dim == 3 && S->nb == 27) {
        if (S->nf == 1) launch_side_gen_nb<3, 1, 27>(S, u, st);
        else launch_side_gen_nb<3, 3, 27>(S, u, st);
    } else if (S->dim == 3) {
        if (S->nf == 1) launch_side_gen_nb<3, 1, 8>(S, u, st);
        else launch_side_gen_nb<3, 3, 8>(S, u, st);
    } else {
        if (S->nf == 1) launch_side_gen_nb<2, 1, 4>(S, u, st);
        else launch_side_gen_nb<2, 2, 4>(S, u, st);
    }
}
}  // namespace

extern "C" mgpu_status mgpu_sideset_assemble(mgpu_sideset* S, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!S) return set_error(MGPU_ERR_INVALID_ARG, "null side set");
    if (S->v.n_facets == 0) return MGPU_OK;
    if (!u || !vals || !rhs) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    hipStream_t st = S->bulk->stream;
    if (S->ghost) {
        const dim3 g((unsigned)S->hv.n_facets), b(256);
        if (S->dim == 3 && S->nf == 1) hipLaunchKernelGGL((k_ghost<3, 1>), g, b, 0, st, S->hv, S->bulk->g, u);
        else if (S->dim == 3) hipLaunchKernelGGL((k_ghost<3, 3>), g, b, 0, st, S->hv, S->bulk->g, u);
        else if (S->nf == 1) hipLaunchKernelGGL((k_ghost<2, 1>), g, b, 0, st, S->hv, S->bulk->g, u);
        else hipLaunchKernelGGL((k_ghost<2, 2>), g, b, 0, st, S->hv, S->bulk->g, u);
    } else if (S->generic) {
        launch_side_gen(S, u, st);
    } else if (S->dim == 3 && S->nf == 1) launch_side<3, 1>(S, u, st);
    else if (S->dim == 3) launch_side<3, 3>(S, u, st);
    else if (S->nf == 1) launch_side<2, 1>(S, u, st);
    else launch_side<2, 2>(S, u, st);
    HIP_TRY(hipGetLastError());
    const double* kpool = S->ghost ? S->hv.ke : (S->generic ? S->gv.ke : S->v.ke);
    const double* rpool = S->ghost ? S->hv.re : (S->generic ? S->gv.re : S->v.re);
    if (S->n_kt > 0)
        hipLaunchKernelGGL(k_side_gather, dim3((unsigned)((S->n_kt + 255) / 256)), dim3(256), 0, st, S->n_kt, S->kt_ptr,
                           S->kt_dst, S->kt_src, kpool, vals);
    if (S->n_rt > 0)
        hipLaunchKernelGGL(k_side_gather, dim3((unsigned)((S->n_rt + 255) / 256)), dim3(256), 0, st, S->n_rt, S->rt_ptr,
                           S->rt_dst, S->rt_src, rpool, rhs);
    HIP_TRY(hipGetLastError());
    return MGPU_OK;
    MGPU_GUARD_END
}
