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

// out[dst[t]] += sum of pool[src[k]], k in [ptr[t], ptr[t+1]) (one thread per target)
__global__ __launch_bounds__(256) void k_side_gather(int64_t n, const int64_t* ptr, const int64_t* dst,
                                                     const int64_t* src, const double* pool, double* out) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double v = 0.0;
    for (int64_t k = ptr[t]; k < ptr[t + 1]; ++k) v += pool[src[k]];
    out[dst[t]] += v;
}

}  // namespace mgpu

using namespace mgpu;

struct mgpu_sideset {
    mgpu_set* bulk = nullptr;
    int dim = 3, nf = 1, nb = 8, geom = MGPU_GEOM_TRI3;
    SideView v{};
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
        default: return set_error(MGPU_ERR_UNSUPPORTED, "unknown side IWG");
    }
    if (kind < 0) return set_error(MGPU_ERR_INVALID_ARG, "side IWG does not match the bulk set's field");
    const bool geom_ok = dim == 3 ? (d->side_geometry == MGPU_GEOM_TRI3 || d->side_geometry == MGPU_GEOM_QUAD4)
                                  : d->side_geometry == MGPU_GEOM_LINE2;
    if (!geom_ok) return set_error(MGPU_ERR_INVALID_ARG, "side geometry must be TRI3 / QUAD4 in 3D, LINE2 in 2D");
    if (h.generic) return set_error(MGPU_ERR_UNSUPPORTED, "side sets are implemented for lattice sets with nodal T rows");
    if (d->n_facets < 0 || (d->n_facets > 0 && (!d->facet_elem || !d->facet_xi || !d->facet_normal)))
        return set_error(MGPU_ERR_INVALID_ARG, "bad facet arrays");
    if (kind == 0 && !(d->nitsche_alpha >= 0.0)) return set_error(MGPU_ERR_INVALID_ARG, "nitsche_alpha must be >= 0");
    const int nvs = d->side_geometry == MGPU_GEOM_TRI3 ? 3 : (d->side_geometry == MGPU_GEOM_QUAD4 ? 4 : 2);
    const int32_t nfc = d->n_facets;
    for (int32_t k = 0; k < nfc; ++k)
        if (d->facet_elem[k] < 0 || d->facet_elem[k] >= h.n_el) return set_error(MGPU_ERR_INVALID_ARG, "facet element out of range");
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
}  // namespace

extern "C" mgpu_status mgpu_sideset_assemble(mgpu_sideset* S, const double* u, double* vals, double* rhs) {
    MGPU_GUARD_BEGIN
    if (!S) return set_error(MGPU_ERR_INVALID_ARG, "null side set");
    if (S->v.n_facets == 0) return MGPU_OK;
    if (!u || !vals || !rhs) return set_error(MGPU_ERR_INVALID_ARG, "null buffer");
    hipStream_t st = S->bulk->stream;
    if (S->dim == 3 && S->nf == 1) launch_side<3, 1>(S, u, st);
    else if (S->dim == 3) launch_side<3, 3>(S, u, st);
    else if (S->nf == 1) launch_side<2, 1>(S, u, st);
    else launch_side<2, 2>(S, u, st);
    HIP_TRY(hipGetLastError());
    if (S->n_kt > 0)
        hipLaunchKernelGGL(k_side_gather, dim3((unsigned)((S->n_kt + 255) / 256)), dim3(256), 0, st, S->n_kt, S->kt_ptr,
                           S->kt_dst, S->kt_src, S->v.ke, vals);
    if (S->n_rt > 0)
        hipLaunchKernelGGL(k_side_gather, dim3((unsigned)((S->n_rt + 255) / 256)), dim3(256), 0, st, S->n_rt, S->rt_ptr,
                           S->rt_dst, S->rt_src, S->v.re, rhs);
    HIP_TRY(hipGetLastError());
    return MGPU_OK;
    MGPU_GUARD_END
}
