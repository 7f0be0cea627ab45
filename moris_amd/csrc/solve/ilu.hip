// ilu.hip -- incomplete LU preconditioner ILU(k) on the GPU, for the GMRES solve of assembled
// systems.  Target: gfx950 (MI355X), 64-wide wavefronts.
//
// The reference's solver pairs AztecOO GMRES with an incomplete factorisation of each subdomain
// (AZ_precond = AZ_dom_decomp, AZ_subdomain_solve = AZ_ilu, AZ_graph_fill = k: ILU(k),
// projects/SOL/DLA/src/cl_DLA_Linear_Solver_Aztec.cpp:353-380; UT_MDL_XTK_HMR_Linear_Struc_2D.cpp
// :990-993 solves with AZ_gmres + AZ_ilu).  Aztec itself is a Trilinos package that is not part of
// the reference sources; the algorithm restated here is the published ILU(k) (Saad, Iterative
// Methods for Sparse Linear Systems, 2nd ed., Algorithm 10.4 (ILU(0), IKJ order) and Algorithm
// 10.5 / eq. 10.14 (level of fill: lev(a_ij) = min(lev(a_ij), lev(a_ik) + lev(a_kj) + 1), fill
// kept up to level k)).
//
//   symbolic (host, once per pattern)   the factor pattern (the matrix pattern plus the fill of level
//                                       <= k, every diagonal present), the diagonal positions and
//                                       two level schedules: row i's forward level is one more than
//                                       the largest of the rows k < i in its pattern, its backward
//                                       level one more than the largest of the rows j > i
//   numeric (device, per matrix)        k_ilu_fill / k_ilu_scatter copy K into the factor pattern;
//                                       one launch per forward level of k_ilu_row, one workgroup of
//                                       one wavefront per row: the row in LDS, for each k < i in
//                                       order a_ik /= u_kk, then the lanes update the entries j > k
//                                       of the row found in row k's upper part (binary search)
//   apply (device)                      z = U^-1 L^-1 v in place: one launch per forward level of
//                                       k_ilu_lower (one wavefront per row, butterfly sum in a fixed
//                                       order), then per backward level of k_ilu_upper
// Every value is FP64, every sum in a fixed order: the factors and the applied vectors are the
// same on every call.
#include "ilu.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "../../../include/moris_gpu_solve.h"
#include "../setup/internal.hpp"

#define ILU_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace mgpu {
namespace ilu {

constexpr int kMaxRow = 5000;   // longest factor row the numeric kernel keeps in LDS (60 KB)

__global__ void k_ilu_clear(int32_t* err) {
    if (threadIdx.x == 0) err[0] = 0;
}

__global__ __launch_bounds__(256) void k_ilu_fill(int64_t n, double* lu) {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) lu[p] = 0.0;
}

__global__ __launch_bounds__(256) void k_ilu_scatter(int64_t n, const int32_t* pos, const double* vals, double* lu) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) lu[pos[k]] = vals[k];
}

// One row of the numeric factorisation (IKJ order); rows of earlier levels are final.
__global__ __launch_bounds__(64) void k_ilu_row(const int32_t* rows, const int32_t* rp, const int32_t* ci,
                                                const int32_t* diag, double* lu, int32_t* err, int max_row) {
    extern __shared__ double sm[];
    double* v = sm;
    int32_t* c = reinterpret_cast<int32_t*>(sm + max_row);
    const int i = rows[blockIdx.x];
    const int lane = threadIdx.x;
    const int r0 = rp[i], len = rp[i + 1] - r0, dpos = diag[i] - r0;
    for (int q = lane; q < len; q += 64) {
        v[q] = lu[r0 + q];
        c[q] = ci[r0 + q];
    }
    __syncthreads();
    for (int p = 0; p < dpos; ++p) {
        const int k = c[p];
        const double piv = lu[diag[k]];
        const double lik = v[p] / piv;
        if (piv == 0.0 && lane == 0) err[0] = 1;
        __syncthreads();   // every lane has read v[p]
        if (lane == 0) v[p] = lik;
        const int u0 = diag[k] + 1, u1 = rp[k + 1];
        for (int q = p + 1 + lane; q < len; q += 64) {
            const int j = c[q];
            int lo = u0, hi = u1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (lo < u1 && ci[lo] == j) v[q] -= lik * lu[lo];
        }
        __syncthreads();
    }
    for (int q = lane; q < len; q += 64) lu[r0 + q] = v[q];
    if (lane == 0 && v[dpos] == 0.0) err[0] = 1;
}

// y_i -= sum_{k < i} l_ik y_k for the rows of one forward level (one wavefront per row).
__global__ __launch_bounds__(256) void k_ilu_lower(int32_t cnt, const int32_t* rows, const int32_t* rp,
                                                   const int32_t* ci, const int32_t* diag, const double* lu,
                                                   double* y, const int32_t* skip) {
    if (skip && *skip) return;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= cnt) return;
    const int i = rows[w];
    double s = 0.0;
    for (int p = rp[i] + lane; p < diag[i]; p += 64) s += lu[p] * y[ci[p]];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) y[i] -= s;
}

// z_i = (z_i - sum_{j > i} u_ij z_j) / u_ii for the rows of one backward level.
__global__ __launch_bounds__(256) void k_ilu_upper(int32_t cnt, const int32_t* rows, const int32_t* rp,
                                                   const int32_t* ci, const int32_t* diag, const double* lu,
                                                   double* z, const int32_t* skip) {
    if (skip && *skip) return;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= cnt) return;
    const int i = rows[w];
    double s = 0.0;
    for (int p = diag[i] + 1 + lane; p < rp[i + 1]; p += 64) s += lu[p] * z[ci[p]];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) z[i] = (z[i] - s) / lu[diag[i]];
}

__global__ __launch_bounds__(256) void k_ilu_copy(int32_t n, const double* src, double* dst) {
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) dst[r] = src[r];
}

template <class T>
mgpu_status dev_upload(mgpu_ilu* f, const std::vector<T>& h, T** d) {
    *d = nullptr;
    if (h.empty()) return MGPU_OK;
    void* p = nullptr;
    ILU_TRY(hipMalloc(&p, h.size() * sizeof(T)));
    f->allocs.push_back(p);
    ILU_TRY(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *d = static_cast<T*>(p);
    return MGPU_OK;
}

// Rows ordered by level (counting sort, ascending row within a level); offsets per level.
void schedule(const std::vector<int32_t>& level, std::vector<int32_t>& rows, std::vector<int32_t>& off) {
    const int32_t n = (int32_t)level.size();
    int32_t n_lev = 0;
    for (int32_t v : level) n_lev = std::max(n_lev, v + 1);
    off.assign((size_t)n_lev + 1, 0);
    for (int32_t v : level) off[(size_t)v + 1]++;
    for (int32_t l = 0; l < n_lev; ++l) off[(size_t)l + 1] += off[(size_t)l];
    rows.assign((size_t)n, 0);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int32_t i = 0; i < n; ++i) rows[(size_t)fill[(size_t)level[(size_t)i]]++] = i;
}

}  // namespace ilu

void ilu_apply_launch(const mgpu_ilu* f, double* v, const int32_t* skip, hipStream_t s) {
    const dim3 t(256);
    for (size_t l = 0; l + 1 < f->lev_f.size(); ++l) {
        const int32_t cnt = f->lev_f[l + 1] - f->lev_f[l];
        hipLaunchKernelGGL(ilu::k_ilu_lower, dim3((unsigned)((cnt + 3) / 4)), t, 0, s, cnt, f->rows_f + f->lev_f[l],
                           f->rp, f->ci, f->diag, f->lu, v, skip);
    }
    for (size_t l = 0; l + 1 < f->lev_b.size(); ++l) {
        const int32_t cnt = f->lev_b[l + 1] - f->lev_b[l];
        hipLaunchKernelGGL(ilu::k_ilu_upper, dim3((unsigned)((cnt + 3) / 4)), t, 0, s, cnt, f->rows_b + f->lev_b[l],
                           f->rp, f->ci, f->diag, f->lu, v, skip);
    }
}

}  // namespace mgpu

using namespace mgpu;

extern "C" mgpu_status mgpu_ilu_create(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t fill_level,
                                       void* stream, mgpu_ilu** out) {
    MGPU_GUARD_BEGIN
    if (!out) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (n < 0 || fill_level < 0) return set_error(MGPU_ERR_INVALID_ARG, "negative size or fill level");
    if (n > 0 && (!row_ptr || !col_idx)) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the matrix pattern on the host
    std::vector<int32_t> rp((size_t)n + 1, 0);
    if (n > 0) {
        ILU_TRY(hipMemcpyAsync(rp.data(), row_ptr, ((size_t)n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        ILU_TRY(hipStreamSynchronize(s));
    }
    if (rp[0] != 0) return set_error(MGPU_ERR_INVALID_ARG, "row_ptr[0] must be 0");
    for (int32_t i = 0; i < n; ++i)
        if (rp[(size_t)i + 1] < rp[(size_t)i]) return set_error(MGPU_ERR_INVALID_ARG, "row_ptr not monotone");
    const int64_t nnz_in = rp[(size_t)n];
    std::vector<int32_t> ci((size_t)nnz_in);
    if (nnz_in > 0) {
        ILU_TRY(hipMemcpyAsync(ci.data(), col_idx, (size_t)nnz_in * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        ILU_TRY(hipStreamSynchronize(s));
    }
    for (int32_t i = 0; i < n; ++i)
        for (int32_t p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) {
            if (ci[(size_t)p] < 0 || ci[(size_t)p] >= n) return set_error(MGPU_ERR_INVALID_ARG, "column out of range");
            if (p > rp[(size_t)i] && ci[(size_t)p] <= ci[(size_t)p - 1])
                return set_error(MGPU_ERR_INVALID_ARG, "columns not sorted ascending");
        }
    // factor pattern: ILU(k) fill by levels, row by row; a missing diagonal is added
    std::vector<int32_t> frp((size_t)n + 1, 0), fci;
    fci.reserve((size_t)nnz_in + (size_t)n);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> upper(fill_level > 0 ? (size_t)n : 0);
    std::map<int32_t, int32_t> row;
    for (int32_t i = 0; i < n; ++i) {
        if (fill_level == 0) {
            bool has_diag = false;
            for (int32_t p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) {
                const int32_t c = ci[(size_t)p];
                if (!has_diag && c > i) {
                    fci.push_back(i);
                    has_diag = true;
                }
                if (c == i) has_diag = true;
                fci.push_back(c);
            }
            if (!has_diag) fci.push_back(i);
        } else {
            row.clear();
            for (int32_t p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) row[ci[(size_t)p]] = 0;
            row.emplace(i, 0);
            for (auto it = row.begin(); it != row.end() && it->first < i; ++it) {
                const int32_t lik = it->second;
                for (const auto& kj : upper[(size_t)it->first]) {
                    const int32_t lev = lik + kj.second + 1;
                    if (lev > fill_level) continue;
                    auto f = row.find(kj.first);
                    if (f == row.end()) row.emplace(kj.first, lev);
                    else f->second = std::min(f->second, lev);
                }
            }
            for (const auto& e : row) {
                fci.push_back(e.first);
                if (e.first > i) upper[(size_t)i].push_back(e);
            }
        }
        if (fci.size() >= (size_t)INT32_MAX) return set_error(MGPU_ERR_INVALID_ARG, "factor pattern exceeds int32");
        frp[(size_t)i + 1] = (int32_t)fci.size();
    }
    upper.clear();
    auto F = std::make_unique<mgpu_ilu>();
    struct Cleanup {
        mgpu_ilu* f;
        ~Cleanup() {
            if (f)
                for (void* p : f->allocs) (void)hipFree(p);
        }
    } cleanup{F.get()};
    F->n = n;
    F->nnz = (int64_t)fci.size();
    F->nnz_in = nnz_in;
    F->fill_level = fill_level;
    std::vector<int32_t> diag((size_t)n), pos((size_t)nnz_in), lev_f((size_t)n, 0), lev_b((size_t)n, 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t a = frp[(size_t)i], b = frp[(size_t)i + 1];
        F->max_row = std::max(F->max_row, b - a);
        diag[(size_t)i] = (int32_t)(std::lower_bound(fci.begin() + a, fci.begin() + b, i) - fci.begin());
        // matrix entries into the factor row (both sorted)
        int32_t q = a;
        for (int32_t p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) {
            while (fci[(size_t)q] < ci[(size_t)p]) ++q;
            pos[(size_t)p] = q;
        }
        int32_t lf = 0;
        for (int32_t p = a; p < diag[(size_t)i]; ++p) lf = std::max(lf, lev_f[(size_t)fci[(size_t)p]] + 1);
        lev_f[(size_t)i] = lf;
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        int32_t lb = 0;
        for (int32_t p = diag[(size_t)i] + 1; p < frp[(size_t)i + 1]; ++p) lb = std::max(lb, lev_b[(size_t)fci[(size_t)p]] + 1);
        lev_b[(size_t)i] = lb;
    }
    if (F->max_row > ilu::kMaxRow)
        return set_error(MGPU_ERR_UNSUPPORTED, "a factor row has more than " + std::to_string(ilu::kMaxRow) + " entries");
    std::vector<int32_t> rows_f, rows_b;
    ilu::schedule(lev_f, rows_f, F->lev_f);
    ilu::schedule(lev_b, rows_b, F->lev_b);
    mgpu_status rc;
    if ((rc = ilu::dev_upload(F.get(), frp, &F->rp)) || (rc = ilu::dev_upload(F.get(), fci, &F->ci)) ||
        (rc = ilu::dev_upload(F.get(), diag, &F->diag)) || (rc = ilu::dev_upload(F.get(), pos, &F->pos)) ||
        (rc = ilu::dev_upload(F.get(), rows_f, &F->rows_f)) || (rc = ilu::dev_upload(F.get(), rows_b, &F->rows_b)))
        return rc;
    void* p = nullptr;
    if (F->nnz > 0) {
        ILU_TRY(hipMalloc(&p, (size_t)F->nnz * sizeof(double)));
        F->allocs.push_back(p);
        F->lu = static_cast<double*>(p);
    }
    ILU_TRY(hipMalloc(&p, sizeof(int32_t)));
    F->allocs.push_back(p);
    F->err = static_cast<int32_t*>(p);
    cleanup.f = nullptr;
    *out = F.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_ilu_factor(mgpu_ilu* f, const double* vals, void* stream) {
    MGPU_GUARD_BEGIN
    if (!f) return set_error(MGPU_ERR_INVALID_ARG, "null factor");
    if (f->nnz_in > 0 && !vals) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    f->factored = false;
    if (f->n == 0) {
        f->factored = true;
        return MGPU_OK;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 t(256), g(1024);
    hipLaunchKernelGGL(ilu::k_ilu_clear, dim3(1), dim3(64), 0, s, f->err);
    hipLaunchKernelGGL(ilu::k_ilu_fill, g, t, 0, s, f->nnz, f->lu);
    if (f->nnz_in > 0) hipLaunchKernelGGL(ilu::k_ilu_scatter, g, t, 0, s, f->nnz_in, f->pos, vals, f->lu);
    const size_t lds = (size_t)f->max_row * (sizeof(double) + sizeof(int32_t));
    for (size_t l = 0; l + 1 < f->lev_f.size(); ++l) {
        const int32_t cnt = f->lev_f[l + 1] - f->lev_f[l];
        hipLaunchKernelGGL(ilu::k_ilu_row, dim3((unsigned)cnt), dim3(64), lds, s, f->rows_f + f->lev_f[l], f->rp, f->ci,
                           f->diag, f->lu, f->err, f->max_row);
    }
    ILU_TRY(hipGetLastError());
    int32_t err = 0;
    ILU_TRY(hipMemcpyAsync(&err, f->err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    ILU_TRY(hipStreamSynchronize(s));
    if (err) return set_error(MGPU_ERR_INVALID_ARG, "zero pivot in the incomplete factorisation");
    f->factored = true;
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_ilu_apply(const mgpu_ilu* f, const double* v, double* z, void* stream) {
    MGPU_GUARD_BEGIN
    if (!f) return set_error(MGPU_ERR_INVALID_ARG, "null factor");
    if (!f->factored) return set_error(MGPU_ERR_STATE, "the factor has not been computed");
    if (f->n == 0) return MGPU_OK;
    if (!v || !z) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (v != z) hipLaunchKernelGGL(ilu::k_ilu_copy, dim3(1024), dim3(256), 0, s, f->n, v, z);
    ilu_apply_launch(f, z, nullptr, s);
    ILU_TRY(hipGetLastError());
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_ilu_factors(const mgpu_ilu* f, int32_t* n, int64_t* nnz, const int32_t** row_ptr,
                                        const int32_t** col_idx, const double** lu, int32_t* n_levels) {
    if (!f) return set_error(MGPU_ERR_INVALID_ARG, "null factor");
    if (n) *n = f->n;
    if (nnz) *nnz = f->nnz;
    if (row_ptr) *row_ptr = f->rp;
    if (col_idx) *col_idx = f->ci;
    if (lu) *lu = f->lu;
    if (n_levels) {
        n_levels[0] = f->lev_f.empty() ? 0 : (int32_t)f->lev_f.size() - 1;
        n_levels[1] = f->lev_b.empty() ? 0 : (int32_t)f->lev_b.size() - 1;
    }
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_ilu_destroy(mgpu_ilu* f) {
    if (!f) return MGPU_OK;
    (void)hipDeviceSynchronize();
    for (void* p : f->allocs) (void)hipFree(p);
    delete f;
    return MGPU_OK;
}
