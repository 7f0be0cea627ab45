// gmres.hip -- restarted GMRES with a right Jacobi preconditioner on an assembled CSR system in
// device memory, for the non-symmetric systems of the UNSYMMETRIC Nitsche side sets.  Target:
// gfx950 (MI355X), 64-wide wavefronts.
//
// The reference's default linear solver is AztecOO GMRES (Linear_Solver_Aztec::solve_linear_system,
// projects/SOL/DLA/src/cl_DLA_Linear_Solver_Aztec.cpp:88-137; AZ_gmres with AZ_kspace the restart
// length and AZ_precond among its options, :260-356) with Aztec's default convergence test
// ||r_k|| <= rtol ||r_0||, r_0 the residual of the initial guess.  With the preconditioner on the
// right (K D^-1 y = b, x = D^-1 y) the Arnoldi residual is the true one.
//
// One Arnoldi step j is a handful of kernels on the caller's stream, every scalar on the device:
//   k_gm_spmv       w = K D^-1 v_j (L lanes per row)
//   k_gm_dots       block partials of w . v_i, i <= j, eight projections per pass over w
//   k_gm_hcol       one block per i: fixed-order sum -> h_i (added into column j of H)
//   k_gm_orth       w -= sum_i h_i v_i, block partials of w . w
//   (k_gm_dots, k_gm_hcol, k_gm_orth twice: classical Gram-Schmidt with one reorthogonalisation)
//   k_gm_givens     one thread: h_{j+1,j} = ||w||, the previous Givens rotations applied to column j,
//                   the new rotation, the residual estimate |g_{j+1}|, the stopping test
//   k_gm_scale      v_{j+1} = w / h_{j+1,j}
// At the end of a cycle k_gm_solve_y (one thread) solves the triangular system for y and
// k_gm_update adds D^-1 V y to x.  Every kernel returns at once when the cycle or the solve has
// ended.  Partial sums come from a fixed grid and are reduced in a fixed order, so results are
// bit-identical run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../../include/moris_gpu_solve.h"
#include "../setup/internal.hpp"
#include "ilu.hpp"

#define GM_TRY(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace mgpu {
namespace gm {

constexpr int kThreads = 256;
constexpr int kBlocks = 1024;      // fixed grid of the vector kernels (4 blocks per CU)
constexpr int kMaxRestart = 64;

struct State {
    double H[kMaxRestart + 1][kMaxRestart];   // Hessenberg matrix of the cycle (rotated in place)
    double cs[kMaxRestart], sn[kMaxRestart];
    double g[kMaxRestart + 1];
    double hcol[kMaxRestart + 1];              // scratch: this pass's projections
    double y[kMaxRestart];
    double r0;         // ||b - K x_0|| of the initial guess
    double res;        // current residual estimate
    double rtol;
    double res_true;   // ||b - K x|| at the end
    double inv_h;      // 1 / h_{j+1,j} for k_gm_scale
    int32_t m;         // restart length
    int32_t j;         // columns of the current cycle
    int32_t cycle_end; // 1: the cycle stops (converged, limit, breakdown)
    int32_t done;      // 0 running, 1 converged, 2 iteration limit
    int32_t iter;
    int32_t max_iter;
};

__device__ inline double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(kThreads) void k_gm_dinv(int32_t n, const int32_t* rp, const int32_t* ci,
                                                      const double* vals, double* dinv) {
    const int32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    int32_t lo = rp[r], hi = rp[r + 1];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (ci[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    const double d = (lo < rp[r + 1] && ci[lo] == r) ? vals[lo] : 0.0;
    dinv[r] = d != 0.0 ? 1.0 / d : 1.0;   // a zero diagonal is left unscaled
}

// y = b - K x (RESID), y = K D^-1 x (JAC) or y = K x; partial[blk] = block sum of y.y (RESID only).
template <int L, bool RESID, bool JAC>
__global__ __launch_bounds__(kThreads) void k_gm_spmv(int32_t n, const int32_t* rp, const int32_t* ci,
                                                      const double* vals, const double* dinv, const double* x,
                                                      const double* b, double* y, double* partial, const State* st) {
    __shared__ double sh[kThreads];
    if (!RESID && st->cycle_end) return;
    constexpr int G = kThreads / L;
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    double dot = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * G + g; row < n; row += (int64_t)gridDim.x * G) {
        const int32_t r = (int32_t)row;
        double s = 0.0;
        for (int32_t k = rp[r] + lane; k < rp[r + 1]; k += L) {
            const int32_t c = ci[k];
            s += (RESID || !JAC) ? vals[k] * x[c] : vals[k] * (dinv[c] * x[c]);
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) {
            const double v = RESID ? b[r] - s : s;
            y[r] = v;
            dot += v * v;
        }
    }
    if (RESID) {
        const double t = block_sum(dot, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = t;
    }
}

// Cycle start: beta = ||r|| from the partials (fixed order); the first cycle records r0; the
// stopping test; v_0 scaled by 1 / beta in k_gm_scale.
__global__ __launch_bounds__(kThreads) void k_gm_start(const double* partial, State* st, int first) {
    __shared__ double sh[kThreads];
    double a = 0.0;
    for (int k = threadIdx.x; k < kBlocks; k += kThreads) a += partial[k];
    const double beta = std::sqrt(block_sum(a, sh));
    if (threadIdx.x != 0) return;
    if (first) st->r0 = beta;
    st->res = beta;
    st->j = 0;
    st->g[0] = beta;
    if (beta <= st->rtol * st->r0 || beta == 0.0) {
        st->done = 1;
        st->cycle_end = 1;
        return;
    }
    if (st->iter >= st->max_iter) {
        st->done = 2;
        st->cycle_end = 1;
        return;
    }
    st->cycle_end = 0;
    st->inv_h = 1.0 / beta;
}

// partial[i][blk] = block sum of w . v_i for the chunk i = i0 .. min(i0 + 8, j + 1) - 1 (one pass over
// w for the chunk; the sums in registers).
constexpr int kChunk = 8;
__global__ __launch_bounds__(kThreads) void k_gm_dots(int32_t n, const double* V, int64_t ldv, const double* w,
                                                      double* partial, const State* st, int j, int i0) {
    __shared__ double sh[kThreads];
    if (st->cycle_end) return;
    double acc[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) acc[k] = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) {
        const double wr = w[r];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (i0 + k <= j) acc[k] += wr * V[(i0 + k) * ldv + r];
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
        if (i0 + k > j) break;
        const double t = block_sum(acc[k], sh);
        if (threadIdx.x == 0) partial[(int64_t)(i0 + k) * kBlocks + blockIdx.x] = t;
        __syncthreads();
    }
}

// Block i: hcol[i] = fixed-order sum of the partials; added into H[i][j].
__global__ __launch_bounds__(kThreads) void k_gm_hcol(const double* partial, State* st) {
    __shared__ double sh[kThreads];
    if (st->cycle_end) return;
    const int i = blockIdx.x;
    if (i > st->j) return;
    double a = 0.0;
    for (int k = threadIdx.x; k < kBlocks; k += kThreads) a += partial[(int64_t)i * kBlocks + k];
    const double h = block_sum(a, sh);
    if (threadIdx.x == 0) {
        st->hcol[i] = h;
        st->H[i][st->j] += h;
    }
}

// w -= sum_{i <= j} hcol[i] v_i; partial[blk] = block sum of w . w.
__global__ __launch_bounds__(kThreads) void k_gm_orth(int32_t n, const double* V, int64_t ldv, double* w,
                                                      double* partial, const State* st, int j) {
    __shared__ double sh[kThreads];
    if (st->cycle_end) return;
    double nn = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) {
        double v = w[r];
        for (int i = 0; i <= j; ++i) v -= st->hcol[i] * V[i * ldv + r];
        w[r] = v;
        nn += v * v;
    }
    const double t = block_sum(nn, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Zero column j of H before its two projection passes add into it.
__global__ void k_gm_zero_col(State* st) {
    if (st->cycle_end) return;
    const int i = threadIdx.x;
    if (i <= st->j + 1 && i <= kMaxRestart) st->H[i][st->j] = 0.0;
}

// One thread after the reduction of ||w||^2: h_{j+1,j}, the rotations, the residual estimate, the
// stopping test; j advances.
__global__ __launch_bounds__(kThreads) void k_gm_givens(const double* partial, State* st) {
    __shared__ double sh[kThreads];
    if (st->cycle_end) return;
    double a = 0.0;
    for (int k = threadIdx.x; k < kBlocks; k += kThreads) a += partial[k];
    const double hn = std::sqrt(block_sum(a, sh));
    if (threadIdx.x != 0) return;
    const int j = st->j;
    double (*H)[kMaxRestart] = st->H;
    for (int i = 0; i < j; ++i) {
        const double t = st->cs[i] * H[i][j] + st->sn[i] * H[i + 1][j];
        H[i + 1][j] = -st->sn[i] * H[i][j] + st->cs[i] * H[i + 1][j];
        H[i][j] = t;
    }
    const double hjj = H[j][j];
    const double rr = std::sqrt(hjj * hjj + hn * hn);
    double c = 1.0, s = 0.0;
    if (rr > 0.0) {
        c = hjj / rr;
        s = hn / rr;
    }
    st->cs[j] = c;
    st->sn[j] = s;
    H[j][j] = rr;
    H[j + 1][j] = 0.0;
    st->g[j + 1] = -s * st->g[j];
    st->g[j] = c * st->g[j];
    st->res = std::fabs(st->g[j + 1]);
    st->iter += 1;
    st->j = j + 1;
    st->inv_h = hn > 0.0 ? 1.0 / hn : 0.0;
    if (st->res <= st->rtol * st->r0 || hn == 0.0) {   // converged (hn == 0: the Krylov space holds x)
        st->done = 1;
        st->cycle_end = 1;
    } else if (st->iter >= st->max_iter) {
        st->done = 2;
        st->cycle_end = 1;
    } else if (st->j >= st->m) {
        st->cycle_end = 1;
    }
}

// v *= inv_h (v_{j+1} = w / h_{j+1,j}, or v_0 = r / beta).
__global__ __launch_bounds__(kThreads) void k_gm_scale(int32_t n, double* v, const State* st) {
    if (st->cycle_end) return;
    const double s = st->inv_h;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) v[r] *= s;
}

// One thread: y = R^-1 g over the cycle's j columns.
__global__ void k_gm_solve_y(State* st) {
    const int j = st->j;
    for (int i = j - 1; i >= 0; --i) {
        double v = st->g[i];
        for (int k = i + 1; k < j; ++k) v -= st->H[i][k] * st->y[k];
        st->y[i] = st->H[i][i] != 0.0 ? v / st->H[i][i] : 0.0;
    }
}

// x += D^-1 sum_i y_i v_i.
__global__ __launch_bounds__(kThreads) void k_gm_update(int32_t n, const double* V, int64_t ldv, const double* dinv,
                                                        double* x, const State* st) {
    const int j = st->j;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) {
        double s = 0.0;
        for (int i = 0; i < j; ++i) s += st->y[i] * V[i * ldv + r];
        x[r] += dinv[r] * s;
    }
}

// ILU path: u = sum_i y_i v_i (then u <- M^-1 u and x += u).
__global__ __launch_bounds__(kThreads) void k_gm_vy(int32_t n, const double* V, int64_t ldv, double* u, const State* st) {
    const int j = st->j;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) {
        double s = 0.0;
        for (int i = 0; i < j; ++i) s += st->y[i] * V[i * ldv + r];
        u[r] = s;
    }
}

__global__ __launch_bounds__(kThreads) void k_gm_add(int32_t n, const double* u, double* x) {
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) x[r] += u[r];
}

// ILU path: z = v_j before M^-1 is applied to it in place (nothing once the cycle has ended).
__global__ __launch_bounds__(kThreads) void k_gm_copy(int32_t n, const double* v, double* z, const State* st) {
    if (st->cycle_end) return;
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += (int64_t)gridDim.x * kThreads) z[r] = v[r];
}

__global__ __launch_bounds__(kThreads) void k_gm_final(const double* partial, State* st) {
    __shared__ double sh[kThreads];
    double a = 0.0;
    for (int k = threadIdx.x; k < kBlocks; k += kThreads) a += partial[k];
    const double t = std::sqrt(block_sum(a, sh));
    if (threadIdx.x == 0) st->res_true = t;
}

int lanes_for(int32_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / n : 0.0;
    int L = 4;
    while (L < 64 && 4.0 * L < avg) L *= 2;
    return L;
}

template <bool RESID, bool JAC>
void launch_spmv_t(int L, int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* dinv,
                   const double* x, const double* b, double* y, double* partial, const State* st, hipStream_t s) {
    const dim3 g(kBlocks), t(kThreads);
    switch (L) {
        case 4: hipLaunchKernelGGL((k_gm_spmv<4, RESID, JAC>), g, t, 0, s, n, rp, ci, vals, dinv, x, b, y, partial, st); break;
        case 8: hipLaunchKernelGGL((k_gm_spmv<8, RESID, JAC>), g, t, 0, s, n, rp, ci, vals, dinv, x, b, y, partial, st); break;
        case 16: hipLaunchKernelGGL((k_gm_spmv<16, RESID, JAC>), g, t, 0, s, n, rp, ci, vals, dinv, x, b, y, partial, st); break;
        case 32: hipLaunchKernelGGL((k_gm_spmv<32, RESID, JAC>), g, t, 0, s, n, rp, ci, vals, dinv, x, b, y, partial, st); break;
        default: hipLaunchKernelGGL((k_gm_spmv<64, RESID, JAC>), g, t, 0, s, n, rp, ci, vals, dinv, x, b, y, partial, st); break;
    }
}

// b - K x; K D^-1 x (Jacobi) or K x (the ILU path applies its factors before)
template <bool RESID>
void launch_spmv(int L, int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* dinv,
                 const double* x, const double* b, double* y, double* partial, const State* st, hipStream_t s,
                 bool jacobi = true) {
    if (RESID || jacobi) launch_spmv_t<RESID, true>(L, n, rp, ci, vals, dinv, x, b, y, partial, st, s);
    else launch_spmv_t<RESID, false>(L, n, rp, ci, vals, dinv, x, b, y, partial, st, s);
}

}  // namespace gm
}  // namespace mgpu

using namespace mgpu;

namespace {

// Restarted GMRES, right-preconditioned by Jacobi (pc == nullptr) or by incomplete LU factors.
mgpu_status gmres_run(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* b, double* x,
                      double rtol, int32_t restart, int32_t max_iter, const mgpu_ilu* pc, void* stream,
                      mgpu_cg_info* info) {
    if (n < 0) return set_error(MGPU_ERR_INVALID_ARG, "negative size");
    if (n > 0 && (!rp || !ci || !vals || !b || !x)) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    if (!(rtol >= 0.0)) return set_error(MGPU_ERR_INVALID_ARG, "rtol must be >= 0");
    if (restart < 1 || restart > gm::kMaxRestart)
        return set_error(MGPU_ERR_INVALID_ARG, "restart must be in 1 .. " + std::to_string(gm::kMaxRestart));
    if (info) *info = mgpu_cg_info{0, 1, 0.0, 0.0};
    if (n == 0) return MGPU_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t ldv = ((int64_t)n + 31) / 32 * 32;
    // work: V [(restart + 1) ldv], dinv [ldv], z [ldv] (ILU), partials [(restart + 1) kBlocks], the state
    const size_t n_doubles = (size_t)(restart + 3) * ldv + (size_t)(gm::kMaxRestart + 1) * gm::kBlocks;
    void* base = nullptr;
    void* stp = nullptr;
    GM_TRY(hipMalloc(&base, n_doubles * sizeof(double)));
    hipError_t e = hipMalloc(&stp, sizeof(gm::State));
    if (e != hipSuccess) {
        (void)hipFree(base);
        return set_error(MGPU_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    struct Free {
        void* a;
        void* b;
        hipStream_t s;
        ~Free() {
            (void)hipStreamSynchronize(s);
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } guard{base, stp, s};
    double* V = static_cast<double*>(base);
    double* dinv = V + (size_t)(restart + 1) * ldv;
    double* zb = dinv + ldv;
    double* partial = zb + ldv;
    const bool jac = pc == nullptr;
    gm::State* st = static_cast<gm::State*>(stp);
    int32_t nnz = 0;
    GM_TRY(hipMemcpyAsync(&nnz, rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GM_TRY(hipStreamSynchronize(s));
    const int L = gm::lanes_for(n, nnz);
    auto init = std::make_unique<gm::State>();
    std::memset(init.get(), 0, sizeof(gm::State));
    init->rtol = rtol;
    init->m = restart;
    init->max_iter = max_iter;
    GM_TRY(hipMemcpyAsync(st, init.get(), sizeof(gm::State), hipMemcpyHostToDevice, s));
    GM_TRY(hipStreamSynchronize(s));
    const dim3 t(gm::kThreads), gv(gm::kBlocks);
    hipLaunchKernelGGL(gm::k_gm_dinv, dim3((unsigned)((n + gm::kThreads - 1) / gm::kThreads)), t, 0, s, n, rp, ci, vals,
                       dinv);
    int32_t done = 0, iters = 0;
    for (int cycle = 0; !done; ++cycle) {
        // v_0 = r = b - K x, beta = ||r||
        gm::launch_spmv<true>(L, n, rp, ci, vals, dinv, x, b, V, partial, st, s);
        hipLaunchKernelGGL(gm::k_gm_start, dim3(1), t, 0, s, partial, st, cycle == 0 ? 1 : 0);
        hipLaunchKernelGGL(gm::k_gm_scale, gv, t, 0, s, n, V, st);
        for (int j = 0; j < restart; ++j) {
            double* vj = V + (size_t)j * ldv;
            double* w = V + (size_t)(j + 1) * ldv;
            if (jac) {
                gm::launch_spmv<false>(L, n, rp, ci, vals, dinv, vj, nullptr, w, partial, st, s);
            } else {
                // w = K M^-1 v_j
                hipLaunchKernelGGL(gm::k_gm_copy, gv, t, 0, s, n, vj, zb, st);
                ilu_apply_launch(pc, zb, &st->cycle_end, s);
                gm::launch_spmv<false>(L, n, rp, ci, vals, dinv, zb, nullptr, w, partial, st, s, false);
            }
            hipLaunchKernelGGL(gm::k_gm_zero_col, dim3(1), dim3(gm::kMaxRestart + 1), 0, s, st);
            for (int pass = 0; pass < 2; ++pass) {
                for (int i0 = 0; i0 <= j; i0 += gm::kChunk)
                    hipLaunchKernelGGL(gm::k_gm_dots, gv, t, 0, s, n, V, ldv, w, partial, st, j, i0);
                hipLaunchKernelGGL(gm::k_gm_hcol, dim3((unsigned)(j + 1)), t, 0, s, partial, st);
                hipLaunchKernelGGL(gm::k_gm_orth, gv, t, 0, s, n, V, ldv, w, partial, st, j);
            }
            hipLaunchKernelGGL(gm::k_gm_givens, dim3(1), t, 0, s, partial, st);
            hipLaunchKernelGGL(gm::k_gm_scale, gv, t, 0, s, n, w, st);
        }
        hipLaunchKernelGGL(gm::k_gm_solve_y, dim3(1), dim3(1), 0, s, st);
        if (jac) {
            hipLaunchKernelGGL(gm::k_gm_update, gv, t, 0, s, n, V, ldv, dinv, x, st);
        } else {
            // x += M^-1 V y
            hipLaunchKernelGGL(gm::k_gm_vy, gv, t, 0, s, n, V, ldv, zb, st);
            ilu_apply_launch(pc, zb, nullptr, s);
            hipLaunchKernelGGL(gm::k_gm_add, gv, t, 0, s, n, zb, x);
        }
        GM_TRY(hipGetLastError());
        GM_TRY(hipMemcpyAsync(&done, &st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GM_TRY(hipMemcpyAsync(&iters, &st->iter, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GM_TRY(hipStreamSynchronize(s));
        if (!done && iters >= max_iter) break;
    }
    // true residual of the result
    gm::launch_spmv<true>(L, n, rp, ci, vals, dinv, x, b, V, partial, st, s);
    hipLaunchKernelGGL(gm::k_gm_final, dim3(1), t, 0, s, partial, st);
    GM_TRY(hipGetLastError());
    double r0 = 0.0, res = 0.0, res_true = 0.0;
    GM_TRY(hipMemcpyAsync(&r0, &st->r0, sizeof(double), hipMemcpyDeviceToHost, s));
    GM_TRY(hipMemcpyAsync(&res, &st->res, sizeof(double), hipMemcpyDeviceToHost, s));
    GM_TRY(hipMemcpyAsync(&res_true, &st->res_true, sizeof(double), hipMemcpyDeviceToHost, s));
    GM_TRY(hipMemcpyAsync(&iters, &st->iter, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GM_TRY(hipMemcpyAsync(&done, &st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GM_TRY(hipStreamSynchronize(s));
    if (info) {
        info->iterations = iters;
        info->converged = done == 1 ? 1 : 0;
        info->rel_residual = r0 > 0.0 ? res / r0 : 0.0;
        info->true_rel_residual = r0 > 0.0 ? res_true / r0 : 0.0;
    }
    return MGPU_OK;
}

}  // namespace

extern "C" mgpu_status mgpu_gmres_solve(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals,
                                        const double* b, double* x, double rtol, int32_t restart, int32_t max_iter,
                                        void* stream, mgpu_cg_info* info) {
    MGPU_GUARD_BEGIN
    return gmres_run(n, rp, ci, vals, b, x, rtol, restart, max_iter, nullptr, stream, info);
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_gmres_solve_ilu(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals,
                                            const double* b, double* x, double rtol, int32_t restart,
                                            int32_t max_iter, const mgpu_ilu* ilu, void* stream, mgpu_cg_info* info) {
    MGPU_GUARD_BEGIN
    if (!ilu) return set_error(MGPU_ERR_INVALID_ARG, "null factor");
    if (ilu->n != n) return set_error(MGPU_ERR_INVALID_ARG, "the factor is of another size");
    if (!ilu->factored) return set_error(MGPU_ERR_STATE, "the factor has not been computed");
    return gmres_run(n, rp, ci, vals, b, x, rtol, restart, max_iter, ilu, stream, info);
    MGPU_GUARD_END
}
