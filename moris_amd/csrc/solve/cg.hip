// cg.hip -- Jacobi-preconditioned conjugate gradients on an assembled CSR system in device memory,
// and strong Dirichlet elimination.  Target: gfx950 (MI355X), 64-wide wavefronts.
//
// The reference solves K u = R after fill_matrix_and_RHS with AztecOO (Linear_Solver_Aztec::
// solve_linear_system, projects/SOL/DLA/src/cl_DLA_Linear_Solver_Aztec.cpp:88-137; AZ_cg /
// AZ_Jacobi are among its AZ_solver / AZ_precond options, :260-356; convergence test
// ||r_k|| <= tol ||r_0||, Aztec's default).  One iteration here is five kernels on the caller's
// stream, every scalar kept on the device:
//   k_spmv<L, kIter>   q = K p (L lanes per row, sorted columns, rows in a grid-stride order) and
//                      block partials of p.q
//   k_reduce<kAlpha>   one block sums the partials in a fixed order: alpha = rho / p.q
//   k_update           x += alpha p, r -= alpha q, z = D^-1 r, block partials of r.z and r.r
//   k_reduce<kBeta>    rho' = r.z, convergence test on r.r, beta = rho' / rho
//   k_pupdate          p = z + beta p
// Every kernel returns at once when the state says the solve has ended, so the host enqueues
// iterations in batches and reads the state once per batch.  Partial sums come from a fixed grid
// and are reduced in a fixed order: results are bit-identical run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../../include/moris_gpu_solve.h"
#include "../setup/internal.hpp"

#define CG_TRY(expr)                                                                               \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

namespace mgpu {
namespace cg {

constexpr int kThreads = 256;
constexpr int kBlocks = 2048;      // fixed grid of the vector kernels (8 blocks per CU)
constexpr int kBatch = 16;         // iterations enqueued between two reads of the state

// Solver state on the device (doubles first).
struct State {
    double rr0;     // ||r_0||^2
    double rho;     // r.z of the current iterate
    double alpha, beta;
    double rr;      // ||r_k||^2
    double rtol2;   // rtol^2
    double rr_true; // ||b - K x||^2 at the end
    int32_t done;   // 0 running, 1 converged, 2 iteration limit, 3 breakdown (p.Kp <= 0)
    int32_t iter;
    int32_t max_iter;
    int32_t pad;
};

// Work space of one solve: five vectors of n doubles, the partial sums, the state.
struct Work {
    double* r;
    double* z;
    double* p;
    double* q;
    double* dinv;
    double* partial;   // [2][kBlocks]
    State* st;
};

constexpr size_t work_doubles(int64_t n) {
    // five 256-byte aligned vectors, the partial sums, the state
    return 5 * (((size_t)n + 31) / 32 * 32) + 2 * kBlocks + (sizeof(State) + 7) / 8 + 32;
}

inline Work carve(double* base, int64_t n) {
    const size_t vec = ((size_t)n + 31) / 32 * 32;
    Work w;
    w.r = base;
    w.z = base + vec;
    w.p = base + 2 * vec;
    w.q = base + 3 * vec;
    w.dinv = base + 4 * vec;
    w.partial = base + 5 * vec;
    w.st = reinterpret_cast<State*>(base + 5 * vec + 2 * kBlocks);
    return w;
}

// Deterministic block sum of v over kThreads threads; the result is valid in thread 0.
__device__ inline double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// inverse diagonal (1 where the diagonal is missing or not positive: such a row makes K indefinite
// and the solve reports a breakdown)
__global__ __launch_bounds__(kThreads) void k_diag_inv(int32_t n, const int32_t* rp, const int32_t* ci,
                                                       const double* vals, double* dinv) {
    const int32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    int32_t lo = rp[r], hi = rp[r + 1];
    while (lo < hi) {   // sorted columns: binary search for the diagonal
        const int32_t m = (lo + hi) >> 1;
        if (ci[m] < r) lo = m + 1;
        else hi = m;
    }
    const double d = (lo < rp[r + 1] && ci[lo] == r) ? vals[lo] : 0.0;
    dinv[r] = d > 0.0 ? 1.0 / d : 1.0;
}

enum { kIter = 0, kResidual = 1 };

// MODE kIter: y = K x, partial[blk] = sum x.y.   MODE kResidual: y = b - K x, partial = sum y.y.
// L lanes per row; each lane sums every L-th entry of the row, then a butterfly within the L lanes
// (a fixed order).
template <int L, int MODE>
__global__ __launch_bounds__(kThreads) void k_spmv(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals,
                                                   const double* x, const double* b, double* y, double* partial,
                                                   const State* st) {
    __shared__ double sh[kThreads];
    if (MODE == kIter && st->done) return;
    constexpr int G = kThreads / L;
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    double dot = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * G + g; row < n; row += (int64_t)gridDim.x * G) {
        const int32_t r = (int32_t)row;
        double s = 0.0;
        for (int32_t k = rp[r] + lane; k < rp[r + 1]; k += L) s += vals[k] * x[ci[k]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) {
            if (MODE == kIter) {
                y[r] = s;
                dot += x[r] * s;
            } else {
                const double v = b[r] - s;
                y[r] = v;
                dot += v * v;
            }
        }
    }
    const double t = block_sum(dot, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

enum { kInit = 0, kAlpha = 1, kBeta = 2, kFinal = 3 };

// One block: fixed-order sums of the two partial streams, then the scalar step of OP.
template <int OP>
__global__ __launch_bounds__(kThreads) void k_reduce(const double* partial, State* st) {
    __shared__ double sh[kThreads];
    if ((OP == kAlpha || OP == kBeta) && st->done) return;
    double a = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < kBlocks; k += kThreads) {
        a += partial[k];
        c += partial[kBlocks + k];
    }
    const double s0 = block_sum(a, sh);
    __syncthreads();
    const double s1 = block_sum(c, sh);
    if (threadIdx.x != 0) return;
    if (OP == kInit) {          // s0 = r0.r0 (residual stream), s1 = r0.z0
        st->rr0 = s0;
        st->rr = s0;
        st->rho = s1;
        st->iter = 0;
        st->done = s0 == 0.0 ? 1 : 0;
    } else if (OP == kAlpha) {  // s0 = p.Kp
        if (!(s0 > 0.0)) {
            st->done = 3;
            return;
        }
        st->alpha = st->rho / s0;
    } else if (OP == kBeta) {   // s0 = r.z, s1 = r.r
        st->iter += 1;
        st->rr = s1;
        if (s1 <= st->rtol2 * st->rr0) {
            st->done = 1;
            return;
        }
        if (st->iter >= st->max_iter) {
            st->done = 2;
            return;
        }
        st->beta = s0 / st->rho;
        st->rho = s0;
    } else {                    // kFinal: s0 = ||b - K x||^2
        st->rr_true = s0;
    }
}

// z0 = D^-1 r0, p0 = z0, partial streams: [0] unused here (r.r came from the residual SpMV), [1] r.z
__global__ __launch_bounds__(kThreads) void k_init_z(int32_t n, const double* r, const double* dinv, double* z,
                                                     double* p, double* partial) {
    __shared__ double sh[kThreads];
    double rz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double zi = dinv[i] * r[i];
        z[i] = zi;
        p[i] = zi;
        rz += r[i] * zi;
    }
    const double t = block_sum(rz, sh);
    if (threadIdx.x == 0) partial[kBlocks + blockIdx.x] = t;
}

// x += alpha p, r -= alpha q, z = D^-1 r; partial streams [0] r.z, [1] r.r
__global__ __launch_bounds__(kThreads) void k_update(int32_t n, double* x, double* r, double* z, const double* p,
                                                     const double* q, const double* dinv, double* partial,
                                                     const State* st) {
    __shared__ double sh[kThreads];
    if (st->done) return;
    const double alpha = st->alpha;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double zi = dinv[i] * ri;
        z[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    const double a = block_sum(rz, sh);
    __syncthreads();
    const double c = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = a;
        partial[kBlocks + blockIdx.x] = c;
    }
}

// p = z + beta p
__global__ __launch_bounds__(kThreads) void k_pupdate(int32_t n, double* p, const double* z, const State* st) {
    if (st->done) return;
    const double beta = st->beta;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        p[i] = z[i] + beta * p[i];
}

// Strong Dirichlet elimination, one thread per row.
__global__ __launch_bounds__(kThreads) void k_dirichlet(int32_t n, const int32_t* rp, const int32_t* ci, double* vals,
                                                        double* rhs, const uint8_t* fixed, const double* value) {
    const int32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    if (fixed[r]) {
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) vals[k] = ci[k] == r ? 1.0 : 0.0;
        rhs[r] = value ? value[r] : 0.0;
        return;
    }
    double s = 0.0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
        const int32_t c = ci[k];
        if (fixed[c]) {
            if (value) s += vals[k] * value[c];
            vals[k] = 0.0;
        }
    }
    rhs[r] -= s;
}

// Lanes per row from the mean row length (a power of two in [4, 64]).
int lanes_for(int32_t n, int64_t nnz) {
    const double avg = n > 0 ? (double)nnz / n : 0.0;
    int L = 4;
    while (L < 64 && 4.0 * L < avg) L *= 2;
    return L;
}

template <int MODE>
void launch_spmv(int L, int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* x,
                 const double* b, double* y, double* partial, const State* st, hipStream_t s) {
    const dim3 g(kBlocks), t(kThreads);
    switch (L) {
        case 4: hipLaunchKernelGGL((k_spmv<4, MODE>), g, t, 0, s, n, rp, ci, vals, x, b, y, partial, st); break;
        case 8: hipLaunchKernelGGL((k_spmv<8, MODE>), g, t, 0, s, n, rp, ci, vals, x, b, y, partial, st); break;
        case 16: hipLaunchKernelGGL((k_spmv<16, MODE>), g, t, 0, s, n, rp, ci, vals, x, b, y, partial, st); break;
        case 32: hipLaunchKernelGGL((k_spmv<32, MODE>), g, t, 0, s, n, rp, ci, vals, x, b, y, partial, st); break;
        default: hipLaunchKernelGGL((k_spmv<64, MODE>), g, t, 0, s, n, rp, ci, vals, x, b, y, partial, st); break;
    }
}

// The solve on a caller-provided work space (base: work_doubles(n) doubles, 256-byte aligned).
mgpu_status solve(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* b, double* x,
                  double rtol, int32_t max_iter, hipStream_t s, double* base, mgpu_cg_info* info) {
    Work w = carve(base, n);
    int32_t nnz_host = 0;
    CG_TRY(hipMemcpyAsync(&nnz_host, rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CG_TRY(hipStreamSynchronize(s));
    const int L = lanes_for(n, nnz_host);
    State init{};
    init.rtol2 = rtol * rtol;
    init.max_iter = max_iter;
    CG_TRY(hipMemcpyAsync(w.st, &init, sizeof(State), hipMemcpyHostToDevice, s));
    const dim3 gr((unsigned)((n + kThreads - 1) / kThreads)), t(kThreads);
    hipLaunchKernelGGL(k_diag_inv, gr, t, 0, s, n, rp, ci, vals, w.dinv);
    // r0 = b - K x0 (stream 0 of the partials: r0.r0), z0 = D^-1 r0, p0 = z0 (stream 1: r0.z0)
    launch_spmv<kResidual>(L, n, rp, ci, vals, x, b, w.r, w.partial, w.st, s);
    hipLaunchKernelGGL(k_init_z, dim3(kBlocks), t, 0, s, n, w.r, w.dinv, w.z, w.p, w.partial);
    hipLaunchKernelGGL((k_reduce<kInit>), dim3(1), t, 0, s, w.partial, w.st);
    CG_TRY(hipGetLastError());
    State h{};
    for (;;) {
        CG_TRY(hipMemcpyAsync(&h, w.st, sizeof(State), hipMemcpyDeviceToHost, s));
        CG_TRY(hipStreamSynchronize(s));
        if (h.done || max_iter <= 0) break;
        for (int k = 0; k < kBatch; ++k) {
            launch_spmv<kIter>(L, n, rp, ci, vals, w.p, nullptr, w.q, w.partial, w.st, s);
            hipLaunchKernelGGL((k_reduce<kAlpha>), dim3(1), t, 0, s, w.partial, w.st);
            hipLaunchKernelGGL(k_update, dim3(kBlocks), t, 0, s, n, x, w.r, w.z, w.p, w.q, w.dinv, w.partial, w.st);
            hipLaunchKernelGGL((k_reduce<kBeta>), dim3(1), t, 0, s, w.partial, w.st);
            hipLaunchKernelGGL(k_pupdate, dim3(kBlocks), t, 0, s, n, w.p, w.z, w.st);
        }
        CG_TRY(hipGetLastError());
    }
    // true residual of the result
    launch_spmv<kResidual>(L, n, rp, ci, vals, x, b, w.q, w.partial, w.st, s);
    hipLaunchKernelGGL((k_reduce<kFinal>), dim3(1), t, 0, s, w.partial, w.st);
    CG_TRY(hipGetLastError());
    CG_TRY(hipMemcpyAsync(&h, w.st, sizeof(State), hipMemcpyDeviceToHost, s));
    CG_TRY(hipStreamSynchronize(s));
    if (info) {
        info->iterations = h.iter;
        info->converged = h.done == 1 ? 1 : 0;
        info->rel_residual = h.rr0 > 0.0 ? std::sqrt(h.rr / h.rr0) : 0.0;
        info->true_rel_residual = h.rr0 > 0.0 ? std::sqrt(h.rr_true / h.rr0) : 0.0;
    }
    if (h.done == 3) return set_error(MGPU_ERR_STATE, "conjugate gradients broke down (p.Kp <= 0: K not positive definite)");
    return MGPU_OK;
}

mgpu_status check_args(int32_t n, const int32_t* rp, const int32_t* ci, const double* vals, const double* b,
                       const double* x, double rtol) {
    if (n < 0) return set_error(MGPU_ERR_INVALID_ARG, "negative size");
    if (n > 0 && (!rp || !ci || !vals || !b || !x)) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    if (!(rtol >= 0.0)) return set_error(MGPU_ERR_INVALID_ARG, "rtol must be >= 0");
    return MGPU_OK;
}

}  // namespace cg
}  // namespace mgpu

// Work space of mgpu_cg_solve_noalloc: one fixed array per GPU (the code object's copy on that
// device), MGPU_CG_NOALLOC_MAX_ROWS rows at most (about 100 MB).  Every part of it is written before
// it is read within a call (residual, search direction, preconditioner, partial sums, state), so
// nothing needs clearing between calls.
namespace {
constexpr size_t kNoallocDoubles = mgpu::cg::work_doubles(MGPU_CG_NOALLOC_MAX_ROWS);
}  // namespace
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
// Device address of `scratch` on each GPU, looked up once per device with that device current.
// In host code the name `scratch` is only a handle for hipGetSymbolAddress; the kernels receive
// the looked-up address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

mgpu_status noalloc_scratch(double** out) {
    int dev = 0;
    CG_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        CG_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return mgpu::set_error(MGPU_ERR_HIP, "hipGetSymbolAddress returned no address for the work space");
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return MGPU_OK;
}
}  // namespace

using namespace mgpu;

extern "C" mgpu_status mgpu_csr_apply_dirichlet(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                                double* vals, double* rhs, const uint8_t* fixed, const double* value,
                                                void* stream) {
    MGPU_GUARD_BEGIN
    if (n_rows < 0) return set_error(MGPU_ERR_INVALID_ARG, "negative size");
    if (n_rows == 0) return MGPU_OK;
    if (!row_ptr || !col_idx || !vals || !rhs || !fixed) return set_error(MGPU_ERR_INVALID_ARG, "null device pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(cg::k_dirichlet, dim3((unsigned)((n_rows + cg::kThreads - 1) / cg::kThreads)),
                       dim3(cg::kThreads), 0, s, n_rows, row_ptr, col_idx, vals, rhs, fixed, value);
    CG_TRY(hipGetLastError());
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_cg_solve(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                     const double* vals, const double* b, double* x, double rtol, int32_t max_iter,
                                     void* stream, mgpu_cg_info* info) {
    MGPU_GUARD_BEGIN
    mgpu_status rc = cg::check_args(n_rows, row_ptr, col_idx, vals, b, x, rtol);
    if (rc != MGPU_OK) return rc;
    if (info) *info = mgpu_cg_info{0, 1, 0.0, 0.0};
    if (n_rows == 0) return MGPU_OK;
    void* base = nullptr;
    CG_TRY(hipMalloc(&base, cg::work_doubles(n_rows) * sizeof(double)));
    rc = cg::solve(n_rows, row_ptr, col_idx, vals, b, x, rtol, max_iter, static_cast<hipStream_t>(stream),
                   static_cast<double*>(base), info);
    (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
    (void)hipFree(base);
    return rc;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_cg_solve_noalloc(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                             const double* vals, const double* b, double* x, double rtol,
                                             int32_t max_iter, void* stream, mgpu_cg_info* info) {
    MGPU_GUARD_BEGIN
    mgpu_status rc = cg::check_args(n_rows, row_ptr, col_idx, vals, b, x, rtol);
    if (rc != MGPU_OK) return rc;
    if (n_rows > MGPU_CG_NOALLOC_MAX_ROWS)
        return set_error(MGPU_ERR_INVALID_ARG, "mgpu_cg_solve_noalloc: " + std::to_string(n_rows) +
                                                   " rows exceed its fixed work space of " +
                                                   std::to_string(MGPU_CG_NOALLOC_MAX_ROWS) + " rows");
    if (info) *info = mgpu_cg_info{0, 1, 0.0, 0.0};
    if (n_rows == 0) return MGPU_OK;
    double* base = nullptr;
    if ((rc = noalloc_scratch(&base)) != MGPU_OK) return rc;
    return cg::solve(n_rows, row_ptr, col_idx, vals, b, x, rtol, max_iter, static_cast<hipStream_t>(stream), base,
                     info);
    MGPU_GUARD_END
}
