/*
 * moris_gpu_solve.h -- solve of the assembled system on the GPU (libmoris_gpu.so).
 *
 * After fill_matrix_and_RHS the reference solves K u = R with an iterative Trilinos solver,
 * Linear_Solver_Aztec::solve_linear_system (projects/SOL/DLA/src/cl_DLA_Linear_Solver_Aztec.cpp
 * :88-137: AztecOO::Iterate( AZ_max_iter, rel_residual ), Aztec's default convergence test
 * ||r_k|| <= rel_residual * ||r_0||, solver and preconditioner chosen by AZ_solver / AZ_precond,
 * :260-356; AZ_gmres (the default), AZ_cg and AZ_Jacobi among its options).  These entry points
 * solve systems held in device memory on one GPU: conjugate gradients (symmetric positive definite
 * K) and restarted GMRES (any non-singular K), both with a Jacobi preconditioner, every reduction
 * in a fixed order (bit-identical results run to run), and the convergence test of Aztec's
 * default.  Strong Dirichlet conditions, which the reference's weak
 * (Nitsche) formulation does not need, can be applied to the system first.  Same conventions as
 * moris_gpu.h; every pointer below is a DEVICE pointer unless said otherwise.
 */
#ifndef MORIS_GPU_SOLVE_H
#define MORIS_GPU_SOLVE_H

#include "moris_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Solver report. */
typedef struct mgpu_cg_info {
    int32_t iterations;        /* iterations taken                                                */
    int32_t converged;         /* 1: ||r_k|| <= rtol ||r_0|| reached                              */
    double  rel_residual;      /* ||r_k|| / ||r_0|| of the recurrence                             */
    double  true_rel_residual; /* ||b - K x|| / ||b - K x_0|| recomputed at the end                */
} mgpu_cg_info;

/* Strong Dirichlet conditions on an assembled system (no counterpart in the reference, which
   imposes them weakly): for every row r with fixed[r] != 0 the row and column of r become those of
   the identity and rhs[r] = value[r]; for every other row, rhs[r] -= sum over fixed columns c of
   K[r][c] * value[c] (value may be NULL: homogeneous).  vals and rhs are modified in place on the
   stream (hipStream_t as void*, NULL = default); fixed[n_rows] (uint8).                         */
mgpu_status mgpu_csr_apply_dirichlet(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                     double* vals, double* rhs, const uint8_t* fixed, const double* value,
                                     void* stream);

/* Jacobi-preconditioned conjugate gradients for K x = b (K symmetric positive definite, CSR with
   sorted columns and a stored diagonal).  x holds the initial guess on entry and the solution on
   return.  Stops when ||r_k|| <= rtol ||r_0|| or after max_iter iterations.  The library
   allocates its work vectors (5 n_rows doubles) for the call; synchronous.                     */
mgpu_status mgpu_cg_solve(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                          const double* b, double* x, double rtol, int32_t max_iter, void* stream,
                          mgpu_cg_info* info);

/* The same solve without any device allocation.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).
   The work vectors live in a fixed device array sized for MGPU_CG_NOALLOC_MAX_ROWS rows (about
   100 MB), one copy per GPU, held from the library's first use on that GPU until the process
   ends; larger systems are refused with MGPU_ERR_INVALID_ARG.  All calls on one GPU share that
   array: calls that overlap in time on one GPU (several host threads or streams) race and must be
   serialised by the caller.  Results equal those of mgpu_cg_solve bit for bit.                */
/* Restarted GMRES(restart) with a right Jacobi preconditioner for K x = b, K any non-singular CSR
   matrix with sorted columns (the non-symmetric systems of UNSYMMETRIC Nitsche side sets): the
   reference's default solver, AztecOO AZ_gmres with AZ_kspace = restart (cl_DLA_Linear_Solver_
   Aztec.cpp:88-137, :260-356).  Classical Gram-Schmidt with one reorthogonalisation, Givens
   rotations; stops when ||r_k|| <= rtol ||r_0|| (r_0 of the initial guess) or after max_iter
   Arnoldi steps; restart in 1 .. 64.  x holds the initial guess on entry and the solution on
   return; info as for CG (iterations = Arnoldi steps).  The library allocates (restart + 2)
   n_rows doubles for the call; synchronous; bit-identical run to run.                           */
mgpu_status mgpu_gmres_solve(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                             const double* b, double* x, double rtol, int32_t restart, int32_t max_iter,
                             void* stream, mgpu_cg_info* info);

#define MGPU_CG_NOALLOC_MAX_ROWS 2500000
mgpu_status mgpu_cg_solve_noalloc(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                  const double* vals, const double* b, double* x, double rtol, int32_t max_iter,
                                  void* stream, mgpu_cg_info* info);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_SOLVE_H */
