/*
 * moris_gpu_solve.h -- solve of the assembled system on the GPU (libmoris_gpu.so).
 *
 * After fill_matrix_and_RHS the reference solves K u = R with an iterative Trilinos solver,
 * Linear_Solver_Aztec::solve_linear_system (projects/SOL/DLA/src/cl_DLA_Linear_Solver_Aztec.cpp
 * :88-137: AztecOO::Iterate( AZ_max_iter, rel_residual ), Aztec's default convergence test
 * ||r_k|| <= rel_residual * ||r_0||, solver and preconditioner chosen by AZ_solver / AZ_precond,
 * :260-356; AZ_gmres (the default), AZ_cg and AZ_Jacobi among its options).  These entry points
 * solve systems held in device memory on one GPU: conjugate gradients (symmetric positive definite
 * K) and restarted GMRES (any non-singular K), with a Jacobi or (GMRES) an ILU(k) preconditioner, every reduction
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

/* Incomplete LU factors ILU(k) of a CSR matrix on the GPU, the preconditioner the reference pairs
   with GMRES (AZ_precond = AZ_dom_decomp, AZ_subdomain_solve = AZ_ilu, AZ_graph_fill = k;
   cl_DLA_Linear_Solver_Aztec.cpp:353-380, UT_MDL_XTK_HMR_Linear_Struc_2D.cpp:990-993), restated
   from the published algorithm (Saad, Iterative Methods for Sparse Linear Systems, Alg. 10.4 and
   the level-of-fill rule lev(a_ij) = min(lev(a_ij), lev(a_ik) + lev(a_kj) + 1), fill kept up to
   level k) since AztecOO is not part of the reference sources.  Symbolic and numeric parts are
   split as a solver reuses one pattern for many matrices:
     mgpu_ilu_create   the factor pattern (the matrix pattern, every diagonal, the fill of level
                       <= fill_level) and the level schedules, computed on the host from the
                       device pattern (row_ptr, col_idx: sorted columns); synchronous;
     mgpu_ilu_factor   the numeric factors of vals (same pattern as at create), one launch per
                       level of the forward schedule; synchronous; MGPU_ERR_INVALID_ARG on a zero
                       pivot;
     mgpu_ilu_apply    z = (L U)^-1 v (v == z allowed), asynchronous on the stream;
     mgpu_ilu_factors  the factor pattern and values (DEVICE pointers owned by the handle: the
                       strict lower part holds L, whose unit diagonal is implied, the rest U) and
                       the forward / backward level counts (n_levels[2], optional).
   Rows of at most 5000 factor entries.  Results are the same on every call.                   */
typedef struct mgpu_ilu mgpu_ilu;
mgpu_status mgpu_ilu_create(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, int32_t fill_level,
                            void* stream, mgpu_ilu** out);
mgpu_status mgpu_ilu_factor(mgpu_ilu* ilu, const double* vals, void* stream);
mgpu_status mgpu_ilu_apply(const mgpu_ilu* ilu, const double* v, double* z, void* stream);
mgpu_status mgpu_ilu_factors(const mgpu_ilu* ilu, int32_t* n_rows, int64_t* nnz, const int32_t** row_ptr,
                             const int32_t** col_idx, const double** lu, int32_t* n_levels);
mgpu_status mgpu_ilu_destroy(mgpu_ilu* ilu);

/* mgpu_gmres_solve right-preconditioned by factors from mgpu_ilu_factor of the same matrix (or of
   one with the same pattern): K M^-1 y = b, x = M^-1 y, M = L U.  Allocates one more vector.  */
mgpu_status mgpu_gmres_solve_ilu(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                                 const double* b, double* x, double rtol, int32_t restart, int32_t max_iter,
                                 const mgpu_ilu* ilu, void* stream, mgpu_cg_info* info);

/* The same solve without any device allocation.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).
   The work vectors live in a fixed device array sized for MGPU_CG_NOALLOC_MAX_ROWS rows (about
   100 MB), one copy per GPU, held from the library's first use on that GPU until the process
   ends; larger systems are refused with MGPU_ERR_INVALID_ARG.  All calls on one GPU share that
   array: calls that overlap in time on one GPU (several host threads or streams) race and must be
   serialised by the caller.  Results equal those of mgpu_cg_solve bit for bit.                */
#define MGPU_CG_NOALLOC_MAX_ROWS 2500000
mgpu_status mgpu_cg_solve_noalloc(int32_t n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                                  const double* vals, const double* b, double* x, double rtol, int32_t max_iter,
                                  void* stream, mgpu_cg_info* info);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_SOLVE_H */
