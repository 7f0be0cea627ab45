/*
 * moris_gpu_learned.h -- learned cut-cell quadrature of libmoris_gpu.so.
 *
 * These entry points have no counterpart in the reference: MORIS integrates cut cells only by XTK
 * subdivision (auto rules TRI_7 / TET_11 / TET_35, projects/FEM/INT/src/CORE/cl_FEM_Set.cpp:2983-3100).
 * They add a small network that predicts weights on the HEX_3x3x3 Gauss points of a hex8 IP cell
 * from its 8 nodal level-set values, a moment fit that makes the rule reproduce the moments of
 * degree <= 4 (a, b, c <= 2) of the element's material region, and a per-element fallback to the
 * subdivision when the fit residual or the weights' cancellation exceeds its bound.  Same
 * conventions as moris_gpu.h.
 */
#ifndef MORIS_GPU_LEARNED_H
#define MORIS_GPU_LEARNED_H

#include "moris_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgpu_mlp mgpu_mlp;   /* 8 -> 64 -> 64 -> 27 SiLU network, FP32 master weights */

/* Level-set families of the training data (bit mask). */
#define MGPU_LQ_PLANES   1
#define MGPU_LQ_SPHERES  2
#define MGPU_LQ_GYROIDS  4

/* Training data: n samples of seeded random cut cells decomposed by the XTK restatement.
   x[n][8]: nodal level-set values / max |phi|; y[n][27]: target weights / 8 (the unique rule on
   the 27 points that reproduces all 27 tri-quadratic moments of the material region, integrated
   with a degree-7 Grundmann-Moeller rule on the sub-tets); moments[n][27] (optional): those
   moments in parametric units (monomials xi^a eta^b zeta^c, a, b, c <= 2, a fastest).           */
mgpu_status mgpu_lq_dataset(int64_t n, uint64_t seed, int32_t kinds, float* x, float* y, double* moments);
/* Fit operators: V[23][27] (monomials at the points), P[27][23] = V^T (V V^T)^-1, Gauss weights[27]. */
mgpu_status mgpu_lq_operators(double* V, double* P, double* gauss_w);
/* The degree-7 Grundmann-Moeller tetrahedron rule: barycentric points [35][4], weights [35]. */
mgpu_status mgpu_lq_gm7(double* lambda, double* w);
/* Moments m27 of the `phase` region of one unit cell [-1,1]^3 from the level set at its 15
   regular-subdivision vertices (corners in hex8 order, faces 0..5, centre). */
mgpu_status mgpu_lq_cell_moments(const double* phi15, int32_t phase, double* m27, int32_t* n_sub);
/* Host version of the moment fit (same algorithm as the device kernel). */
mgpu_status mgpu_lq_fit(const double* m, const double* w0, int32_t iters, double* w, double* residual);

/* Network life cycle, parameters (layout W1[8][64] b1 W2[64][64] b2 W3[64][27] b3, FP32). */
mgpu_status mgpu_mlp_create(uint64_t seed, mgpu_mlp** out);
mgpu_status mgpu_mlp_destroy(mgpu_mlp* m);
mgpu_status mgpu_mlp_num_params(int32_t* n);
mgpu_status mgpu_mlp_get_params(const mgpu_mlp* m, float* host);
mgpu_status mgpu_mlp_set_params(mgpu_mlp* m, const float* host);
/* Inference on host arrays (BF16 MFMA, FP32 accumulation): y[n][27] for x[n][8]. */
mgpu_status mgpu_mlp_forward(const mgpu_mlp* m, const float* x, int64_t n, float* y);
/* Mean-squared-error loss and parameter gradients on n samples (no update). */
mgpu_status mgpu_mlp_gradients(mgpu_mlp* m, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* Adam training with the project's HIP forward/backward kernels; history[epochs] = epoch loss. */
mgpu_status mgpu_mlp_train(mgpu_mlp* m, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                           float lr, uint64_t seed, float* history);

/* Attach a network to a 3D hex8 set (the set descriptor must carry vertex_phi); then
   mgpu_set_options(set, precision, MGPU_QUAD_LEARNED) selects it.  The network must outlive the
   set's assemble calls.  An element keeps its learned rule when the FP32 weights w reproduce the
   moments to ||V w - m|| / ||m|| <= tol and sum |w| <= kappa * volume (kappa >= 1 bounds the
   cancellation, and with it the FP32 rounding of the 27-point products relative to K_e); iters > 0
   alternates the fit with the nonnegative orthant.  Typical: tol 1e-6, iters 0, kappa 4.       */
mgpu_status mgpu_set_quadrature_model(mgpu_set* set, const mgpu_mlp* mlp, double tol, int32_t iters, double kappa);
/* Last learned assemble: w_pred[n_cut][27] (network output / 8), w_fit[n_cut][27] (parametric
   weights used), flags[n_cut] (1 learned, 0 subdivision fallback).  Any pointer may be NULL.  */
mgpu_status mgpu_set_lq_result(const mgpu_set* set, float* w_pred, float* w_fit, int32_t* flags);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_LEARNED_H */
