// lq_view.hpp -- device view of the learned cut-cell quadrature state of a set, and the launches
// the assembly drives (lq_assemble.hip, mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../set/set_device.hpp"

struct mgpu_mlp;

namespace mgpu {

struct LqView {
    int32_t n_cut;
    int32_t iters;
    double tol;                     // moment residual tolerance
    double kappa;                   // bound on sum |w| / volume
    double detJ_ip;                 // prod(h / 2) of the lattice cell
    const double* vphi;             // [n_cut][8] nodal level-set values of the cut elements
    float* w_pred;                  // [n_cut][27] MLP output (weights / 8)
    double* mom;                    // [n_cut][24] moments of the material sub-cells (23 + padding)
    int32_t n_mom_waves;            // wavefronts of the moment kernel
    const int32_t* mom_owner;       // [n_mom_waves][12]: per 16-lane quarter the cut element (or
                                    // -1), its first sub-cell and its sub-cell count
    float* w_fit;                   // [n_cut][27] fitted parametric weights
    int32_t* flag;                  // [n_cut] 1 learned, 0 fallback
    const double* Vt;               // [27][23]  V^T
    const double* Pt;               // [23][27]  P^T
    const float* g27;               // [27][NG] physical gradients at the 27 points
};

void lq_launch_fit(const DevView& s, const LqView& lv, hipStream_t st);
// MGPU_QUAD_MOMENT_FIT on lattice sets: the 27 weights from the sub-tets' Lagrange moments, every
// element accepted
void lq_launch_mfit8(const DevView& s, const LqView& lv, hipStream_t st);
void lq_launch_element(const DevView& s, const LqView& lv, int nf, bool f32_store, hipStream_t st);
void mlp_forward_launch(const mgpu_mlp* m, const float* x, int64_t n, float* y, hipStream_t st);
// the same with the inputs formed from the cells' 8 nodal level-set values (vphi[n][8])
void mlp_forward_phi_launch(const mgpu_mlp* m, const double* vphi, int64_t n, float* y, hipStream_t st);

}  // namespace mgpu
