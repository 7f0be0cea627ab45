/*
 * moris_gpu_setup.h -- the set-up of a set on the GPU (libmoris_gpu.so): XTK cut generation, MSI
 * adof numbering, HMR T rows, element topology and the CSR pattern, from a problem spec, for the
 * per-design re-meshing of design variants.
 *
 * The same products as mgpu_problem_build + mgpu_set_create (host), computed by kernels:
 *   cut of every intersected background cell   cl_XTK_Integration_Mesh_Generator.cpp:53-173,
 *                                              cl_XTK_Node_Hierarchy_Interface.cpp:160-733 (one
 *                                              thread per intersected cell, the host's per-cell
 *                                              routine common/xtk_cut.hpp)
 *   adof numbering                             cl_MSI_Dof_Manager.cpp:703-745, :820 (used bases in
 *                                              ascending index, by dof type)
 *   T rows (HMR level 0, p = 1 nodal, p = 2)   cl_HMR_T_Matrix_Base.cpp:306-457
 *   sparsity pattern                           cl_MSI_Solver_Interface.cpp:309,
 *                                              cl_Sparse_Matrix_EpetraFECrs.cpp:195-240
 * Scope: HEX8 / QUAD4 IP cells, bspline_order 1 or 2, no refinement (refine_levels 0).  The level
 * set is evaluated on the device without FMA contraction, as on the host: for planes and spheres
 * the arrays equal the host set-up's bit for bit; the gyroid's sine and cosine come from the
 * device math library (last-bit differences in the level set, hence in the intersection points).
 * Same conventions as moris_gpu.h.
 */
#ifndef MORIS_GPU_SETUP_H
#define MORIS_GPU_SETUP_H

#include "moris_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgpu_gpu_setup mgpu_gpu_setup;

/* Build on the stream (hipStream_t as void*, NULL = default); synchronous. */
mgpu_status mgpu_setup_build(const mgpu_problem_spec* spec, void* hip_stream, mgpu_gpu_setup** out);
mgpu_status mgpu_setup_destroy(mgpu_gpu_setup* s);
/* [0] elements, [1] cut elements, [2] IG sub-cells, [3] adofs per type, [4] rows, [5] nnz,
   [6] T entries, [7] bases per element.                                                        */
mgpu_status mgpu_setup_counts(const mgpu_gpu_setup* s, int64_t counts[8]);
/* Milliseconds of the device work of the build (HIP events around it).                         */
mgpu_status mgpu_setup_time(const mgpu_gpu_setup* s, double* ms);
/* Host copies of the device arrays (any pointer may be NULL; sizes from mgpu_setup_counts):
   elem_cell[n_el], ig_ptr[n_el + 1], ig_xi[n_ig][dim + 1][dim], material[n_el][1 | 2],
   ip_box[n_el][2 dim], vertex_phi[n_el][nb], t_ptr[n_el nb + 1], t_adof[n_t], t_weight[n_t],
   topo[n_el][nf na] (na = (p + 1)^dim bases per dof type, 27 on HEX27 cells, ascending; type f offset by f N),
   row_ptr[rows + 1], col_idx[nnz].  The descriptor (host pointers into copies owned by the
   handle, valid until destroy) feeds mgpu_set_create.                                          */
mgpu_status mgpu_setup_copy(const mgpu_gpu_setup* s, int32_t* elem_cell, int32_t* ig_ptr, double* ig_xi,
                            double* material, double* ip_box, double* vertex_phi, int32_t* t_ptr, int32_t* t_adof,
                            double* t_weight, int32_t* topo, int32_t* row_ptr, int32_t* col_idx);
mgpu_status mgpu_setup_desc(mgpu_gpu_setup* s, mgpu_set_desc* desc);
/* A set made directly from the set-up's device arrays, with no host round trip: the same set as
   mgpu_set_create(mgpu_setup_desc(...)) for bspline_order 1 on HEX8 / QUAD4 cells (nodal T rows; others return
   MGPU_ERR_UNSUPPORTED).  The set uses the set-up's device arrays in place, so the set-up must
   outlive the set.  Its host copies of the pattern and topology (mgpu_pattern_get,
   mgpu_element_topology, side sets) are fetched from the device on first use.  No counterpart in
   the reference (its set-up is on the host); for design loops that re-mesh per design.        */
mgpu_status mgpu_set_create_from_setup(mgpu_gpu_setup* s, mgpu_set** out);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_SETUP_H */
