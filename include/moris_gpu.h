/*
 * moris_gpu.h -- C ABI of the MI355X element K+R assembly library (libmoris_gpu.so).
 *
 * The library replaces the per-object loop of MORIS's bulk assembly
 *
 *   Solver_Interface::fill_matrix_and_RHS                projects/SOL/DLA/src/cl_DLA_Solver_Interface.cpp:317-365
 *     -> MSI::Equation_Object::get_egn_obj_jacobian_and_residual
 *                                                      projects/FEM/MSI/src/cl_MSI_Equation_Object.cpp:789-842
 *        -> fem::Interpolation_Element::compute_jacobian_and_residual
 *                                                      projects/FEM/INT/src/ELEM/cl_FEM_Interpolation_Element.cpp:632-690
 *     -> Sparse_Matrix_EpetraFECrs::fill_matrix       projects/SOL/SOL_CORE/src/cl_Sparse_Matrix_EpetraFECrs.cpp:103-133
 *
 * for bulk sets whose IWGs are IWG_Diffusion_Bulk or IWG_Isotropic_Struc_Linear_Bulk, with one
 * set-level batched call.  Everything crossing this boundary is plain C: POD structs, raw pointers,
 * sizes.  No call throws; every entry point returns an mgpu_status and mgpu_last_error_string()
 * describes the last failure of the calling thread.
 *
 * Index base is 0, ids are int32 (MORIS_USE_32BIT, CMakeLists.txt:156-157), all reals are FP64.
 */
#ifndef MORIS_GPU_H
#define MORIS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------------
 * status codes.  Replaces MORIS_ERROR (projects/MRS/ASR/src/fn_log_assert.hpp:41,92), which throws.
 * ---------------------------------------------------------------------------------------------- */
typedef int32_t mgpu_status;
#define MGPU_OK                 0
#define MGPU_ERR_INVALID_ARG    1   /* bad descriptor / pointer / size                          */
#define MGPU_ERR_HIP            2   /* a HIP runtime call failed                                */
#define MGPU_ERR_UNSUPPORTED    3   /* combination not implemented (e.g. IWG not in the set)    */
#define MGPU_ERR_NO_DEVICE      4   /* no gfx950 device visible                                 */
#define MGPU_ERR_STATE          5   /* call order violated (e.g. assemble before create)       */
#define MGPU_ERR_GEOMETRY       6   /* level set vertex on the isocontour (|phi| <= 1e-12)      */

/* IWG types (projects/FEM/INT/src/IWG/cl_FEM_IWG_Factory.cpp:153,191). */
#define MGPU_IWG_SPATIALDIFF_BULK   0   /* IWG_Diffusion_Bulk, CM_Diffusion_Linear_Isotropic       */
#define MGPU_IWG_STRUC_LINEAR_BULK  1   /* IWG_Isotropic_Struc_Linear_Bulk, CM_Struc_Linear_Isotropic
                                           (3D full, 2D plane stress)                              */

/* Cell geometries (mtk::Geometry_Type + interpolation order). */
#define MGPU_GEOM_QUAD4  0
#define MGPU_GEOM_HEX8   1
#define MGPU_GEOM_TRI3   2
#define MGPU_GEOM_TET4   3
#define MGPU_GEOM_HEX27  4   /* quadratic Lagrange hex (cl_MTK_Interpolation_Function_Lagrange_Hex27.hpp
                                node order); IP cell of the generic path, HEX_3x3x3 uncut, TET_35 on
                                its TET4 sub-cells (cl_FEM_Set.cpp:3048,3091)                      */

/* Arithmetic of the element integrands. */
#define MGPU_PREC_FP64   0   /* plain path: everything FP64                                        */
#define MGPU_PREC_MIXED  1   /* FP32 integrands with FP32 accumulation; FP64 cut-cell Jacobians and
                                FP64 global accumulation                                           */

/* Cut-cell quadrature. */
#define MGPU_QUAD_SUBDIVISION  0   /* XTK regular subdivision + node hierarchy, TET_11 / TRI_7       */
#define MGPU_QUAD_LEARNED      1   /* MLP weights on the HEX_3x3x3 Gauss points, moment-checked,
                                      falling back to subdivision per element                      */
#define MGPU_QUAD_MOMENT_FIT   2   /* HEX27 cells (generic path): a cut cell's TET_35 points replaced
                                      by the 5 x 5 x 5 Gauss-Legendre nodes, weighted by the cell's
                                      degree-4 Lagrange moments over those points; the same sum as the
                                      subdivision for integrands of degree <= 4 per axis (the box
                                      HEX27 stiffness), i.e. the same K up to rounding, from 125
                                      points per cut cell.  Generic HEX8 / QUAD4 cells: the 3^dim
                                      tensor Gauss points, degree-2 Lagrange moments of the TET_11 /
                                      TRI_7 points.  3D lattice (hex8) sets: the 27 HEX_3x3x3
                                      points weighted by the degree-2 Lagrange moments over the
                                      TET_11 points (exact for the trilinear stiffness), no model,
                                      no fallback; the 27-point product runs as in
                                      MGPU_QUAD_LEARNED (FP32 MFMA)                                  */

/* Level-set kinds (GEN analytic geometries). */
#define MGPU_LS_NONE     0   /* no geometry: every cell is uncut material                           */
#define MGPU_LS_SPHERES  1   /* union of spherical (circular in 2D) holes: phi = min_i(|x-c_i|-r_i) */
#define MGPU_LS_GYROID   2   /* phi = sin x' cos y' + sin y' cos z' + sin z' cos x', x' = s*x       */
#define MGPU_LS_PLANE    3   /* phi = n.x + d                                                        */

/* ------------------------------------------------------------------------------------------------
 * Workload description: a level-0 HMR background grid, one GEN level set, XTK cut, MSI numbering.
 * mgpu_problem_build() runs this set-up on the host and produces the set descriptor below.  It is
 * the input side of the hot path (cl_XTK_Integration_Mesh_Generator.cpp:53-173,
 * cl_MSI_Dof_Manager.cpp:684-990), not timed.
 * ---------------------------------------------------------------------------------------------- */
typedef struct mgpu_problem_spec {
    int32_t dim;               /* 2 or 3                                                         */
    int32_t n[3];              /* cells per axis (n[2] ignored in 2D)                            */
    double  lo[3];             /* domain lower corner                                            */
    double  hi[3];             /* domain upper corner                                            */
    int32_t iwg;               /* MGPU_IWG_*                                                     */
    int32_t ls_kind;           /* MGPU_LS_*                                                      */
    int32_t n_spheres;         /* MGPU_LS_SPHERES: number of spheres                             */
    const double* spheres;     /* MGPU_LS_SPHERES: [n_spheres][4] = cx, cy, cz, r               */
    double  gyroid_scale;      /* MGPU_LS_GYROID: s                                              */
    double  plane[4];          /* MGPU_LS_PLANE: nx, ny, nz, d                                   */
    int32_t material_phase;    /* 1: phi > 0 is material (phase 1), 0: phi < 0 is material      */
    double  E, nu;             /* elasticity                                                     */
    double  k_mu, k_sigma;     /* diffusion: k_e = exp(k_mu + k_sigma * N(0,1)) per element     */
    uint64_t seed;             /* splitmix64 seed of the per-element conductivity                */
    int32_t bspline_order;     /* 1 (or 0): T rows of weight 1 per Lagrange node (p = 1);
                                  2: quadratic B-splines of the lattice evaluated at the Lagrange
                                  nodes (HMR T-matrix at level 0, cl_HMR_T_Matrix_Base.cpp:306-457):
                                  2 splines of weight 1/2 per axis at a lattice node, 3 of weights
                                  1/8, 3/4, 1/8 at a cell-centre coordinate                     */
    int32_t lagrange_order;    /* 1 (or 0): HEX8 / QUAD4 IP cells; 2: HEX27 IP cells (3D only),
                                  whose 27 nodes sit on the half-step grid of the lattice        */
    int32_t refine_levels;     /* 0: level-0 lattice only; 1 (3D, bspline_order 2): one THB level:
                                  every level-0 cell the level set intersects is split into 8
                                  level-1 cells; truncated hierarchical quadratic B-splines
                                  (cl_HMR_T_Matrix_Base.cpp:127-258); such sets have no lattice
                                  hint and take the generic path                                */
    int32_t enrich;            /* 0: one element per material cell (a cut cell's material
                                  sub-cells form one element; neighbouring cells share bases);
                                  1: XTK basis-cluster enrichment (cl_XTK_Enrichment.cpp:276-400,
                                  level-0 lattices): a cut cell's material sub-cells split into
                                  subphases (face-connected components), one element each, and
                                  every basis is replicated once per connected component of the
                                  material subphases in its support, so material pieces that do
                                  not touch inside a basis' support share no adof.  Sets with a
                                  replicated basis or two elements in one cell take the generic
                                  path.                                                          */
    int32_t ghost;             /* 1: the set descriptor couples the two elements of every ghost
                                  facet (mgpu_problem_facets with MGPU_SIDE_GHOST) in its pattern,
                                  so that a ghost side set can be assembled into it; such sets
                                  take the generic path (level-0 lattices, HEX8 / QUAD4, without
                                  enrichment)                                                    */
} mgpu_problem_spec;

/* Uniform lattice the elements of a set sit on (HMR level 0).  n[0] == 0: no lattice. */
typedef struct mgpu_lattice {
    int32_t n[3];              /* cells per axis                                                 */
    double  lo[3];             /* lower corner                                                   */
    double  h[3];              /* cell size                                                      */
} mgpu_lattice;

/* ------------------------------------------------------------------------------------------------
 * Set descriptor: element-major POD arrays of one or more bulk blocks (uncut *_n_p and cut *_c_p)
 * assembled into one matrix.  Mirrors what fem::Set holds per equation object
 * (projects/FEM/INT/src/CORE/cl_FEM_Set.hpp:64; clusters cl_FEM_Set.cpp:267-320).
 * ---------------------------------------------------------------------------------------------- */
typedef struct mgpu_set_desc {
    int32_t spatial_dim;       /* 2 | 3                                                          */
    int32_t iwg;               /* MGPU_IWG_*                                                     */
    int32_t ip_geometry;       /* MGPU_GEOM_QUAD4 | MGPU_GEOM_HEX8 | MGPU_GEOM_HEX27 (IP cell)    */
    int32_t ig_geometry;       /* MGPU_GEOM_TRI3 | MGPU_GEOM_TET4 (cut sub-cells)              */
    int32_t n_elements;        /* equation objects                                               */
    int32_t n_bases;           /* IP-cell bases per field (4 | 8 | 27)                           */
    int32_t n_adofs_per_type;  /* adofs of one dof type; total adofs = n_fields * this           */

    /* IP cell: axis-aligned box, [n_elements][2*dim] = lower corner then extent.                */
    const double* ip_box;
    /* material: diffusion [n_elements][1] = k; elasticity [n_elements][2] = E, nu.             */
    const double* material;

    /* T-matrix rows per (element, basis): type-local adof ids and weights (cl_MSI_Pdof_Host.cpp
       :309-337 via cl_HMR_T_Matrix_Base.cpp:306-457).  Row r = e*n_bases + b spans
       [t_ptr[r], t_ptr[r+1]).  The same rows serve every dof type of an element.               */
    const int32_t* t_ptr;      /* [n_elements*n_bases + 1]                                       */
    const int32_t* t_adof;
    const double*  t_weight;

    /* Cut sub-cells (IG cells of the element's material subphase).  ig_ptr[e] == ig_ptr[e+1]
       means uncut: the IP cell itself is integrated with the auto rule (QUAD_2x2 / HEX_2x2x2 /
       HEX_3x3x3 for HEX27); cut sub-cells take TRI_7 / TET_11, or TET_35 under HEX27.
       ig_xi holds every IG vertex's parametric coordinate wrt the IP cell,
       [n_ig][n_vertices(ig_geometry)][spatial_dim]
       (Cluster::get_primary_cell_local_coords_on_side_wrt_interp_cell).                      */
    const int32_t* ig_ptr;     /* [n_elements + 1]                                               */
    const double*  ig_xi;

    /* Lattice hint: elem_cell[e] = lattice cell (i + n0*(j + n1*k)).  HEX8 / QUAD4 sets on a
       lattice with nodal T rows (one entry of weight 1 per basis) take the owner-computes stencil
       path; every other set (no lattice, T rows of several entries such as p = 2 B-splines, HEX27
       IP cells) takes the generic element path (K_e per element, deterministic gather).        */
    mgpu_lattice   lattice;
    const int32_t* elem_cell;  /* [n_elements] or NULL                                          */

    /* Per-element level-set values at the IP-cell vertices, [n_elements][n_bases] (optional,
       needed by MGPU_QUAD_LEARNED).                                                             */
    const double*  vertex_phi;

    /* Element pairs whose adofs are coupled in the pattern beyond the elements themselves: the
       leader and follower of each facet of a double-sided side set (ghost facets), [n][2]
       (optional; the graph of Solver_Interface::build_graph covers every equation object, single-
       and double-sided, cl_Sparse_Matrix_EpetraFECrs.cpp:195-240).  Sets with couplings take the
       generic path.                                                                              */
    int32_t        n_couplings;
    const int32_t* couplings;
} mgpu_set_desc;

typedef struct mgpu_problem mgpu_problem;   /* host-side workload (cut mesh, numbering)        */
typedef struct mgpu_set     mgpu_set;       /* device-resident set + CSR pattern + scratch     */

/* Host set-up. */
mgpu_status mgpu_problem_build(const mgpu_problem_spec* spec, mgpu_problem** out);
mgpu_status mgpu_problem_desc(const mgpu_problem* p, mgpu_set_desc* out);
/* Counts: [0] elements, [1] cut elements, [2] IG sub-cells, [3] adofs per type, [4] lattice cells,
   [5] void cells.                                                                               */
mgpu_status mgpu_problem_counts(const mgpu_problem* p, int64_t counts[6]);
mgpu_status mgpu_problem_destroy(mgpu_problem* p);

/* Elemental extraction operators (T-matrices from the integration cells to the adofs) of a set
   descriptor, written to a file: the content of the reference's T-matrix output
   (Integration_Mesh::save_elemental_T_matrices_to_file, projects/MTK/src/mesh/
   cl_MTK_Integration_Mesh_write_T_matrices.cpp:38-160, 684-809; Workflow_HMR_XTK::
   output_T_matrices, projects/WRK/src/cl_WRK_Workflow_HMR_XTK.cpp:542-575), which stores HDF5.
   For every integration cell c (a cut element's sub-cells in order, or an uncut element's IP-cell
   box as one cell with its 2^dim corners) the weights E_c[v][a] = sum_b N_b(xi_v) T[b][a] map the
   element's adofs to the cell's vertices.  Host arrays only; synchronous.  Little-endian layout:
     char magic[8] = "MGPUEXT1"; int32 dim, n_bases, n_adofs_per_type, 0; int64 n_cells;
     per cell: int32 element, n_vertices, n_adofs; int32 adof[n_adofs] (ascending, type-local);
               double E[n_vertices][n_adofs].
   *n_cells (optional) receives the cell count.                                                 */
mgpu_status mgpu_write_extraction_operators(const mgpu_set_desc* desc, const char* path, int64_t* n_cells);

/* Set life cycle (fem::Set + MSI::Equation_Set scratch + SOL graph; cl_FEM_Set.cpp:35-100,
   :132 initialize_set, cl_DLA_Solver_Interface.cpp:20-54 build_graph).                          */
mgpu_status mgpu_set_create(const mgpu_set_desc* desc, mgpu_set** out);
mgpu_status mgpu_set_destroy(mgpu_set* set);
/* Precision and cut-cell quadrature of later assemble calls. */
mgpu_status mgpu_set_options(mgpu_set* set, int32_t precision, int32_t quadrature);
/* HIP stream (hipStream_t passed as void*) all work of the set is issued on; NULL = default. */
mgpu_status mgpu_set_stream(mgpu_set* set, void* hip_stream);

/* CSR pattern of the set's matrix (host arrays owned by the handle, valid until destroy):
   union of element topologies (cl_Sparse_Matrix_EpetraFECrs.cpp:195-240), rows and columns
   sorted ascending.                                                                             */
mgpu_status mgpu_pattern_get(const mgpu_set* set, int32_t* n_rows, int64_t* nnz,
                             const int32_t** row_ptr, const int32_t** col_idx);
/* Element topology (cl_MSI_Equation_Object.cpp:505-554): sorted unique adof ids per dof type,
   concatenated over types.  ids: [n_elements][n_fields*n_bases] (host, owned by the handle).  */
mgpu_status mgpu_element_topology(const mgpu_set* set, int32_t* n_per_element, const int32_t** ids);
/* The same for sets whose elements differ in size (e.g. truncated hierarchical rows), one call
   per set instead of get_element_topology per object (cl_MSI_Solver_Interface.hpp:369-385):
   element e's ids are ids[offsets[e] .. offsets[e+1]); offsets [n_elements + 1] (host arrays
   owned by the handle).                                                                         */
mgpu_status mgpu_element_topology_csr(const mgpu_set* set, const int64_t** offsets, const int32_t** ids);

/* K+R assembly of all elements of the set into caller-owned DEVICE arrays:
   csr_vals[nnz] and rhs[n_rows] are overwritten (zeroed then summed, cl_DLA_Linear_Problem.cpp
   :131-138 mat_put_scalar(0) + fill_matrix_and_RHS); u_adof[n_rows] is the adof vector
   (compute_my_pdof_values, cl_MSI_Equation_Object.cpp:1100-1131).  Asynchronous on the set's
   stream: a kernel-side failure (a stencil row whose length disagrees with the pattern, whose
   values are then left unwritten) is reported by the next synchronising call on the set
   (mgpu_set_assemble_host, mgpu_set_time_assemble, mgpu_set_autotune, mgpu_batch_synchronize),
   once; the flag is cleared when it is reported.  csr_vals must be ordinary device memory
   (hipMalloc): HEX27 sets in scatter mode add into it with FP64 atomics, which fine-grained or
   host-mapped memory does not take.                                                            */
mgpu_status mgpu_set_assemble(mgpu_set* set, const double* u_adof, double* csr_vals, double* rhs);
/* Same with HOST arrays; copies in and out, synchronous. */
mgpu_status mgpu_set_assemble_host(mgpu_set* set, const double* u_adof, double* csr_vals, double* rhs);

/* Statistics of the last assemble: [0] elements, [1] cut elements integrated by subdivision,
   [2] cut elements integrated by learned quadrature, [3] learned elements rejected by the moment
   check, [4] integration points evaluated.                                                      */
mgpu_status mgpu_set_stats(const mgpu_set* set, int64_t stats[5]);
/* The same statistics and more, the first n_stats of: [0]..[4] as mgpu_set_stats, [5] integration
   points skipped because their sub-cell has detJ < -1e-6 (Element_Bulk::compute_jacobian_and_
   residual, cl_FEM_Element_Bulk.cpp:291 -- a reported statistic, not an error), [6] such
   sub-cells.  Entries beyond [6] are set to 0.                                                */
mgpu_status mgpu_set_statistics(const mgpu_set* set, int64_t* stats, int32_t n_stats);

/* Device helpers (so callers need no HIP headers). */
mgpu_status mgpu_device_count(int32_t* n);
mgpu_status mgpu_malloc(void** ptr, size_t bytes);
mgpu_status mgpu_free(void* ptr);
mgpu_status mgpu_memcpy_htod(void* dst, const void* src, size_t bytes);
mgpu_status mgpu_memcpy_dtoh(void* dst, const void* src, size_t bytes);
mgpu_status mgpu_synchronize(void);
/* Time `iters` back-to-back mgpu_set_assemble calls with HIP events on the set's stream, after
   one untimed call.  per_call_ms receives the mean; kernel_ms[5] the mean per-call durations, from events around
   each stage of separate calls: [0] network (k_mlp_forward, inputs formed from the level set),
   [1] moment fit (k_lq_moments, k_lq_project), [2] cut-element integrands (k_lq_element and the subdivision
   kernel), [3] cell records (k_prep), [4] stencil scatter (k_stencil); 0 where a stage does not
   run.                                                                                          */
mgpu_status mgpu_set_time_assemble(mgpu_set* set, const double* u_adof, double* csr_vals,
                                   double* rhs, int32_t iters, double* per_call_ms,
                                   double kernel_ms[5]);


/* ---- launch shapes and the autotuner (no counterpart in the reference) -------------------------
   A set's launch shape is the stencil variant and the order of the cut-element workgroups
   (0 natural, 1 longest first by sub-cell count).  Stencil variants differ in rows per workgroup
   and in how a row reaches HBM: the first half of the variants stage the workgroup's CSR range
   in LDS and write it coalesced, the second half (same row counts) store each row directly.
   Generic-path sets have no stencil: their shape is the order of the element workgroups of the
   gather mode (0 element order, 1 most sub-cells first; scatter mode orders each colour most
   sub-cells first itself), and the autotuner times only (0, 0) and (0, 1).
   Results are bit-identical for every shape.                                                   */
mgpu_status mgpu_set_tuning(mgpu_set* set, int32_t stencil_variant, int32_t cut_order);
mgpu_status mgpu_set_get_tuning(const mgpu_set* set, int32_t* stencil_variant, int32_t* cut_order);
/* Number of stencil variants of the set and the rows per workgroup of each (rows[] optional,
   at least 6 entries).                                                                         */
mgpu_status mgpu_set_tuning_variants(const mgpu_set* set, int32_t* n_variants, int32_t* rows);
/* Time every launch shape with `iters` assemble calls on the given device buffers and keep the
   fastest.  The choice is cached per process by set shape (dimension, fields, rows, cut count
   rounded to a power of two, precision, quadrature, path); flags bit 0 ignores the cache.
   ms[2 * n_variants] (optional) receives the per-call mean of shape (variant, order) at
   [2 * variant + order], or -1 where the cache answered.                                       */
mgpu_status mgpu_set_autotune(mgpu_set* set, const double* u_adof, double* csr_vals, double* rhs,
                              int32_t iters, int32_t flags, double* ms);

/* ---- batches of sets captured into one HIP graph (no counterpart in the reference) ------------
   Replaces the loop over sets of Solver_Interface::fill_matrix_and_RHS
   (projects/SOL/DLA/src/cl_DLA_Solver_Interface.cpp:317-365) for sets that are assembled
   together, e.g. design variants: one assemble call of each set, on its current options and
   launch shape, captured once with fixed device buffers u[i], vals[i], rhs[i] (which must stay
   valid) and replayed as one graph launch.  The sets are spread over n_branches parallel graph
   branches (0: autotuned among 1, 2, 4, 8).  The batch does not own the sets; a set must not be
   reconfigured (options, tuning, model) while a batch holds it.                                */
typedef struct mgpu_batch mgpu_batch;
mgpu_status mgpu_batch_create(mgpu_set* const* sets, int32_t n_sets, const double* const* u_adof,
                              double* const* csr_vals, double* const* rhs, int32_t n_branches,
                              void* hip_stream, mgpu_batch** out);
mgpu_status mgpu_batch_destroy(mgpu_batch* batch);
/* Enqueue one replay of the graph on the batch's stream (asynchronous). */
mgpu_status mgpu_batch_launch(mgpu_batch* batch);
/* Wait for the batch's stream and check every set's row-length flag. */
mgpu_status mgpu_batch_synchronize(mgpu_batch* batch);
/* Mean duration of `iters` back-to-back replays (HIP events), and the branch count in use. */
mgpu_status mgpu_batch_time(mgpu_batch* batch, int32_t iters, double* per_launch_ms, int32_t* n_branches);

const char* mgpu_last_error_string(void);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_H */
