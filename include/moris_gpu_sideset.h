/*
 * moris_gpu_sideset.h -- side-set IWGs of libmoris_gpu.so: weak Dirichlet (Nitsche) and Neumann
 * conditions on the sides of the bulk set's IG cells (domain boundary faces and the level-set
 * interface), assembled on the GPU into the bulk set's CSR values and residual.
 *
 * Reference path: Element_Sideset::compute_jacobian_and_residual
 * (projects/FEM/INT/src/ELEM/cl_FEM_Element_Sideset.cpp:245-330: per side point w* = w detJ_side,
 * skipped below detJ -1e-6, the side's outward normal from the mesh) calling
 *   IWG_Diffusion_Dirichlet_Nitsche        projects/FEM/INT/src/IWG/cl_FEM_IWG_Diffusion_Dirichlet_Nitsche.cpp
 *   IWG_Isotropic_Struc_Linear_Dirichlet   projects/FEM/INT/src/IWG/cl_FEM_IWG_Isotropic_Struc_Linear_Dirichlet.cpp:52-240
 *       R += w* ( -N^T t(u) + beta T^T (u - u_D) + gamma N^T (u - u_D) ),
 *       K += w* ( -N^T T + beta T^T N + gamma N^T N ),  T = n_flat D B (diffusion: k n^T B),
 *       beta = -1 for the SYMMETRIC and +1 for the UNSYMMETRIC variant (cl_FEM_IWG_Factory.cpp:156-159,196-199)
 *   IWG_Diffusion_Neumann / IWG_Isotropic_Struc_Linear_Neumann
 *                                          projects/FEM/INT/src/IWG/cl_FEM_IWG_Isotropic_Struc_Linear_Neumann.cpp:38-85
 *       R -= w* N^T t_N
 * with the Nitsche parameter gamma = alpha * material / h of SP_Dirichlet_Nitsche
 * (projects/FEM/INT/src/SP/cl_FEM_SP_Dirichlet_Nitsche.cpp:75-115; material = k or E; h the cell
 * length measure of the side cluster, sqrt(4 A / pi) in 3D and the side length in 2D, A the total
 * side measure of the element's facets in the set, cl_FEM_Cluster.cpp:858-919) and the reference's
 * auto side rules TRI_7 (tet faces), QUAD_2x2 (hex faces), BAR_3 (2D sides) (cl_FEM_Set.cpp
 * :2983-3160).  Same conventions as moris_gpu.h.
 */
#ifndef MORIS_GPU_SIDESET_H
#define MORIS_GPU_SIDESET_H

#include "moris_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Side IWGs (cl_FEM_IWG_Factory.cpp:156-168, 196-220). */
#define MGPU_IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE    2
#define MGPU_IWG_SPATIALDIFF_DIRICHLET_UNSYMMETRIC_NITSCHE  3
#define MGPU_IWG_SPATIALDIFF_NEUMANN                        4
#define MGPU_IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE   5
#define MGPU_IWG_STRUC_LINEAR_DIRICHLET_UNSYMMETRIC_NITSCHE 6
#define MGPU_IWG_STRUC_LINEAR_NEUMANN                       7
/* Double-sided ghost stabilisation on the faces between a cut element and its neighbours
   (SPATIALDIFF_VW_GHOST and STRUC_LINEAR_VW_GHOST, cl_FEM_IWG_Factory.cpp:183, 223;
   IWG_Diffusion_Virtual_Work_Ghost, projects/FEM/INT/src/IWG/cl_FEM_IWG_Diffusion_Virtual_Work_
   Ghost.cpp; IWG_Isotropic_Struc_Linear_Virtual_Work_Ghost, cl_FEM_IWG_Isotropic_Struc_Linear_
   Virtual_Work_Ghost.cpp:50-255), order 1, with SP_Ghost_Virtual_Work gamma = alpha h
   (projects/FEM/INT/src/SP/cl_FEM_SP_Ghost_Virtual_Work.cpp:30-41; h the cell length measure of
   the double-sided cluster, sqrt(4 A / pi) in 3D, A in 2D, A the face measure):
     R_L += w* gamma T_L^T [[t]],  R_F -= w* gamma T_F^T [[t]],  [[t]] = N_n (D_L B_L u_L - D_F B_F u_F),
   T_X = N_n B_X the test traction of side X without its material (diffusion: n^T grad N_X,
   [[t]] = n^T (k_L grad u_L - k_F grad u_F)), N_n the Voigt traction operator of the leader's
   outward normal n.                                                                             */
#define MGPU_IWG_SPATIALDIFF_GHOST                          8
#define MGPU_IWG_STRUC_LINEAR_GHOST                         9

/* Side geometry: the facet kind of one side set (mtk side geometry of the IG cells). */
#define MGPU_GEOM_LINE2  5   /* 2D: a side of a QUAD4 or TRI3 IG cell, BAR_3                          */
/* 3D: MGPU_GEOM_TRI3 (a face of a TET4 sub-cell, TRI_7) or MGPU_GEOM_QUAD4 (a face of an uncut
   HEX8 cell, QUAD_2x2).                                                                          */

/* Where facets are taken from (mgpu_problem_facets). */
#define MGPU_SIDE_INTERFACE  0   /* level-set interface: faces of material sub-cells that lie on no
                                    face of their cell and are shared with no other material sub-cell */
#define MGPU_SIDE_XMIN       1   /* domain boundary faces x = lo[0], x = hi[0], y = ..., z = ...      */
#define MGPU_SIDE_XMAX       2
#define MGPU_SIDE_YMIN       3
#define MGPU_SIDE_YMAX       4
#define MGPU_SIDE_ZMIN       5
#define MGPU_SIDE_ZMAX       6
#define MGPU_SIDE_GHOST     -1   /* ghost facets: the faces between two material elements of a level-0
                                    lattice of which at least one is cut and whose shared face holds a
                                    lattice node strictly in the material; leader the lower cell index,
                                    the facet the leader's face (QUAD4 / LINE2), n its outward normal */

/* A side set: facets of one geometry, each on an element of the bulk set.  Facet vertices are given
   in the parametric coordinates of the element's IP cell, [n_facets][n_vertices][spatial_dim]
   (LINE2 2, TRI3 3, QUAD4 4 vertices in cyclic order); normals are the outward unit normals of
   the material domain, [n_facets][spatial_dim].                                                 */
typedef struct mgpu_sideset_desc {
    int32_t iwg;               /* MGPU_IWG_* side IWG of the bulk set's field (diffusion / elasticity) */
    int32_t side_geometry;     /* MGPU_GEOM_TRI3 | MGPU_GEOM_QUAD4 (3D), MGPU_GEOM_LINE2 (2D)       */
    int32_t n_facets;
    const int32_t* facet_elem; /* [n_facets] element of the bulk set                               */
    const double*  facet_xi;   /* [n_facets][n_vertices][dim]                                       */
    const double*  facet_normal; /* [n_facets][dim]                                                  */
    double nitsche_alpha;      /* Dirichlet: SP parameter alpha of gamma = alpha * material / h;
                                  ghost: alpha of gamma = alpha h                                   */
    double value[3];           /* Dirichlet: prescribed u_D per field; Neumann: flux / traction t_N  */
    const int32_t* facet_follower; /* ghost IWGs: [n_facets] the follower element.  The facet is the
                                  whole face the two IP cells share: equal boxes, the follower's
                                  the leader's shifted along the facet normal (an axis)             */
} mgpu_sideset_desc;

typedef struct mgpu_sideset mgpu_sideset;

/* A side set on the bulk set `bulk` (which must outlive it): lattice sets with nodal T rows, and
   generic-path sets (p = 2 B-spline rows on HEX8 / QUAD4 cells, HEX27 cells, THB rows) with up to
   64 adofs per dof type and element.  Ghost side sets need HEX8 / QUAD4 cells, at most 64 adofs
   per dof type on the two elements of a facet together, and a bulk set whose descriptor couples
   every facet's two elements (mgpu_set_desc.couplings).                                         */
mgpu_status mgpu_sideset_create(mgpu_set* bulk, const mgpu_sideset_desc* desc, mgpu_sideset** out);
mgpu_status mgpu_sideset_destroy(mgpu_sideset* ss);
/* ADD the side set's K and R into the bulk set's CSR values and R (DEVICE arrays; call after
   mgpu_set_assemble, which overwrites them).  Asynchronous on the bulk set's stream; each value
   receives its side contributions in ascending facet order (deterministic).                   */
mgpu_status mgpu_sideset_assemble(mgpu_sideset* ss, const double* u_adof, double* csr_vals, double* rhs);

/* Facets of a problem (host set-up, as the reference's side sets of the XTK integration mesh):
   `where` one of MGPU_SIDE_*, `side_geometry` the facet kind (cut cells give TRI3 / LINE2, uncut
   cells QUAD4 / LINE2).  Elements ascending; within a cut element its material sub-cells in order,
   a sub-cell's sides in the order of the vertex they omit (0, 1, 2[, 3]), the remaining vertices in
   ascending local order.  The arrays are owned by the returned handle.                          */
typedef struct mgpu_facets mgpu_facets;
mgpu_status mgpu_problem_facets(const mgpu_problem* p, int32_t where, int32_t side_geometry, mgpu_facets** out);
mgpu_status mgpu_facets_get(const mgpu_facets* f, int32_t* n_facets, const int32_t** facet_elem,
                            const double** facet_xi, const double** facet_normal);
mgpu_status mgpu_facets_destroy(mgpu_facets* f);
/* The follower element of every facet of a MGPU_SIDE_GHOST facet list (NULL for other lists). */
mgpu_status mgpu_facets_followers(const mgpu_facets* f, const int32_t** facet_follower);

#ifdef __cplusplus
}
#endif
#endif /* MORIS_GPU_SIDESET_H */
