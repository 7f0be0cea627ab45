// set_device.hpp -- device-side view of an mgpu_set and the element-level building blocks shared
// by the assembly kernels (subdivision, learned quadrature, stencil scatter).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common/fem_tables.hpp"

namespace mgpu {

// ------------------------------------------------------------------------------------------------
// device-side view of a set
// ------------------------------------------------------------------------------------------------
struct DevView {
    int32_t n_el, n_cut, N, n_rows;
    int32_t nx, ny, nz;
    double h[3];
    const double* material;     // [n_el][nm]
    const int32_t* cell_elem;   // [n_cells]
    const int32_t* elem_cut;    // [n_el]
    const int32_t* node_adof;   // [n_nodes]
    const int32_t* row_node;    // [N]
    const int32_t* row_ptr;     // [n_rows + 1]
    const int32_t* cut_elems;   // [n_cut]
    const int32_t* ig_ptr;      // [n_el + 1]
    const double* ig_xi;        // [n_ig][nv][dim]
    double* ke_cut;             // [n_cut][na][na]: rows (vertex, field) node-major, columns (field, vertex)
    float* ke_cut_f;            // the same in FP32 (mixed precision: FP32 integrands)
    int32_t* err_flag;          // [1]
    double* cell_rec;           // per lattice cell: diffusion [1] {k | -(cut+1) | 0 void},
                                // elasticity [4] {flag, D_nn, D_nn', D_shear}
    double* u_node;             // per lattice node: u of its adofs, planar [nf][n_nodes] (0 if unused)
    const int32_t* skip;        // [n_cut] nonzero: element integrated elsewhere (learned), or NULL
    const int32_t* cut_order;   // [n_cut] cut element processed by block b (longest first), or NULL
};

// ------------------------------------------------------------------------------------------------
// Integrals over the reference cell [-1,1]^DIM of dN_a/dxi_i * dN_b/dxi_j for the trilinear /
// bilinear Lagrange basis N_a = prod_d (1 + xi_d s_ad) / 2^DIM (s_ad = +-1, the vertex signs of
// the HEX8 / QUAD4 vertex order).  The integrand factorises over the axes; per axis d the factor
// is 2 s_ad s_bd (d = i = j), 2 s_ad (d = i != j), 2 s_bd (d = j != i) or 2 + 2/3 s_ad s_bd
// (otherwise), times 4^-DIM overall.  HEX_2x2x2 / QUAD_2x2 integrate these products (degree <= 2
// per variable) exactly, so on an axis-aligned cell of size h the moments of an uncut element are
//   M[(i,a),(j,b)] = G[i][a][j][b] * 4 / (h_i h_j) * prod_d h_d / 2.
// ------------------------------------------------------------------------------------------------
__host__ __device__ constexpr int lagrange_vertex_sign(int dim, int a, int d) {
    // HEX8: 0 (-,-,-) 1 (+,-,-) 2 (+,+,-) 3 (-,+,-) 4..7 the same at z = +1; QUAD4: 0..3 as z = -1
    return d == 0 ? ((a % 4 == 1 || a % 4 == 2) ? 1 : -1)
                  : d == 1 ? ((a % 4 == 2 || a % 4 == 3) ? 1 : -1) : (dim == 3 && a >= 4 ? 1 : -1);
}

template <int DIM>
struct RefGradIntegrals {
    static constexpr int NB = DIM == 3 ? 8 : 4;
    double v[DIM][NB][DIM][NB];
    constexpr RefGradIntegrals() : v{} {
        for (int i = 0; i < DIM; ++i)
            for (int a = 0; a < NB; ++a)
                for (int j = 0; j < DIM; ++j)
                    for (int b = 0; b < NB; ++b) {
                        double p = 1.0;
                        for (int d = 0; d < DIM; ++d) {
                            const double sa = lagrange_vertex_sign(DIM, a, d), sb = lagrange_vertex_sign(DIM, b, d);
                            if (d == i && d == j) p *= 2.0 * sa * sb;
                            else if (d == i) p *= 2.0 * sa;
                            else if (d == j) p *= 2.0 * sb;
                            else p *= 2.0 + (2.0 / 3.0) * sa * sb;
                        }
                        v[i][a][j][b] = p / (DIM == 3 ? 64.0 : 16.0);
                    }
    }
};
template <int DIM>
constexpr RefGradIntegrals<DIM> kRefGrad{};

// ------------------------------------------------------------------------------------------------
// device-side view of a set on the generic element path (any T rows, no lattice needed)
// ------------------------------------------------------------------------------------------------
struct GenView {
    int32_t n_el, N, n_rows;
    int64_t nnz;
    const double* ip_box;        // [n_el][2 dim] lower corner, extent
    const double* material;      // [n_el][nm]
    const int32_t* ig_ptr;       // [n_el + 1]
    const double* ig_xi;         // [n_ig][nv][dim]
    const int32_t* t_ptr;        // [n_el nb + 1]
    const int32_t* t_col;        // element-local column of each T entry
    const double* t_w;           // weight of each T entry
    const int32_t* na_t;         // [n_el] adofs per dof type
    const int64_t* ke_off;       // [n_el + 1]
    double* ke;                  // element matrices, (nf na_t)^2 per element, row-major
    const int64_t* gather_ptr;   // [nnz + 1]
    const int64_t* gather_src;   // K_e entries summed into each CSR value
    const int32_t* gather_src32; // the same in 32 bits when the pool has fewer than 2^31 entries
    const int32_t* row_ptr;      // [n_rows + 1]
    const int32_t* col_idx;      // [nnz]
    // scatter mode (HEX27 sets without a K_e pool)
    const int32_t* topo;         // rows of each element, nf * na_t from topo_ptr[e]
    const int64_t* topo_ptr;     // [n_el + 1]
    const int32_t* colour_el;    // [n_el] elements by colour
    // gather mode: the processing order of the element workgroups (most sub-cells first), or
    // null for element order; the autotuner picks one
    const int32_t* order;
    const int64_t* rank_off;     // [n_el + 1]
    const uint16_t* rank;        // [na_t][na_t] per element: column of adof y among adof x's neighbours
};

// Storage of the cut-element matrices: FP64 on the plain path, FP32 with the FP32 integrands.
template <typename KT>
__device__ inline KT* ke_store(const DevView& s) {
    if constexpr (sizeof(KT) == 4) return s.ke_cut_f;
    else return s.ke_cut;
}

// Voigt index of the symmetric pair (i, j): 3D [xx,yy,zz,yz,xz,xy]
// (CM_Struc_Linear::eval_strain_3d, cl_FEM_CM_Struc_Linear.cpp:448-453); 2D [xx,yy,xy].
template <int DIM>
__host__ __device__ constexpr int voigt(int i, int j) {
    return DIM == 3 ? (i == j ? i : (i + j == 3 ? 3 : (i + j == 2 ? 4 : 5))) : (i == j ? i : 2);
}

// Structural non-zeros of the isotropic D: the normal block and the shear diagonal.
template <int DIM>
__host__ __device__ constexpr bool d_nonzero(int a, int b) {
    return DIM == 3 ? ((a < 3 && b < 3) || a == b) : ((a < 2 && b < 2) || a == b);
}

// Constitutive matrix D (Voigt, engineering shear):
//   3D  CM_Struc_Linear_Isotropic::full_3d, cl_FEM_CM_Struc_Linear_Isotropic.cpp:374-390
//   2D  CM_Struc_Linear_Isotropic::full_plane_stress, :248-259
template <int DIM>
__device__ inline void elastic_D(double E, double nu, double (*D)[DIM == 3 ? 6 : 3]) {
    constexpr int NS = DIM == 3 ? 6 : 3;
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) D[i][j] = 0.0;
    if constexpr (DIM == 3) {
        const double pre = E / (1.0 + nu) / (1.0 - 2.0 * nu);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i][j] = i == j ? pre * (1.0 - nu) : pre * nu;
        for (int i = 3; i < 6; ++i) D[i][i] = pre * (1.0 - 2.0 * nu) / 2.0;
    } else {
        const double pre = E / (1.0 - nu * nu);
        D[0][0] = pre;
        D[1][1] = pre;
        D[0][1] = pre * nu;
        D[1][0] = pre * nu;
        D[2][2] = pre * 0.5 * (1.0 - nu);
    }
}

// dN/dxi of the linear Lagrange IP cell at xi: out[i][a].
template <int DIM>
__device__ inline void lagrange_dNdxi(const double* xi, double (*out)[DIM == 3 ? 8 : 4]) {
    if constexpr (DIM == 3) {
        for (int a = 0; a < 8; ++a) {
            const double sx = kHex8XiHat[a][0], sy = kHex8XiHat[a][1], sz = kHex8XiHat[a][2];
            const double fx = 1.0 + xi[0] * sx, fy = 1.0 + xi[1] * sy, fz = 1.0 + xi[2] * sz;
            out[0][a] = 0.125 * sx * fy * fz;
            out[1][a] = 0.125 * fx * sy * fz;
            out[2][a] = 0.125 * fx * fy * sz;
        }
    } else {
        for (int a = 0; a < 4; ++a) {
            const double sx = kQuad4XiHat[a][0], sy = kQuad4XiHat[a][1];
            out[0][a] = 0.25 * sx * (1.0 + xi[1] * sy);
            out[1][a] = 0.25 * (1.0 + xi[0] * sx) * sy;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Integration point ip of the cut element's IG sub-cells: its parametric position in the IP cell
// (Space_Interpolator::map_integration_point, cl_MTK_Space_Interpolator.cpp:1381-1397) and its
// weight w* = w detJ (0 where detJ < -1e-6, cl_FEM_Element_Bulk.cpp:291), all in FP64.
// ------------------------------------------------------------------------------------------------
template <int DIM>
__device__ inline double cut_point_map(const DevView& s, int ig0, int ip, const double* hh, double* xip) {
    constexpr int NV = DIM + 1;
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    const int sc = ig0 + ip / NQ;
    const int q = ip % NQ;
    const double* X = s.ig_xi + (size_t)sc * NV * DIM;
    double N[NV];
    if constexpr (DIM == 3) {
        N[0] = kTet11Xi[q][0];
        N[1] = kTet11Xi[q][1];
        N[2] = 1.0 - kTet11Xi[q][0] - kTet11Xi[q][1] - kTet11Xi[q][2];
        N[3] = kTet11Xi[q][2];
    } else {
        N[0] = kTri7Xi[q][0];
        N[1] = kTri7Xi[q][1];
        N[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
    }
    for (int d = 0; d < DIM; ++d) {
        double v = 0.0;
        for (int k = 0; k < NV; ++k) v += N[k] * X[k * DIM + d];
        xip[d] = v;
    }
    double detJ;
    if constexpr (DIM == 3) {
        double J[3][3];
        const int rr[3] = {0, 1, 3};
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) J[r][d] = (X[rr[r] * 3 + d] - X[2 * 3 + d]) * (0.5 * hh[d]);
        detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
    } else {
        double J[2][2];
        for (int r = 0; r < 2; ++r)
            for (int d = 0; d < 2; ++d) J[r][d] = (X[r * 2 + d] - X[2 * 2 + d]) * (0.5 * hh[d]);
        detJ = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / 2.0;
    }
    return detJ < kDetJLowerLimit ? 0.0 : (DIM == 3 ? kTet11W[q] : kTri7W[q]) * detJ;
}

// ------------------------------------------------------------------------------------------------
// Integration point ip of the cut element's IG sub-cells: weight w* and IP-cell gradients g[i][a].
// ------------------------------------------------------------------------------------------------
template <int DIM>
__device__ inline double cut_point(const DevView& s, int ig0, int ip, const double* hh,
                                   double (*g)[DIM == 3 ? 8 : 4]) {
    constexpr int NV = DIM + 1;
    constexpr int NQ = DIM == 3 ? kTet11Points : kTri7Points;
    const int sc = ig0 + ip / NQ;
    const int q = ip % NQ;
    const double* X = s.ig_xi + (size_t)sc * NV * DIM;
    double N[NV];
    if constexpr (DIM == 3) {
        N[0] = kTet11Xi[q][0];
        N[1] = kTet11Xi[q][1];
        N[2] = 1.0 - kTet11Xi[q][0] - kTet11Xi[q][1] - kTet11Xi[q][2];
        N[3] = kTet11Xi[q][2];
    } else {
        N[0] = kTri7Xi[q][0];
        N[1] = kTri7Xi[q][1];
        N[2] = 1.0 - kTri7Xi[q][0] - kTri7Xi[q][1];
    }
    double xip[DIM];
    for (int d = 0; d < DIM; ++d) {
        double v = 0.0;
        for (int k = 0; k < NV; ++k) v += N[k] * X[k * DIM + d];
        xip[d] = v;
    }
    // detJ of the simplex in physical space: rows X_r - X_2 (r = 0, 1 [, 3]) scaled by h/2
    double detJ;
    if constexpr (DIM == 3) {
        double J[3][3];
        const int rr[3] = {0, 1, 3};
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) J[r][d] = (X[rr[r] * 3 + d] - X[2 * 3 + d]) * (0.5 * hh[d]);
        detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
    } else {
        double J[2][2];
        for (int r = 0; r < 2; ++r)
            for (int d = 0; d < 2; ++d) J[r][d] = (X[r * 2 + d] - X[2 * 2 + d]) * (0.5 * hh[d]);
        detJ = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / 2.0;
    }
    const double w = detJ < kDetJLowerLimit ? 0.0 : (DIM == 3 ? kTet11W[q] : kTri7W[q]) * detJ;
    lagrange_dNdxi<DIM>(xip, g);
    for (int i = 0; i < DIM; ++i)
        for (int a = 0; a < (DIM == 3 ? 8 : 4); ++a) g[i][a] *= 2.0 / hh[i];
    return w;
}

}  // namespace mgpu
