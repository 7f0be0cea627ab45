// quad.hpp -- the learned cut-cell quadrature's fixed point set and moment space.
//
// Points: the HEX_3x3x3 Gauss points of the IP cell (cl_MTK_Integration_Coeffs_Hex_3x3x3.hpp).
// Moment space: monomials xi^a eta^b zeta^c with a, b, c <= 2 and a + b + c <= 4.  Every product
// dN_a/dx_i * dN_b/dx_j of trilinear (hex8) shape functions lies in this 23-dimensional space, so a
// rule on the 27 points that reproduces these moments of the material region reproduces the
// element matrices B^T D B of the subdivision quadrature (TET_11 is exact to degree 4 on the affine
// sub-tets) up to rounding.
#pragma once
#include <cmath>
#include <cstdint>

#include "../common/fem_tables.hpp"

namespace mgpu {

constexpr int kLqPoints = 27;       // HEX_3x3x3
constexpr int kLqMoments = 23;      // monomials of the moment space
constexpr int kLqMomentsPad = 24;   // stored per element (zero-padded for the 16 x 16 x 4 MFMA tiles)
constexpr int kLqInputs = 8;        // nodal level-set values of the hex8 IP cell
constexpr int kLqHidden = 64;
// MLP parameter layout (FP32): W1[8][64], b1[64], W2[64][64], b2[64], W3[64][27], b3[27]
constexpr int kLqOffW1 = 0;
constexpr int kLqOffB1 = kLqOffW1 + kLqInputs * kLqHidden;
constexpr int kLqOffW2 = kLqOffB1 + kLqHidden;
constexpr int kLqOffB2 = kLqOffW2 + kLqHidden * kLqHidden;
constexpr int kLqOffW3 = kLqOffB2 + kLqHidden;
constexpr int kLqOffB3 = kLqOffW3 + kLqHidden * kLqPoints;
constexpr int kLqParams = kLqOffB3 + kLqPoints;   // 6491

// Exponents (a, b, c) of moment k; a fastest.
MGPU_HD inline void lq_exponents(int k, int* a, int* b, int* c) {
    int n = 0;
    for (int cc = 0; cc <= 2; ++cc)
        for (int bb = 0; bb <= 2; ++bb)
            for (int aa = 0; aa <= 2; ++aa) {
                if (aa + bb + cc > 4) continue;
                if (n == k) { *a = aa; *b = bb; *c = cc; return; }
                ++n;
            }
    *a = *b = *c = 0;
}

// All 23 monomials at xi.
MGPU_HD inline void lq_monomials(const double* xi, double* p) {
    const double x[3] = {1.0, xi[0], xi[0] * xi[0]};
    const double y[3] = {1.0, xi[1], xi[1] * xi[1]};
    const double z[3] = {1.0, xi[2], xi[2] * xi[2]};
    int n = 0;
    for (int cc = 0; cc <= 2; ++cc)
        for (int bb = 0; bb <= 2; ++bb)
            for (int aa = 0; aa <= 2; ++aa) {
                if (aa + bb + cc > 4) continue;
                p[n++] = x[aa] * y[bb] * z[cc];
            }
}

// All 27 tri-quadratic monomials (a, b, c <= 2, a fastest) at xi; the first-listed 23 of
// lq_monomials are the subset with a + b + c <= 4.
MGPU_HD inline void lq_monomials27(const double* xi, double* p) {
    const double x[3] = {1.0, xi[0], xi[0] * xi[0]};
    const double y[3] = {1.0, xi[1], xi[1] * xi[1]};
    const double z[3] = {1.0, xi[2], xi[2] * xi[2]};
    int n = 0;
    for (int cc = 0; cc <= 2; ++cc)
        for (int bb = 0; bb <= 2; ++bb)
            for (int aa = 0; aa <= 2; ++aa) p[n++] = x[aa] * y[bb] * z[cc];
}

// Grundmann-Moeller rule of degree 7 on the unit tetrahedron (35 points, barycentric points and
// weights summing to 1/6), used for the degree-6 moments of the training targets.
constexpr int kGm7Points = 35;

// Fit operators built on the host: V[k][q] = p_k(xi_q) and the minimum-norm correction
// P = V^T (V V^T)^{-1} (27 x 23); V27inv: inverse of the 27 x 27 tri-quadratic moment matrix.
struct LqOperators {
    double V[kLqMoments][kLqPoints];
    double P[kLqPoints][kLqMoments];
    double V27inv[kLqPoints][kLqPoints];
    double gauss_w[kLqPoints];      // HEX_3x3x3 weights (sum 8)
    double gm_lambda[kGm7Points][4];
    double gm_w[kGm7Points];
};
void lq_build_operators(LqOperators* op);

// Minimum-norm fit of weights to moments m starting from w0 scaled to the volume moment m[0]
// (iters = 0), optionally alternated `iters` times with the nonnegative orthant; returns the
// relative residual ||V w - m|| / ||m|| (FP64 weights; the device fit also rounds them to FP32).
double lq_fit_nonnegative(const LqOperators& op, const double* m, const double* w0, int iters, double* w);

// Physical gradients dN_a/dx_i of the hex8 basis at the 27 points of a cell of size h:
// g27[q][i * 8 + a] (FP32).
void lq_point_gradients(const double* h, float* g27);

}  // namespace mgpu
