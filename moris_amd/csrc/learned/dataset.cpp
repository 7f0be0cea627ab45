// dataset.cpp -- training data of the learned cut-cell quadrature and the host side of its fit.
//
// A sample is one hex8 cell (parametric box [-1,1]^3) cut by a seeded random level set: a plane
// through the cell, a sphere (radius in [0.1, 1.5] h, centre within 2h, either orientation) or a
// gyroid patch (random phase, period about four cells).  The cell is decomposed by the same XTK
// restatement as the assembly set-up (common/xtk_cut.hpp, regular subdivision + node hierarchy),
// all 27 tri-quadratic moments of the material sub-tets are integrated with a degree-7
// Grundmann-Moeller rule (exact to degree 6), and the target weights on the HEX_3x3x3 points are
// the unique rule that reproduces them (the tri-quadratic moment matrix at the 27 points is
// invertible).  Inputs are the 8 corner level-set values divided by their largest magnitude.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../common/levelset.hpp"
#include "../common/xtk_cut.hpp"
#include "../setup/internal.hpp"
#include "quad.hpp"

namespace mgpu {

namespace {
// Gauss-Jordan inverse of an n x n matrix (row-major, n <= 27); returns false if singular.
bool invert(int n, const double* A, double* Ainv) {
    double M[27][54];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            M[i][j] = A[i * n + j];
            M[i][n + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) std::swap(M[c][j], M[piv][j]);
        const double d = M[c][c];
        for (int j = 0; j < 2 * n; ++j) M[c][j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = M[i][n + j];
    return true;
}

// Grundmann & Moeller (SIAM J. Numer. Anal. 15, 1978): degree 2s+1 on the n-simplex,
//   sum_{i=0..s} (-1)^i 2^{-2s} (d+n-2i)^d / (i! (d+n-i)!) sum_{|beta|=s-i} f((2 beta_j + 1)/(d+n-2i))
// here n = 3, s = 3, d = 7.
void gm7(double (*lam)[4], double* w) {
    const int n = 3, s = 3, d = 7;
    int k = 0;
    auto fact = [](int m) { double f = 1.0; for (int i = 2; i <= m; ++i) f *= i; return f; };
    for (int i = 0; i <= s; ++i) {
        const int tot = s - i;
        const double den = (double)(d + n - 2 * i);
        const double wi = (i % 2 ? -1.0 : 1.0) * std::pow(2.0, -2.0 * s) * std::pow(den, d) / (fact(i) * fact(d + n - i));
        for (int b0 = 0; b0 <= tot; ++b0)
            for (int b1 = 0; b0 + b1 <= tot; ++b1)
                for (int b2 = 0; b0 + b1 + b2 <= tot; ++b2) {
                    const int b3 = tot - b0 - b1 - b2;
                    lam[k][0] = (2 * b0 + 1) / den;
                    lam[k][1] = (2 * b1 + 1) / den;
                    lam[k][2] = (2 * b2 + 1) / den;
                    lam[k][3] = (2 * b3 + 1) / den;
                    w[k] = wi;
                    ++k;
                }
    }
}
}  // namespace

void lq_build_operators(LqOperators* op) {
    gm7(op->gm_lambda, op->gm_w);
    {
        double A[kLqPoints * kLqPoints];
        for (int q = 0; q < kLqPoints; ++q) {
            double p[kLqPoints];
            lq_monomials27(kHex3Xi[q], p);
            for (int k = 0; k < kLqPoints; ++k) A[k * kLqPoints + q] = p[k];
        }
        invert(kLqPoints, A, &op->V27inv[0][0]);
    }
    for (int q = 0; q < kLqPoints; ++q) {
        double p[kLqMoments];
        lq_monomials(kHex3Xi[q], p);
        for (int k = 0; k < kLqMoments; ++k) op->V[k][q] = p[k];
        op->gauss_w[q] = kHex3W[q];
    }
    // G = V V^T, inverse by Gauss-Jordan with partial pivoting
    double G[kLqMoments][2 * kLqMoments];
    for (int i = 0; i < kLqMoments; ++i)
        for (int j = 0; j < kLqMoments; ++j) {
            double v = 0.0;
            for (int q = 0; q < kLqPoints; ++q) v += op->V[i][q] * op->V[j][q];
            G[i][j] = v;
            G[i][kLqMoments + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < kLqMoments; ++c) {
        int piv = c;
        for (int r = c + 1; r < kLqMoments; ++r)
            if (std::fabs(G[r][c]) > std::fabs(G[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 2 * kLqMoments; ++j) std::swap(G[c][j], G[piv][j]);
        const double d = G[c][c];
        for (int j = 0; j < 2 * kLqMoments; ++j) G[c][j] /= d;
        for (int r = 0; r < kLqMoments; ++r) {
            if (r == c) continue;
            const double f = G[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * kLqMoments; ++j) G[r][j] -= f * G[c][j];
        }
    }
    for (int q = 0; q < kLqPoints; ++q)
        for (int k = 0; k < kLqMoments; ++k) {
            double v = 0.0;
            for (int j = 0; j < kLqMoments; ++j) v += op->V[j][q] * G[j][kLqMoments + k];
            op->P[q][k] = v;
        }
}

double lq_fit_nonnegative(const LqOperators& op, const double* m, const double* w0, int iters, double* w) {
    double r[kLqMoments];
    // prior scaled to the volume moment m[0] (moment 0 is the monomial 1)
    double sw = 0.0;
    for (int q = 0; q < kLqPoints; ++q) sw += w0[q];
    const double scale = sw > 0.0 && m[0] > 0.0 ? m[0] / sw : 1.0;
    for (int q = 0; q < kLqPoints; ++q) w[q] = w0[q] * scale;
    for (int it = 0; it <= iters; ++it) {
        if (it > 0)
            for (int q = 0; q < kLqPoints; ++q) w[q] = std::max(w[q], 0.0);
        for (int k = 0; k < kLqMoments; ++k) {
            double v = m[k];
            for (int q = 0; q < kLqPoints; ++q) v -= op.V[k][q] * w[q];
            r[k] = v;
        }
        for (int q = 0; q < kLqPoints; ++q) {
            double v = 0.0;
            for (int k = 0; k < kLqMoments; ++k) v += op.P[q][k] * r[k];
            w[q] += v;
        }
    }
    if (iters > 0)
        for (int q = 0; q < kLqPoints; ++q) w[q] = std::max(w[q], 0.0);
    double rn = 0.0, mn = 0.0;
    for (int k = 0; k < kLqMoments; ++k) {
        double v = m[k];
        for (int q = 0; q < kLqPoints; ++q) v -= op.V[k][q] * w[q];
        rn += v * v;
        mn += m[k] * m[k];
    }
    return mn > 0.0 ? std::sqrt(rn / mn) : std::sqrt(rn);
}

void lq_point_gradients(const double* h, float* g27) {
    for (int q = 0; q < kLqPoints; ++q)
        for (int a = 0; a < 8; ++a) {
            const double* xi = kHex3Xi[q];
            const double sx = kHex8XiHat[a][0], sy = kHex8XiHat[a][1], sz = kHex8XiHat[a][2];
            const double fx = 1.0 + xi[0] * sx, fy = 1.0 + xi[1] * sy, fz = 1.0 + xi[2] * sz;
            g27[q * 24 + 0 * 8 + a] = (float)(0.125 * sx * fy * fz * (2.0 / h[0]));
            g27[q * 24 + 1 * 8 + a] = (float)(0.125 * fx * sy * fz * (2.0 / h[1]));
            g27[q * 24 + 2 * 8 + a] = (float)(0.125 * fx * fy * sz * (2.0 / h[2]));
        }
}

// Moments of the material sub-tets of a cell, in parametric units (full cell volume 8).
void lq_moments(const double (*sub)[4][3], int n_sub, double* m) {
    for (int k = 0; k < kLqMoments; ++k) m[k] = 0.0;
    for (int t = 0; t < n_sub; ++t) {
        const double(*X)[3] = sub[t];
        double J[3][3];
        const int rr[3] = {0, 1, 3};
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) J[r][d] = X[rr[r]][d] - X[2][d];
        const double detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
        for (int q = 0; q < kTet11Points; ++q) {
            const double N[4] = {kTet11Xi[q][0], kTet11Xi[q][1], 1.0 - kTet11Xi[q][0] - kTet11Xi[q][1] - kTet11Xi[q][2],
                                 kTet11Xi[q][2]};
            double xi[3];
            for (int d = 0; d < 3; ++d) xi[d] = N[0] * X[0][d] + N[1] * X[1][d] + N[2] * X[2][d] + N[3] * X[3][d];
            double p[kLqMoments];
            lq_monomials(xi, p);
            const double w = kTet11W[q] * detJ;
            for (int k = 0; k < kLqMoments; ++k) m[k] += w * p[k];
        }
    }
}

// All 27 tri-quadratic moments of the material sub-tets (degree-7 rule, exact to degree 6).
void lq_moments27(const LqOperators& op, const double (*sub)[4][3], int n_sub, double* m) {
    for (int k = 0; k < kLqPoints; ++k) m[k] = 0.0;
    for (int t = 0; t < n_sub; ++t) {
        const double(*X)[3] = sub[t];
        double J[3][3];
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) J[r][d] = X[r + 1][d] - X[0][d];
        const double detJ = std::fabs(J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                                      J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                      J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
        for (int q = 0; q < kGm7Points; ++q) {
            double xi[3];
            for (int d = 0; d < 3; ++d)
                xi[d] = op.gm_lambda[q][0] * X[0][d] + op.gm_lambda[q][1] * X[1][d] + op.gm_lambda[q][2] * X[2][d] +
                        op.gm_lambda[q][3] * X[3][d];
            double p[kLqPoints];
            lq_monomials27(xi, p);
            const double w = op.gm_w[q] * detJ;
            for (int k = 0; k < kLqPoints; ++k) m[k] += w * p[k];
        }
    }
}

namespace {

struct Rng {
    uint64_t seed, n = 0;
    double u() { return uniform01_at(seed, n++); }
    double range(double a, double b) { return a + (b - a) * u(); }
};

// vertex table of a unit lattice cell (1 x 1 x 1): ids ordered like an interior lattice cell
void unit_cell_vertices(HexSubdivVertex v[15]) {
    const int corner_gid[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int face_gid[6] = {8 + 2, 8 + 1, 8 + 3, 8 + 0, 8 + 4, 8 + 5};
    for (int c = 0; c < 8; ++c) {
        for (int d = 0; d < 3; ++d) v[c].xi[d] = kHex8XiHat[c][d];
        v[c].gid = (uint32_t)corner_gid[c];
    }
    for (int f = 0; f < 6; ++f) {
        for (int d = 0; d < 3; ++d) v[8 + f].xi[d] = kHexFaceXi[f][d];
        v[8 + f].gid = (uint32_t)face_gid[f];
    }
    for (int d = 0; d < 3; ++d) v[14].xi[d] = 0.0;
    v[14].gid = 14;
}

double sample_phi(int kind, const double* prm, const double* x) {
    switch (kind) {
        case 0:   // plane n.(x - p)
            return prm[0] * (x[0] - prm[3]) + prm[1] * (x[1] - prm[4]) + prm[2] * (x[2] - prm[5]);
        case 1: { // sphere, orientation prm[4]
            const double r = std::sqrt((x[0] - prm[0]) * (x[0] - prm[0]) + (x[1] - prm[1]) * (x[1] - prm[1]) +
                                       (x[2] - prm[2]) * (x[2] - prm[2])) - prm[3];
            return prm[4] * r;
        }
        default: { // gyroid patch
            const double a = prm[0] * x[0] + prm[1], b = prm[0] * x[1] + prm[2], c = prm[0] * x[2] + prm[3];
            return std::sin(a) * std::cos(b) + std::sin(b) * std::cos(c) + std::sin(c) * std::cos(a);
        }
    }
}

}  // namespace

// One sample; returns false if the draw does not give an intersected cell.
bool lq_make_sample(uint64_t seed, int64_t idx, int kinds, const LqOperators& op, float* x, float* y, double* m_out,
                    double* phi15) {
    Rng rng{seed ^ (0x5851F42D4C957F2DULL * (uint64_t)(idx + 1))};
    int avail[3], na = 0;
    for (int k = 0; k < 3; ++k)
        if (kinds & (1 << k)) avail[na++] = k;
    if (na == 0) return false;
    const int kind = avail[std::min(na - 1, (int)(rng.u() * na))];
    double prm[6];
    if (kind == 0) {
        const double z = rng.range(-1.0, 1.0), t = rng.range(0.0, 6.283185307179586);
        const double s = std::sqrt(1.0 - z * z);
        prm[0] = s * std::cos(t); prm[1] = s * std::sin(t); prm[2] = z;
        for (int d = 0; d < 3; ++d) prm[3 + d] = rng.range(-1.0, 1.0);
    } else if (kind == 1) {
        for (int d = 0; d < 3; ++d) prm[d] = rng.range(-4.0, 4.0);   // centre within 2h (h = 2)
        prm[3] = rng.range(0.2, 3.0);                                  // r in [0.1, 1.5] h
        prm[4] = rng.u() < 0.5 ? 1.0 : -1.0;
    } else {
        prm[0] = rng.range(0.5, 1.5) * 6.283185307179586 / 8.0;        // period about four cells
        for (int d = 1; d < 4; ++d) prm[d] = rng.range(0.0, 6.283185307179586);
    }
    HexSubdivVertex v[15];
    unit_cell_vertices(v);
    double pmax = 0.0;
    int r0 = -1;
    bool mixed = false;
    for (int i = 0; i < 15; ++i) {
        v[i].phi = sample_phi(kind, prm, v[i].xi);
        if (ls_region(v[i].phi) == 2) return false;
        if (i < 8) {
            pmax = std::max(pmax, std::fabs(v[i].phi));
            const int r = ls_region(v[i].phi);
            if (r0 < 0) r0 = r;
            else if (r != r0) mixed = true;
        }
    }
    if (!mixed) return false;
    static thread_local double sub[kMaxSubTets][4][3];
    int n_sub = 0;
    if (cut_hex8(v, 1, sub, &n_sub) != kCutOk) return false;
    double m27[kLqPoints];
    lq_moments27(op, sub, n_sub, m27);
    double w[kLqPoints];
    for (int q = 0; q < kLqPoints; ++q) {
        double s = 0.0;
        for (int k = 0; k < kLqPoints; ++k) s += op.V27inv[q][k] * m27[k];
        w[q] = s;
    }
    for (int c = 0; c < 8; ++c) x[c] = (float)(v[c].phi / pmax);
    for (int q = 0; q < kLqPoints; ++q) y[q] = (float)(w[q] / 8.0);
    if (m_out)
        for (int k = 0; k < kLqPoints; ++k) m_out[k] = m27[k];
    if (phi15)
        for (int i = 0; i < 15; ++i) phi15[i] = v[i].phi;
    return true;
}

}  // namespace mgpu

using namespace mgpu;

extern "C" mgpu_status mgpu_lq_dataset(int64_t n, uint64_t seed, int32_t kinds, float* x, float* y, double* moments) {
    // moments (optional): [n][27] tri-quadratic moments, a fastest
    MGPU_GUARD_BEGIN
    if (n < 0 || !x || !y) return set_error(MGPU_ERR_INVALID_ARG, "bad dataset arguments");
    if ((kinds & 7) == 0) return set_error(MGPU_ERR_INVALID_ARG, "no level-set kind selected");
    LqOperators op;
    lq_build_operators(&op);
    unsigned T = std::thread::hardware_concurrency();
    T = std::max(1u, std::min(T, 32u));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&, t]() {
            const int64_t a = n * t / T, b = n * (t + 1) / T;
            for (int64_t i = a; i < b; ++i) {
                // draw until an intersected cell comes out; the draw index depends only on i
                for (int64_t attempt = 0;; ++attempt) {
                    if (lq_make_sample(seed, i * 64 + attempt, kinds, op, x + 8 * i, y + kLqPoints * i,
                                       moments ? moments + kLqPoints * i : nullptr, nullptr))
                        break;
                }
            }
        });
    for (auto& t : th) t.join();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_lq_operators(double* V, double* P, double* gauss_w) {
    LqOperators op;
    lq_build_operators(&op);
    if (V) std::memcpy(V, op.V, sizeof(op.V));
    if (P) std::memcpy(P, op.P, sizeof(op.P));
    if (gauss_w) std::memcpy(gauss_w, op.gauss_w, sizeof(op.gauss_w));
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_lq_gm7(double* lambda, double* w) {
    LqOperators op;
    lq_build_operators(&op);
    if (lambda) std::memcpy(lambda, op.gm_lambda, sizeof(op.gm_lambda));
    if (w) std::memcpy(w, op.gm_w, sizeof(op.gm_w));
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_lq_fit(const double* m, const double* w0, int32_t iters, double* w, double* residual) {
    if (!m || !w0 || !w) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    LqOperators op;
    lq_build_operators(&op);
    const double r = lq_fit_nonnegative(op, m, w0, iters, w);
    if (residual) *residual = r;
    return MGPU_OK;
}

// Moments of one unit cell ([-1,1]^3) from the level set at its 15 regular-subdivision vertices
// (8 corners in hex8 order, faces 0..5, centre): XTK restatement + degree-7 rule.
extern "C" mgpu_status mgpu_lq_cell_moments(const double* phi15, int32_t phase, double* m27, int32_t* n_sub) {
    if (!phi15 || !m27) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    LqOperators op;
    lq_build_operators(&op);
    HexSubdivVertex v[15];
    const int corner_gid[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int face_gid[6] = {8 + 2, 8 + 1, 8 + 3, 8 + 0, 8 + 4, 8 + 5};
    for (int c = 0; c < 8; ++c) {
        for (int d = 0; d < 3; ++d) v[c].xi[d] = kHex8XiHat[c][d];
        v[c].gid = (uint32_t)corner_gid[c];
    }
    for (int f = 0; f < 6; ++f) {
        for (int d = 0; d < 3; ++d) v[8 + f].xi[d] = kHexFaceXi[f][d];
        v[8 + f].gid = (uint32_t)face_gid[f];
    }
    for (int d = 0; d < 3; ++d) v[14].xi[d] = 0.0;
    v[14].gid = 14;
    for (int i = 0; i < 15; ++i) v[i].phi = phi15[i];
    static thread_local double sub[kMaxSubTets][4][3];
    int ns = 0;
    if (cut_hex8(v, phase, sub, &ns) != kCutOk) return set_error(MGPU_ERR_GEOMETRY, "template not found");
    lq_moments27(op, sub, ns, m27);
    if (n_sub) *n_sub = ns;
    return MGPU_OK;
}
