// oracle.cpp -- TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of MORIS's bulk K+R assembly and
// of the set-up that feeds it (XTK cut of a level-0 HMR lattice, MSI adof numbering, CSR graph).
// It is the checker of libmoris_gpu.so: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.  It is written in the reference's own flow (global vertex/edge maps,
// template lookup by permutation id, dense T, dense B/D, per-entry row search on scatter) and
// shares no code with the product except the public C header for the spec/descriptor structs.
//
// Reference functions restated (file:line under projects/):
//   fill_matrix_and_RHS loop           SOL/DLA/src/cl_DLA_Solver_Interface.cpp:317-365
//   K_e = T^T K_p T, R_e = T^T R_p      FEM/MSI/src/cl_MSI_Equation_Object.cpp:789-842, 426-493
//   element topology (sorted unique)   FEM/MSI/src/cl_MSI_Equation_Object.cpp:199-291, 505-554
//   pdof values u_p = T u_adof         FEM/MSI/src/cl_MSI_Equation_Object.cpp:1100-1131
//   IG-cell loop, w* = w detJ, skip    FEM/INT/src/ELEM/cl_FEM_Element_Bulk.cpp:258-355
//   IG->IP point map                   MTK/src/interpolation/cl_MTK_Space_Interpolator.cpp:1381-1397
//   detJ hex / tet / tri                cl_MTK_Space_Interpolator.cpp:1006-1040, 1180-1210, 1120-1128
//   rectangular inverse Jacobian       cl_MTK_Space_Interpolator.cpp:676-700
//   dN/dx                              FEM/INT/src/IP/cl_FEM_Field_Interpolator.cpp:464-496
//   test strain / strain / flux        FEM/INT/src/CM/cl_FEM_CM_Struc_Linear.cpp:280, 349-590
//   D (3D, plane stress)               FEM/INT/src/CM/cl_FEM_CM_Struc_Linear_Isotropic.cpp:248-259, 374-390
//   diffusion flux / test strain       FEM/INT/src/CM/cl_FEM_CM_Diffusion_Linear_Isotropic.cpp:95-119, 302-350
//   IWG residual / jacobian            FEM/INT/src/IWG/cl_FEM_IWG_Isotropic_Struc_Linear_Bulk.cpp:43-205,
//                                      FEM/INT/src/IWG/cl_FEM_IWG_Diffusion_Bulk.cpp:51-323
//   regular subdivision                XTK/src/cl_XTK_Regular_Subdivision_Interface.hpp:110-360
//   node hierarchy                     XTK/src/cl_XTK_Node_Hierarchy_Interface.cpp:160-287, 672-968
//   templates                          XTK/src/cl_XTK_Node_Hierarchy_Template_Library.cpp:29-836
//   edge intersection                  GEN/src/cl_GEN_Interpolation.hpp:107-132,
//                                      GEN/src/PDV/cl_GEN_Intersection_Node.cpp:33,78-100
//   adof ids by type                   FEM/MSI/src/cl_MSI_Dof_Manager.cpp:703-745
//   CSR graph                          SOL/SOL_CORE/src/cl_Sparse_Matrix_EpetraFECrs.cpp:195-240
//   SumIntoGlobalValues                cl_Sparse_Matrix_EpetraFECrs.cpp:103-133
//   quadrature tables                  MTK/src/integration/cl_MTK_Integration_Coeffs_{Quad_2x2,
//                                      Hex_2x2x2,Hex_3x3x3,Tri_7,Tet_11,Tet_35}.hpp
//   HEX27 shape functions              MTK/src/interpolation/cl_MTK_Interpolation_Function_Lagrange_Hex27.hpp
//                                      (node order of get_param_coords; tensor-product form)
//   auto rules of quadratic IP cells   FEM/INT/src/CORE/cl_FEM_Set.cpp:3022-3091
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../include/moris_gpu.h"
#include "../include/moris_gpu_sideset.h"

namespace {

// ------------------------------------------------------------------------------------------------
// quadrature (reference literals, point-major here)
// ------------------------------------------------------------------------------------------------
const double G2 = 0.577350269189626;
const double kQuad2[4][2] = {{-G2, -G2}, {G2, -G2}, {G2, G2}, {-G2, G2}};
const double kHex2[8][3] = {{-G2, -G2, -G2}, {G2, -G2, -G2}, {G2, G2, -G2}, {-G2, G2, -G2},
                            {-G2, -G2, G2},  {G2, -G2, G2},  {G2, G2, G2},  {-G2, G2, G2}};
const double kTri7P[7][2] = {{0.333333333333333, 0.333333333333333}, {0.797426985353087, 0.101286507323456},
                             {0.101286507323456, 0.797426985353087}, {0.101286507323456, 0.101286507323456},
                             {0.059715871789770, 0.470142064105115}, {0.470142064105115, 0.059715871789770},
                             {0.470142064105115, 0.470142064105115}};
const double kTri7W[7] = {0.225000000000000, 0.125939180544827, 0.125939180544827, 0.125939180544827,
                          0.132394152788506, 0.132394152788506, 0.132394152788506};
const double A11 = 0.7857142857142860, B11 = 0.0714285714285714, C11 = 0.3994035761667990,
             D11 = 0.1005964238332010;
const double kTet11P[11][3] = {{A11, B11, B11}, {B11, A11, B11}, {B11, B11, B11}, {B11, B11, A11},
                               {C11, C11, D11}, {C11, D11, D11}, {C11, D11, C11}, {D11, C11, D11},
                               {D11, C11, C11}, {D11, D11, C11}, {0.25, 0.25, 0.25}};
double tet11_weight(int q) {
    const double w[11] = {0.0076222222222222, 0.0076222222222222, 0.0076222222222222, 0.0076222222222222,
                          0.0248888888888889, 0.0248888888888889, 0.0248888888888889, 0.0248888888888889,
                          0.0248888888888889, 0.0248888888888889, -0.0131555555555556};
    return 6.0 * w[q];
}

// hex8 / quad4 parametric vertices
const int kHexV[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                         {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
const int kQuadV[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};

void shape_N(int dim, const double* xi, double* N) {
    if (dim == 3) {
        const double x = xi[0], y = xi[1], z = xi[2];
        // cl_MTK_Interpolation_Function_Lagrange_Hex8.hpp eval_N
        N[0] = -(y - 1.0) * (x - 1.0) * (z - 1.0) * 0.125;
        N[1] = (y - 1.0) * (x + 1.0) * (z - 1.0) * 0.125;
        N[2] = -(y + 1.0) * (x + 1.0) * (z - 1.0) * 0.125;
        N[3] = (y + 1.0) * (x - 1.0) * (z - 1.0) * 0.125;
        N[4] = (y - 1.0) * (x - 1.0) * (z + 1.0) * 0.125;
        N[5] = -(y - 1.0) * (x + 1.0) * (z + 1.0) * 0.125;
        N[6] = (y + 1.0) * (x + 1.0) * (z + 1.0) * 0.125;
        N[7] = -(y + 1.0) * (x - 1.0) * (z + 1.0) * 0.125;
    } else {
        const double x = xi[0], y = xi[1];
        N[0] = (1.0 - x) * (1.0 - y) * 0.25;
        N[1] = (1.0 + x) * (1.0 - y) * 0.25;
        N[2] = (1.0 + x) * (1.0 + y) * 0.25;
        N[3] = (1.0 - x) * (1.0 + y) * 0.25;
    }
}

// dN/dxi [dim][nb]
void shape_dN(int dim, const double* xi, double dN[3][8]) {
    if (dim == 3) {
        for (int a = 0; a < 8; ++a) {
            const double sx = kHexV[a][0], sy = kHexV[a][1], sz = kHexV[a][2];
            dN[0][a] = 0.125 * sx * (1.0 + sy * xi[1]) * (1.0 + sz * xi[2]);
            dN[1][a] = 0.125 * sy * (1.0 + sx * xi[0]) * (1.0 + sz * xi[2]);
            dN[2][a] = 0.125 * sz * (1.0 + sx * xi[0]) * (1.0 + sy * xi[1]);
        }
    } else {
        for (int a = 0; a < 4; ++a) {
            const double sx = kQuadV[a][0], sy = kQuadV[a][1];
            dN[0][a] = 0.25 * sx * (1.0 + sy * xi[1]);
            dN[1][a] = 0.25 * sy * (1.0 + sx * xi[0]);
        }
    }
}

// HEX_3x3x3 (cl_MTK_Integration_Coeffs_Hex_3x3x3.hpp): the tensor Gauss rule; the weight of a point
// is the reference literal of its class (number of zero coordinates: 0, 1, 2, 3).
const double G3 = 0.774596669241483;
const double kHex3Wclass[4] = {0.171467764060357, 0.274348422496571, 0.438957475994513, 0.702331961591221};
void hex3_rule(double (*xi)[3], double* w) {
    const double g1[3] = {-G3, 0.0, G3};
    int q = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                xi[q][0] = g1[a];
                xi[q][1] = g1[b];
                xi[q][2] = g1[c];
                w[q] = kHex3Wclass[(a == 1) + (b == 1) + (c == 1)];
                ++q;
            }
}

// TET_35 (cl_MTK_Integration_Coeffs_Tet_35.hpp): five orbits of barycentric tuples under
// permutation; a point is the first three barycentric coordinates of a distinct permutation.
// Weights as in the reference file (sum 1, detJ = volume).
void tet35_rule(double (*xi)[3], double* w) {
    struct Orbit { double l[4]; double w; };
    const Orbit orb[5] = {
        {{0.9197896733368800, 0.0267367755543735, 0.0267367755543735, 0.0267367755543735}, 0.0021900463965388},
        {{0.1740356302468940, 0.7477598884818090, 0.0391022406356488, 0.0391022406356488}, 0.0143395670177665},
        {{0.4547545999844830, 0.4547545999844830, 0.0452454000155172, 0.0452454000155172}, 0.0250305395686746},
        {{0.5031186450145980, 0.2232010379623150, 0.2232010379623150, 0.0504792790607720}, 0.0479839333057554},
        {{0.2500000000000000, 0.2500000000000000, 0.2500000000000000, 0.2500000000000000}, 0.0931745731195340}};
    int q = 0;
    for (const Orbit& o : orb) {
        std::array<double, 4> l = {o.l[0], o.l[1], o.l[2], o.l[3]};
        std::sort(l.begin(), l.end());
        do {
            xi[q][0] = l[0];
            xi[q][1] = l[1];
            xi[q][2] = l[2];
            w[q] = o.w;
            ++q;
        } while (std::next_permutation(l.begin(), l.end()));
    }
    if (q != 35) throw std::runtime_error("TET_35 orbit count");
}

// HEX27 (cl_MTK_Interpolation_Function_Lagrange_Hex27.hpp): parametric node coordinates in the
// reference's node order, and the functions as tensor products of the 1D quadratic Lagrange
// polynomials through -1, 0, 1.
const int kHex27V[27][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
                            {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1}, {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
                            {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},  {0, 1, 1},   {-1, 0, 1},  {0, 0, 0},
                            {0, 0, -1},   {0, 0, 1},   {-1, 0, 0}, {1, 0, 0},  {0, -1, 0},  {0, 1, 0}};
double lag2(int node, double s) {   // 1D quadratic Lagrange function of node -1 / 0 / 1
    return node == 0 ? (1.0 - s) * (1.0 + s) : 0.5 * s * (s + (double)node);
}
double dlag2(int node, double s) {
    return node == 0 ? -2.0 * s : s + 0.5 * (double)node;
}
void hex27_N(const double* xi, double* N) {
    for (int a = 0; a < 27; ++a) N[a] = lag2(kHex27V[a][0], xi[0]) * lag2(kHex27V[a][1], xi[1]) * lag2(kHex27V[a][2], xi[2]);
}
void hex27_dN(const double* xi, double dN[3][27]) {
    for (int a = 0; a < 27; ++a)
        for (int i = 0; i < 3; ++i) {
            double v = 1.0;
            for (int d = 0; d < 3; ++d) v *= d == i ? dlag2(kHex27V[a][d], xi[d]) : lag2(kHex27V[a][d], xi[d]);
            dN[i][a] = v;
        }
}

// IP-cell bases of nb = 4, 8 or 27
void basis_N(int dim, int nb, const double* xi, double* N) {
    if (nb == 27) hex27_N(xi, N);
    else shape_N(dim, xi, N);
}
void basis_dN(int dim, int nb, const double* xi, double dN[3][27]) {
    if (nb == 27) {
        hex27_dN(xi, dN);
        return;
    }
    double d8[3][8];
    shape_dN(dim, xi, d8);
    for (int i = 0; i < dim; ++i)
        for (int a = 0; a < nb; ++a) dN[i][a] = d8[i][a];
}
// parametric coordinate of basis a of an nb-basis cell (for its nodal position)
void basis_xi(int dim, int nb, int a, double* xi) {
    for (int d = 0; d < dim; ++d) xi[d] = nb == 27 ? kHex27V[a][d] : (dim == 3 ? kHexV[a][d] : kQuadV[a][d]);
}

// ------------------------------------------------------------------------------------------------
// level sets (analytic GEN fields)
// ------------------------------------------------------------------------------------------------
struct LS {
    int kind = MGPU_LS_NONE, dim = 3;
    std::vector<double> spheres;
    double gs = 1.0;
    double plane[4] = {0, 0, 0, 0};
    double eval(const double* x) const {
        if (kind == MGPU_LS_SPHERES) {
            double best = INFINITY;
            for (size_t i = 0; i < spheres.size() / 4; ++i) {
                const double* s = &spheres[4 * i];
                double r2 = 0.0;
                for (int d = 0; d < dim; ++d) r2 += (x[d] - s[d]) * (x[d] - s[d]);
                best = std::min(best, std::sqrt(r2) - s[3]);
            }
            return best;
        }
        if (kind == MGPU_LS_GYROID) {
            const double a = gs * x[0], b = gs * x[1], c = dim == 3 ? gs * x[2] : 0.0;
            return std::sin(a) * std::cos(b) + std::sin(b) * std::cos(c) + std::sin(c) * std::cos(a);
        }
        if (kind == MGPU_LS_PLANE) {
            double v = plane[3];
            for (int d = 0; d < dim; ++d) v += plane[d] * x[d];
            return v;
        }
        return 1.0;
    }
};
int region(double phi) {
    if (std::fabs(phi) <= 1e-12) return 2;
    return phi < 0 ? 0 : 1;
}

uint64_t smix(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double gauss_draw(uint64_t seed, uint64_t cell) {
    const double u1 = 1.0 - (double)(smix(seed, 2 * cell) >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(smix(seed, 2 * cell + 1) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// ------------------------------------------------------------------------------------------------
// node-hierarchy template library: permutation id -> connectivity (cl_XTK_Node_Hierarchy_Template
// _Library.cpp load_3d_template / load_2d_template)
// ------------------------------------------------------------------------------------------------
struct Tmpl { int n; int c[6][4]; };
bool load_3d_template(int id, Tmpl& t) {
    static const int g3a[] = {320, 32, 203, 251, 512, 125, 140, 401, 14, 453, 534, 345};
    static const int g3b[] = {230, 302, 23, 521, 152, 215, 410, 41, 104, 543, 354, 435};
    static const int g4[6][12] = {
        {5420, 5240, 425, 3501, 1503, 3051, 1053, 4312, 2314, 4132, 2134, 245},
        {4502, 4052, 1243, 2504, 2054, 5310, 5130, 135, 315, 3421, 3241, 1423},
        {4250, 2450, 4205, 2405, 5031, 5013, 531, 513, 3124, 1342, 1324, 3142},
        {5024, 524, 5042, 542, 3150, 1350, 3105, 1305, 4231, 2431, 4213, 2413},
        {4520, 2540, 4025, 2045, 5301, 5103, 351, 153, 3412, 3214, 1432, 1234},
        {5402, 452, 5204, 254, 3510, 1530, 3015, 1035, 4321, 2341, 4123, 2143}};
    static const int c3[2][4][4] = {{{0, 4, 5, 6}, {4, 1, 5, 6}, {1, 2, 5, 6}, {1, 3, 2, 6}},
                                    {{4, 0, 5, 6}, {1, 4, 5, 6}, {2, 1, 5, 6}, {3, 1, 2, 6}}};
    static const int c4[6][6][4] = {
        {{0, 3, 7, 6}, {4, 3, 6, 7}, {4, 3, 7, 5}, {1, 2, 7, 5}, {2, 4, 7, 5}, {4, 2, 7, 6}},
        {{3, 0, 7, 6}, {3, 4, 6, 7}, {3, 4, 7, 5}, {2, 1, 7, 5}, {4, 2, 7, 5}, {2, 4, 7, 6}},
        {{4, 0, 6, 7}, {0, 3, 6, 7}, {3, 5, 6, 7}, {5, 1, 6, 7}, {1, 2, 6, 7}, {2, 4, 6, 7}},
        {{0, 4, 6, 7}, {3, 0, 6, 7}, {5, 3, 6, 7}, {1, 5, 6, 7}, {2, 1, 6, 7}, {4, 2, 6, 7}},
        {{4, 0, 5, 7}, {0, 3, 5, 7}, {5, 3, 6, 7}, {1, 2, 6, 7}, {2, 4, 5, 7}, {2, 5, 6, 7}},
        {{0, 4, 5, 7}, {3, 0, 5, 7}, {3, 5, 6, 7}, {2, 1, 6, 7}, {4, 2, 5, 7}, {5, 2, 6, 7}}};
    for (int i = 0; i < 12; ++i) {
        if (g3a[i] == id) { t.n = 4; std::memcpy(t.c, c3[0], sizeof(c3[0])); return true; }
        if (g3b[i] == id) { t.n = 4; std::memcpy(t.c, c3[1], sizeof(c3[1])); return true; }
    }
    for (int g = 0; g < 6; ++g)
        for (int i = 0; i < 12; ++i)
            if (g4[g][i] == id) { t.n = 6; std::memcpy(t.c, c4[g], sizeof(c4[g])); return true; }
    // vertices on the isocontour: one or two intersected edges
    static const int c1[6][2][4] = {{{0, 4, 2, 3}, {4, 1, 2, 3}}, {{0, 1, 4, 3}, {0, 4, 2, 3}},
                                    {{0, 1, 4, 3}, {4, 1, 2, 3}}, {{0, 1, 2, 4}, {4, 1, 2, 3}},
                                    {{0, 4, 2, 3}, {0, 1, 2, 4}}, {{0, 1, 4, 3}, {0, 1, 2, 4}}};
    if (id >= 10000 && id <= 10005) {
        t.n = 2;
        for (int i = 0; i < 2; ++i)
            for (int k = 0; k < 4; ++k) t.c[i][k] = c1[id - 10000][i][k];
        return true;
    }
    struct T2 { int id; int c[3][4]; };
    static const T2 two[24] = {
        {10250, {{0, 1, 4, 5}, {0, 4, 2, 5}, {0, 5, 2, 3}}}, {10520, {{0, 1, 5, 4}, {0, 4, 5, 3}, {0, 5, 2, 3}}},
        {10260, {{0, 1, 4, 5}, {0, 1, 5, 3}, {0, 4, 2, 5}}}, {10620, {{0, 1, 5, 3}, {0, 5, 4, 3}, {0, 5, 2, 4}}},
        {10560, {{0, 1, 5, 4}, {0, 1, 2, 5}, {0, 4, 5, 3}}}, {10650, {{0, 1, 2, 5}, {0, 5, 2, 4}, {0, 5, 4, 3}}},
        {10640, {{0, 1, 2, 5}, {5, 1, 4, 3}, {1, 2, 5, 4}}}, {10460, {{0, 1, 2, 5}, {0, 1, 5, 4}, {4, 1, 5, 3}}},
        {10230, {{5, 4, 2, 3}, {1, 4, 5, 3}, {0, 1, 5, 3}}}, {10320, {{5, 2, 4, 3}, {0, 5, 4, 3}, {0, 1, 5, 3}}},
        {10130, {{5, 1, 2, 3}, {5, 4, 1, 3}, {0, 4, 5, 3}}}, {10310, {{5, 1, 2, 3}, {4, 5, 2, 3}, {0, 5, 4, 3}}},
        {10120, {{0, 5, 2, 3}, {0, 4, 5, 3}, {4, 1, 5, 3}}}, {10210, {{0, 5, 2, 3}, {5, 4, 2, 3}, {5, 1, 4, 3}}},
        {10140, {{0, 4, 2, 5}, {4, 1, 2, 5}, {5, 1, 2, 3}}}, {10410, {{0, 5, 2, 4}, {4, 5, 2, 3}, {5, 1, 2, 3}}},
        {10150, {{0, 4, 2, 5}, {4, 1, 2, 5}, {0, 5, 2, 3}}}, {10510, {{5, 1, 2, 4}, {5, 4, 2, 3}, {0, 5, 2, 3}}},
        {10450, {{0, 5, 2, 4}, {0, 1, 2, 5}, {4, 5, 2, 3}}}, {10540, {{0, 1, 2, 5}, {5, 1, 2, 4}, {5, 4, 2, 3}}},
        {10360, {{0, 1, 5, 3}, {0, 1, 4, 5}, {1, 2, 4, 5}}}, {10630, {{0, 1, 5, 3}, {1, 4, 5, 3}, {1, 2, 5, 4}}},
        {10340, {{0, 1, 4, 5}, {4, 1, 2, 5}, {1, 2, 5, 3}}}, {10430, {{0, 1, 5, 4}, {4, 1, 5, 3}, {1, 2, 5, 3}}}};
    for (const T2& x : two)
        if (x.id == id) {
            t.n = 3;
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 4; ++k) t.c[i][k] = x.c[i][k];
            return true;
        }
    return false;
}
bool load_2d_template(int id, Tmpl& t) {
    const int c1[3][2][3] = {{{0, 3, 2}, {3, 1, 2}}, {{0, 1, 3}, {0, 3, 2}}, {{0, 1, 3}, {1, 2, 3}}};
    if (id >= 10 && id <= 12) {
        t.n = 2;
        for (int i = 0; i < 2; ++i)
            for (int k = 0; k < 3; ++k) t.c[i][k] = c1[id - 10][i][k];
        return true;
    }
    t.n = 3;
    const int c[3][3][3] = {{{0, 3, 2}, {3, 1, 4}, {3, 4, 2}},
                            {{0, 3, 4}, {3, 2, 4}, {3, 1, 2}},
                            {{0, 1, 3}, {0, 3, 4}, {3, 2, 4}}};
    if (id < 1 || id > 3) return false;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) t.c[i][k] = c[id - 1][i][k];
    return true;
}

// regular subdivision tables (cl_XTK_Regular_Subdivision_Interface.hpp)
const int kR24[24][4] = {{0, 8, 1, 14},  {1, 8, 5, 14},  {4, 5, 8, 14},  {0, 4, 8, 14},  {1, 9, 2, 14},
                         {2, 9, 6, 14},  {5, 6, 9, 14},  {1, 5, 9, 14},  {2, 10, 3, 14}, {2, 6, 10, 14},
                         {6, 7, 10, 14}, {3, 10, 7, 14}, {0, 3, 11, 14}, {3, 7, 11, 14}, {4, 11, 7, 14},
                         {0, 11, 4, 14}, {0, 1, 12, 14}, {1, 2, 12, 14}, {2, 3, 12, 14}, {0, 12, 3, 14},
                         {4, 13, 5, 14}, {5, 13, 6, 14}, {6, 13, 7, 14}, {4, 7, 13, 14}};
const double kR24New[7][3] = {{0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 1}, {0, 0, 0}};
const int kR4[4][3] = {{0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4}};
const int kTetE[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
const int kTriE[3][2] = {{0, 1}, {1, 2}, {2, 0}};

int tet_shared(int e1, int e2) {
    for (int a : kTetE[e1])
        for (int b : kTetE[e2])
            if (a == b) return a;
    return -1;
}

// ------------------------------------------------------------------------------------------------
// problem
// ------------------------------------------------------------------------------------------------
struct OracleProblem {
    mgpu_problem_spec spec{};
    LS ls;
    int dim = 3, nb = 8, nf = 1, nm = 1;
    int nb_ip = 8;   // IP-cell bases: nb (the box vertices), or 27 for HEX27
    // THB (one refinement level): refined level-0 cells, each element's level and its cell on
    // that level's lattice
    std::vector<char> refined;
    std::vector<int8_t> elem_level;
    std::vector<int64_t> elem_lcell;
    int n[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    int64_t n_cells = 0, n_nodes = 0;
    std::vector<double> node_phi;
    // elements (material clusters) in lattice order
    std::vector<int32_t> elem_cell;
    std::vector<double> material;
    std::vector<int32_t> ig_ptr;
    std::vector<double> ig_xi;
    std::vector<double> ip_box;
    std::vector<int32_t> t_ptr, t_adof;
    std::vector<double> t_weight;
    int32_t N = 0;
    int64_t n_cut = 0;
    // ghost stabilisation (spec.ghost): (leader, follower) element pairs coupled in the graph
    std::vector<int32_t> couplings;
    // graph
    std::vector<int32_t> row_ptr, col_idx;
};

int64_t nid3(const OracleProblem& p, int i, int j, int k) {
    return i + (int64_t)(p.n[0] + 1) * (j + (int64_t)(p.n[1] + 1) * k);
}
void node_x(const OracleProblem& p, int64_t nid, double* x) {
    const int i = (int)(nid % (p.n[0] + 1));
    const int j = (int)((nid / (p.n[0] + 1)) % (p.n[1] + 1));
    const int k = p.dim == 3 ? (int)(nid / ((int64_t)(p.n[0] + 1) * (p.n[1] + 1))) : 0;
    x[0] = p.lo[0] + (double)i * p.h[0];
    x[1] = p.lo[1] + (double)j * p.h[1];
    x[2] = p.dim == 3 ? p.lo[2] + (double)k * p.h[2] : 0.0;
}
void cell_nodes(const OracleProblem& p, int64_t c, int64_t* nodes) {
    const int i = (int)(c % p.n[0]);
    const int j = (int)((c / p.n[0]) % p.n[1]);
    const int k = p.dim == 3 ? (int)(c / ((int64_t)p.n[0] * p.n[1])) : 0;
    if (p.dim == 3) {
        for (int v = 0; v < 8; ++v)
            nodes[v] = nid3(p, i + (kHexV[v][0] + 1) / 2, j + (kHexV[v][1] + 1) / 2, k + (kHexV[v][2] + 1) / 2);
    } else {
        for (int v = 0; v < 4; ++v)
            nodes[v] = (i + (kQuadV[v][0] + 1) / 2) + (int64_t)(p.n[0] + 1) * (j + (kQuadV[v][1] + 1) / 2);
    }
}
// global coordinates of parametric point xi in cell c (interpolation over the cell vertices)
void cell_x(const OracleProblem& p, int64_t c, const double* xi, double* x) {
    int64_t nodes[8];
    cell_nodes(p, c, nodes);
    double N[8];
    shape_N(p.dim, xi, N);
    x[0] = x[1] = x[2] = 0.0;
    for (int v = 0; v < p.nb; ++v) {
        double xv[3];
        node_x(p, nodes[v], xv);
        for (int d = 0; d < p.dim; ++d) x[d] += N[v] * xv[d];
    }
}
bool intersected(const OracleProblem& p, int64_t c) {
    int64_t nodes[8];
    cell_nodes(p, c, nodes);
    const int r0 = region(p.node_phi[nodes[0]]);
    for (int v = 1; v < p.nb; ++v)
        if (region(p.node_phi[nodes[v]]) != r0) return true;
    return false;
}

// Node hierarchy of one IG cell of the regular subdivision (tet in 3D, tri in 2D; cl_XTK_Node_
// Hierarchy_Interface.cpp:160-287, 672-968): g, f, x are its vertices' global ids, level-set values
// and parametric coordinates.  Appends the vertex coordinates of its sub-cells of phase `phase`
// (region() of their non-interface vertices) to sub.
void nh_cell(int D, const uint32_t* g, const double* f, const std::array<double, 3>* x, int phase,
             std::vector<double>& sub) {
    const int nv = D + 1;
    // intersected edges: (ordinal, key, parametric point)
    struct IE { int ord; uint64_t key; double x[3]; };
    std::vector<IE> ies;
    const int ne = D == 3 ? 6 : 3;
    for (int e = 0; e < ne; ++e) {
        const int l0 = D == 3 ? kTetE[e][0] : kTriE[e][0];
        const int l1 = D == 3 ? kTetE[e][1] : kTriE[e][1];
        const double f0 = f[l0], f1 = f[l1];
        // an endpoint on the interface: no general intersection on this edge
        if (region(f0) == 2 || region(f1) == 2) continue;
        if (region(f0) == region(f1)) continue;
        // parent nodes ordered by global id
        const int la = g[l0] < g[l1] ? l0 : l1;
        const int lb = g[l0] < g[l1] ? l1 : l0;
        const double fa = f[la], fb = f[lb];
        const double loc = (2.0 * 0.0 - fb - fa) / (fb - fa);
        if (!(std::fabs(loc) <= 1.0)) continue;
        IE ie;
        ie.ord = e;
        ie.key = ((uint64_t)g[la] << 32) | g[lb];
        for (int d = 0; d < 3; ++d) ie.x[d] = 0.5 * (1.0 - loc) * x[la][d] + 0.5 * (1.0 + loc) * x[lb][d];
        ies.push_back(ie);
    }
    auto phase_of = [&](int local) { return region(f[local]); };
    if (ies.empty()) {
        int ph = -1;
        for (int q = 0; q < nv && ph < 0; ++q)
            if (phase_of(q) != 2) ph = phase_of(q);
        if (ph == phase)
            for (int q = 0; q < nv; ++q)
                for (int d = 0; d < D; ++d) sub.push_back(x[q][d]);
        return;
    }
    Tmpl tm;
    int nodeord[4] = {0, 1, 2, 3};
    std::vector<IE> sorted = ies;
    if (D == 3) {
        std::stable_sort(sorted.begin(), sorted.end(), [](const IE& a, const IE& b) { return a.key < b.key; });
        int perm = 0;
        if (sorted.size() == 1) {
            perm = 10000 + sorted[0].ord;
        } else if (sorted.size() == 2) {
            perm = 10000 + 100 * (sorted[0].ord + 1) + 10 * (sorted[1].ord + 1);
        } else if (sorted.size() == 3) {
            const int L = sorted[0].ord, M = sorted[1].ord, H = sorted[2].ord;
            const int n0 = tet_shared(L, M);
            auto other = [&](int e) { return kTetE[e][0] == n0 ? kTetE[e][1] : kTetE[e][0]; };
            nodeord[0] = n0; nodeord[1] = other(L); nodeord[2] = other(M); nodeord[3] = other(H);
            perm = L + 10 * M + 100 * H;
        } else if (sorted.size() == 4) {
            const int L = sorted[0].ord, ML = sorted[1].ord, MH = sorted[2].ord, H = sorted[3].ord;
            perm = L + 10 * ML + 100 * MH + 1000 * H;
            const int HL = tet_shared(L, H), HMH = tet_shared(MH, H), HML = tet_shared(ML, H);
            const int LML = tet_shared(L, ML), LMH = tet_shared(L, MH), MLMH = tet_shared(ML, MH);
            if (HL < 0) { nodeord[0] = HMH; nodeord[1] = HML; nodeord[2] = LMH; nodeord[3] = LML; }
            else if (HMH < 0) { nodeord[0] = HL; nodeord[1] = HML; nodeord[2] = LMH; nodeord[3] = MLMH; }
            else { nodeord[0] = HL; nodeord[1] = HMH; nodeord[2] = LML; nodeord[3] = MLMH; }
        } else {
            throw std::runtime_error("unexpected intersected edge count");
        }
        if (!load_3d_template(perm, tm)) throw std::runtime_error("template not found");
    } else {
        std::stable_sort(sorted.begin(), sorted.end(), [](const IE& a, const IE& b) { return a.ord < b.ord; });
        int id;
        if (sorted.size() == 1) id = 10 + sorted[0].ord;
        else if (sorted.size() == 2) id = sorted[0].ord + sorted[1].ord;
        else throw std::runtime_error("unexpected 2D cut");
        if (!load_2d_template(id, tm)) throw std::runtime_error("template not found");
    }
    for (int cidx = 0; cidx < tm.n; ++cidx) {
        int ph = -1;
        for (int q = 0; q < nv && ph < 0; ++q)
            if (tm.c[cidx][q] < nv && phase_of(nodeord[tm.c[cidx][q]]) != 2) ph = phase_of(nodeord[tm.c[cidx][q]]);
        if (ph != phase) continue;
        for (int q = 0; q < nv; ++q) {
            const int o = tm.c[cidx][q];
            if (o < nv) {
                for (int d = 0; d < D; ++d) sub.push_back(x[nodeord[o]][d]);
            } else {
                for (int d = 0; d < D; ++d) sub.push_back(sorted[o - nv].x[d]);
            }
        }
    }
}

// XTK decomposition of all intersected cells; fills per-cell material sub-cells.
void decompose(OracleProblem& p, std::map<int64_t, std::vector<double>>& cell_sub) {
    const int D = p.dim;
    const int nxf = (p.n[0] + 1) * p.n[1] * p.n[2];
    const int nyf = p.n[0] * (p.n[1] + 1) * p.n[2];
    const int64_t nfaces = D == 3 ? (int64_t)nxf + nyf + (int64_t)p.n[0] * p.n[1] * (p.n[2] + 1) : 0;
    // global subdivision vertices created on request: gid -> (phi)
    std::map<uint32_t, double> vphi;
    for (int64_t c = 0; c < p.n_cells; ++c) {
        if (!intersected(p, c)) continue;
        const int i = (int)(c % p.n[0]);
        const int j = (int)((c / p.n[0]) % p.n[1]);
        const int k = D == 3 ? (int)(c / ((int64_t)p.n[0] * p.n[1])) : 0;
        // local vertex table: gid, parametric coordinates wrt this cell
        std::vector<uint32_t> gid;
        std::vector<std::array<double, 3>> xi;
        int64_t nodes[8];
        cell_nodes(p, c, nodes);
        for (int v = 0; v < p.nb; ++v) {
            gid.push_back((uint32_t)nodes[v]);
            std::array<double, 3> a{0, 0, 0};
            for (int d = 0; d < D; ++d) a[d] = D == 3 ? kHexV[v][d] : kQuadV[v][d];
            xi.push_back(a);
            vphi[(uint32_t)nodes[v]] = p.node_phi[nodes[v]];
        }
        if (D == 3) {
            const int64_t fids[6] = {nxf + i + (int64_t)p.n[0] * (j + (int64_t)(p.n[1] + 1) * k),
                                     (i + 1) + (int64_t)(p.n[0] + 1) * (j + (int64_t)p.n[1] * k),
                                     nxf + i + (int64_t)p.n[0] * ((j + 1) + (int64_t)(p.n[1] + 1) * k),
                                     i + (int64_t)(p.n[0] + 1) * (j + (int64_t)p.n[1] * k),
                                     (int64_t)nxf + nyf + i + (int64_t)p.n[0] * (j + (int64_t)p.n[1] * k),
                                     (int64_t)nxf + nyf + i + (int64_t)p.n[0] * (j + (int64_t)p.n[1] * (k + 1))};
            for (int f = 0; f < 7; ++f) {
                const uint32_t g = f < 6 ? (uint32_t)(p.n_nodes + fids[f]) : (uint32_t)(p.n_nodes + nfaces + c);
                std::array<double, 3> a{kR24New[f][0], kR24New[f][1], kR24New[f][2]};
                if (!vphi.count(g)) {   // first requesting cell creates the vertex
                    double x[3];
                    cell_x(p, c, a.data(), x);
                    vphi[g] = p.ls.eval(x);
                }
                gid.push_back(g);
                xi.push_back(a);
            }
        } else {
            const uint32_t g = (uint32_t)(p.n_nodes + c);
            std::array<double, 3> a{0, 0, 0};
            double x[3];
            cell_x(p, c, a.data(), x);
            vphi[g] = p.ls.eval(x);
            gid.push_back(g);
            xi.push_back(a);
        }
        // IG cells of the regular subdivision, then node hierarchy per IG cell
        const int ncell = D == 3 ? 24 : 4;
        const int nv = D + 1;
        std::vector<double> sub;
        for (int t = 0; t < ncell; ++t) {
            uint32_t tg[4];
            double tf[4];
            std::array<double, 3> tx[4];
            for (int q = 0; q < nv; ++q) {
                const int lv = D == 3 ? kR24[t][q] : kR4[t][q];
                tg[q] = gid[lv];
                tf[q] = vphi[gid[lv]];
                tx[q] = xi[lv];
            }
            nh_cell(D, tg, tf, tx, p.spec.material_phase, sub);
        }
        cell_sub[c] = std::move(sub);
    }
}

// Uniform quadratic B-spline on knots 0, 1, 2, 3 (Cox-de Boor, degree 2), zero outside [0, 3].
double bspline2(double t) {
    if (t <= 0.0 || t >= 3.0) return 0.0;
    if (t < 1.0) return 0.5 * t * t;
    if (t < 2.0) return 0.5 * (-2.0 * t * t + 6.0 * t - 3.0);
    return 0.5 * (3.0 - t) * (3.0 - t);
}

// Lattice coordinates (in cell units) of basis a of element e: the Lagrange node of the IP cell.
void basis_pos(const OracleProblem& p, size_t e, int a, double* x) {
    const int64_t c = p.elem_cell[e];
    const int ijk[3] = {(int)(c % p.n[0]), (int)((c / p.n[0]) % p.n[1]),
                        p.dim == 3 ? (int)(c / ((int64_t)p.n[0] * p.n[1])) : 0};
    double xi[3] = {0.0, 0.0, 0.0};
    basis_xi(p.dim, p.nb_ip, a, xi);
    for (int d = 0; d < 3; ++d) x[d] = d < p.dim ? (double)ijk[d] + 0.5 * (xi[d] + 1.0) : 0.0;
}

// T rows of p = 2 B-splines at the Lagrange nodes of each element (HMR level 0 T-matrix,
// cl_HMR_T_Matrix_Base.cpp:306-457, B-spline values at the Lagrange nodes; entries with
// |T| <= 1e-12 dropped as there).  Spline s of an axis starts at knot s - 2 (so n + 2 splines
// cover n cells); adofs are the used splines in ascending index, x fastest.
void bspline2_rows(OracleProblem& p) {
    const int64_t ns[3] = {p.n[0] + 2, p.n[1] + 2, p.dim == 3 ? p.n[2] + 2 : 1};
    // per node position: (spline index, weight) of every spline non-zero there
    auto row = [&](const double* x, std::vector<std::pair<int64_t, double>>& out) {
        std::vector<std::pair<int64_t, double>> ax[3];
        for (int d = 0; d < 3; ++d) {
            if (d >= p.dim) { ax[d].push_back({0, 1.0}); continue; }
            for (int64_t sp = 0; sp < ns[d]; ++sp) {
                const double v = bspline2(x[d] - (double)(sp - 2));
                if (std::fabs(v) > 1e-12) ax[d].push_back({sp, v});
            }
        }
        out.clear();
        for (auto& c : ax[2])
            for (auto& b : ax[1])
                for (auto& a : ax[0]) out.push_back({a.first + ns[0] * (b.first + ns[1] * c.first), a.second * b.second * c.second});
        std::sort(out.begin(), out.end());
    };
    std::map<int64_t, int32_t> adof;
    std::vector<std::pair<int64_t, double>> r;
    double x[3];
    for (size_t e = 0; e < p.elem_cell.size(); ++e)
        for (int v = 0; v < p.nb_ip; ++v) {
            basis_pos(p, e, v, x);
            row(x, r);
            for (auto& q : r) adof[q.first] = 0;
        }
    p.N = 0;
    for (auto& kv : adof) kv.second = p.N++;
    for (size_t e = 0; e < p.elem_cell.size(); ++e)
        for (int v = 0; v < p.nb_ip; ++v) {
            basis_pos(p, e, v, x);
            p.t_ptr.push_back((int32_t)p.t_adof.size());
            row(x, r);
            for (auto& q : r) {
                p.t_adof.push_back(adof.at(q.first));
                p.t_weight.push_back(q.second);
            }
        }
}

// THB T rows (two levels), in the reference's flow: per element the B-spline T-matrix of the
// element's bases -- its own level's 27 splines (identity columns for the active ones) and, on a
// level-1 element, the parent's active level-0 splines truncated to their inactive level-1
// children among the element's splines (calculate_truncated_t_matrix,
// cl_HMR_T_Matrix_Base.cpp:127-258, two-scale weights 1/4, 3/4, 3/4, 1/4 per axis) -- multiplied by
// the Lagrange matrix L (the element's 27 splines at its Lagrange nodes), T = L B, entries with
// |T| <= 1e-12 dropped (evaluate, :306-457).  Activity: a level-0 spline is active unless all
// level-0 cells of its support are refined; a level-1 spline is active if all level-1 cells of its
// support lie in refined level-0 cells.  Adofs: used level-0 splines ascending, then level-1.
void thb_rows(OracleProblem& p) {
    const int64_t n0[3] = {p.n[0], p.n[1], p.n[2]};
    const int64_t ns0[3] = {n0[0] + 2, n0[1] + 2, n0[2] + 2};
    const int64_t ns1[3] = {2 * n0[0] + 2, 2 * n0[1] + 2, 2 * n0[2] + 2};
    auto is_refined = [&](int64_t i, int64_t j, int64_t k) { return p.refined[i + n0[0] * (j + n0[1] * k)] != 0; };
    auto active = [&](int level, const int64_t* t) {
        // support cells of spline t on its level: t - 2 .. t per axis, inside the lattice
        const int64_t nl[3] = {level ? 2 * n0[0] : n0[0], level ? 2 * n0[1] : n0[1], level ? 2 * n0[2] : n0[2]};
        int n_cells = 0, n_refined = 0;
        for (int64_t k = t[2] - 2; k <= t[2]; ++k)
            for (int64_t j = t[1] - 2; j <= t[1]; ++j)
                for (int64_t i = t[0] - 2; i <= t[0]; ++i) {
                    if (i < 0 || j < 0 || k < 0 || i >= nl[0] || j >= nl[1] || k >= nl[2]) continue;
                    ++n_cells;
                    const bool r = level ? is_refined(i / 2, j / 2, k / 2) : is_refined(i, j, k);
                    n_refined += r;
                }
        return level == 0 ? n_refined < n_cells : n_refined == n_cells;
    };
    const double tw[4] = {0.25, 0.75, 0.75, 0.25};
    // per element: rows of (level, global spline index, weight)
    struct E3 { int level; int64_t sp; double w; };
    std::vector<std::vector<std::vector<E3>>> rows(p.elem_cell.size());
    std::map<int64_t, int32_t> used0, used1;
    for (size_t e = 0; e < p.elem_cell.size(); ++e) {
        const int lvl = p.elem_level[e];
        const int64_t nl[3] = {lvl ? 2 * n0[0] : n0[0], lvl ? 2 * n0[1] : n0[1], lvl ? 2 * n0[2] : n0[2]};
        const int64_t c = p.elem_lcell[e];
        const int64_t ijk[3] = {c % nl[0], (c / nl[0]) % nl[1], c / (nl[0] * nl[1])};
        // the element's splines (local j = a + 3 b + 9 c -> spline ijk + (a, b, c))
        int64_t own[27][3];
        for (int j = 0; j < 27; ++j) {
            own[j][0] = ijk[0] + j % 3;
            own[j][1] = ijk[1] + (j / 3) % 3;
            own[j][2] = ijk[2] + j / 9;
        }
        // B: columns (level, spline) -> 27 coefficients on the element's splines
        std::vector<std::pair<std::pair<int, int64_t>, std::array<double, 27>>> B;
        const int64_t* ns = lvl ? ns1 : ns0;
        for (int j = 0; j < 27; ++j)
            if (active(lvl, own[j])) {
                std::array<double, 27> col{};
                col[j] = 1.0;
                B.push_back({{lvl, own[j][0] + ns[0] * (own[j][1] + ns[1] * own[j][2])}, col});
            }
        if (lvl == 1) {
            const int64_t P[3] = {ijk[0] / 2, ijk[1] / 2, ijk[2] / 2};
            for (int jp = 0; jp < 27; ++jp) {
                const int64_t sp[3] = {P[0] + jp % 3, P[1] + (jp / 3) % 3, P[2] + jp / 9};
                if (!active(0, sp)) continue;
                std::array<double, 27> col{};
                // children 2 s - 2 + k, k = 0..3 per axis, that are inactive and among the element's splines
                for (int kz = 0; kz < 4; ++kz)
                    for (int ky = 0; ky < 4; ++ky)
                        for (int kx = 0; kx < 4; ++kx) {
                            const int64_t ch[3] = {2 * sp[0] - 2 + kx, 2 * sp[1] - 2 + ky, 2 * sp[2] - 2 + kz};
                            if (ch[0] < 0 || ch[1] < 0 || ch[2] < 0 || ch[0] >= ns1[0] || ch[1] >= ns1[1] || ch[2] >= ns1[2]) continue;
                            if (active(1, ch)) continue;
                            for (int j = 0; j < 27; ++j)
                                if (own[j][0] == ch[0] && own[j][1] == ch[1] && own[j][2] == ch[2]) {
                                    col[j] += tw[kx] * tw[ky] * tw[kz];
                                    break;
                                }
                        }
                bool nonzero = false;
                for (double v : col) nonzero = nonzero || std::fabs(v) > 1e-12;
                if (nonzero) B.push_back({{0, sp[0] + ns0[0] * (sp[1] + ns0[1] * sp[2])}, col});
            }
        }
        // L: the element's splines at its Lagrange nodes; T = L B
        rows[e].resize(p.nb_ip);
        for (int v = 0; v < p.nb_ip; ++v) {
            double xi[3];
            basis_xi(3, p.nb_ip, v, xi);
            double L[27];
            for (int j = 0; j < 27; ++j) {
                L[j] = 1.0;
                for (int d = 0; d < 3; ++d) L[j] *= bspline2((double)ijk[d] + 0.5 * (xi[d] + 1.0) - (double)(own[j][d] - 2));
            }
            for (const auto& col : B) {
                double t = 0.0;
                for (int j = 0; j < 27; ++j) t += L[j] * col.second[j];
                if (std::fabs(t) > 1e-12) {
                    rows[e][v].push_back({col.first.first, col.first.second, t});
                    (col.first.first == 0 ? used0 : used1)[col.first.second] = 0;
                }
            }
        }
    }
    p.N = 0;
    for (auto& kv : used0) kv.second = p.N++;
    for (auto& kv : used1) kv.second = p.N++;
    for (size_t e = 0; e < p.elem_cell.size(); ++e)
        for (int v = 0; v < p.nb_ip; ++v) {
            std::vector<std::pair<int32_t, double>> r;
            for (const E3& q : rows[e][v]) r.push_back({(q.level == 0 ? used0 : used1).at(q.sp), q.w});
            std::sort(r.begin(), r.end());
            p.t_ptr.push_back((int32_t)p.t_adof.size());
            for (auto& q : r) {
                p.t_adof.push_back(q.first);
                p.t_weight.push_back(q.second);
            }
        }
}

// ------------------------------------------------------------------------------------------------
// XTK basis-cluster enrichment (XTK/src/cl_XTK_Enrichment.cpp:276-400, 622-1510), restated from
// the geometry: subphases are the face-connected components of a cell's material sub-cells; two
// subphases of face-adjacent cells are neighbours when a material facet of one on the common face
// has its centroid inside a material facet of the other (the facets are compared in physical
// coordinates); a basis' support is the set of elements whose T rows hold it, and its enrichment
// levels are the components of that support under the neighbour relation, numbered in order of
// their first element.  Enriched adofs: basis by basis in ascending adof, level by level.
// ------------------------------------------------------------------------------------------------

// Components of the material sub-cells of one cell, [n_sub][D+1][D] in `sub`: sub-cells with D
// common vertices share a face.  Components in order of their first sub-cell.
std::vector<std::vector<int>> subphase_groups(int D, const std::vector<double>& sub) {
    const int nv = D + 1;
    const int n = (int)(sub.size() / (size_t)(nv * D));
    auto vtx_eq = [&](int s, int a, int t, int b) {
        for (int d = 0; d < D; ++d)
            if (sub[((size_t)s * nv + a) * D + d] != sub[((size_t)t * nv + b) * D + d]) return false;
        return true;
    };
    std::vector<int> comp(n, -1);
    std::vector<std::vector<int>> groups;
    for (int s0 = 0; s0 < n; ++s0) {
        if (comp[s0] >= 0) continue;
        const int g = (int)groups.size();
        groups.push_back({});
        std::vector<int> stack{s0};
        comp[s0] = g;
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            for (int t = 0; t < n; ++t) {
                if (comp[t] >= 0) continue;
                int common = 0;
                for (int a = 0; a < nv; ++a)
                    for (int b = 0; b < nv; ++b) common += vtx_eq(s, a, t, b) ? 1 : 0;
                if (common == D) {
                    comp[t] = g;
                    stack.push_back(t);
                }
            }
        }
    }
    for (int s = 0; s < n; ++s) groups[comp[s]].push_back(s);
    return groups;
}

// Material facets of an element on face (axis d, side +-1) of its cell, as physical vertex
// coordinates [facet][vertex][3]; an uncut element's face is the whole face (two triangles, or one
// segment in 2D).
std::vector<std::array<std::array<double, 3>, 3>> face_facets(const OracleProblem& p, size_t e, int d, double side) {
    const int D = p.dim, nv = D + 1;
    const double* box = p.ip_box.data() + e * 2 * D;
    auto phys = [&](const double* xi, std::array<double, 3>& x) {
        x = {0.0, 0.0, 0.0};
        for (int q = 0; q < D; ++q) x[q] = box[q] + 0.5 * (xi[q] + 1.0) * box[D + q];
    };
    std::vector<std::array<std::array<double, 3>, 3>> out;
    const int32_t s0 = p.ig_ptr[e], s1 = p.ig_ptr[e + 1];
    if (s0 == s1) {
        // the face's corners in cyclic order
        int ax[2] = {0, 0}, na = 0;
        for (int q = 0; q < D; ++q)
            if (q != d) ax[na++] = q;
        double cyc[4][3];
        const double pat[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (int k = 0; k < 4; ++k) {
            cyc[k][d] = side;
            cyc[k][ax[0]] = pat[k][0];
            if (D == 3) cyc[k][ax[1]] = pat[k][1];
        }
        std::array<std::array<double, 3>, 3> f{};
        if (D == 3) {
            const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
            for (auto& t : tri) {
                for (int k = 0; k < 3; ++k) phys(cyc[t[k]], f[k]);
                out.push_back(f);
            }
        } else {
            phys(cyc[0], f[0]);
            phys(cyc[1], f[1]);
            out.push_back(f);
        }
        return out;
    }
    for (int32_t s = s0; s < s1; ++s) {
        const double* X = p.ig_xi.data() + (size_t)s * nv * D;
        for (int omit = 0; omit < nv; ++omit) {
            std::array<std::array<double, 3>, 3> f{};
            int m = 0;
            bool on = true;
            for (int k = 0; k < nv; ++k) {
                if (k == omit) continue;
                on = on && std::fabs(X[k * D + d] - side) < 1e-12;
                phys(X + k * D, f[m++]);
            }
            if (on) out.push_back(f);
        }
    }
    return out;
}

// Point (on the face plane normal to axis d) inside a facet of that plane, with a tolerance.
bool in_facet(int D, int d, const std::array<double, 3>& x, const std::array<std::array<double, 3>, 3>& f) {
    int ax[2] = {0, 0}, na = 0;
    for (int q = 0; q < D; ++q)
        if (q != d) ax[na++] = q;
    if (D == 2) {
        const double a = f[0][ax[0]], b = f[1][ax[0]], lo = std::min(a, b), hi = std::max(a, b);
        const double tol = 1e-9 * (hi - lo);
        return x[ax[0]] > lo - tol && x[ax[0]] < hi + tol;
    }
    auto P = [&](const std::array<double, 3>& v, int k) { return v[ax[k]]; };
    const double x1 = P(f[0], 0), y1 = P(f[0], 1), x2 = P(f[1], 0), y2 = P(f[1], 1), x3 = P(f[2], 0), y3 = P(f[2], 1);
    const double det = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3);
    if (std::fabs(det) == 0.0) return false;
    const double l1 = ((y2 - y3) * (P(x, 0) - x3) + (x3 - x2) * (P(x, 1) - y3)) / det;
    const double l2 = ((y3 - y1) * (P(x, 0) - x3) + (x1 - x3) * (P(x, 1) - y3)) / det;
    const double l3 = 1.0 - l1 - l2;
    const double tol = 1e-9;
    return l1 > -tol && l2 > -tol && l3 > -tol;
}

// Enrich the T rows of p in place: the adof of each row entry becomes (basis, level).
void enrich_rows(OracleProblem& p) {
    const int D = p.dim;
    const size_t n_el = p.elem_cell.size();
    // elements of each cell
    std::map<int64_t, std::vector<size_t>> cell_elems;
    for (size_t e = 0; e < n_el; ++e) cell_elems[p.elem_cell[e]].push_back(e);
    // neighbour pairs across the faces between cells c and c + e_d
    std::vector<std::vector<size_t>> nbr(n_el);
    for (const auto& kv : cell_elems) {
        const int64_t c = kv.first;
        const int64_t ijk[3] = {c % p.n[0], (c / p.n[0]) % p.n[1], c / ((int64_t)p.n[0] * p.n[1])};
        for (int d = 0; d < D; ++d) {
            if (ijk[d] + 1 >= p.n[d]) continue;
            const int64_t cn = c + (d == 0 ? 1 : d == 1 ? p.n[0] : (int64_t)p.n[0] * p.n[1]);
            auto it = cell_elems.find(cn);
            if (it == cell_elems.end()) continue;
            for (size_t a : kv.second) {
                const auto fa = face_facets(p, a, d, 1.0);
                for (size_t b : it->second) {
                    const auto fb = face_facets(p, b, d, -1.0);
                    bool touch = false;
                    for (const auto& f : fa) {
                        std::array<double, 3> ctr{0.0, 0.0, 0.0};
                        for (int k = 0; k < D; ++k)
                            for (int q = 0; q < 3; ++q) ctr[q] += f[k][q] / D;
                        for (const auto& g : fb)
                            if (in_facet(D, d, ctr, g)) { touch = true; break; }
                        if (touch) break;
                    }
                    if (touch) {
                        nbr[a].push_back(b);
                        nbr[b].push_back(a);
                    }
                }
            }
        }
    }
    // support (elements) of each adof, ascending
    const int nb = p.nb_ip;
    std::vector<std::vector<size_t>> support(p.N);
    for (size_t e = 0; e < n_el; ++e)
        for (int32_t q = p.t_ptr[e * nb]; q < p.t_ptr[(e + 1) * nb]; ++q) {
            auto& s = support[p.t_adof[q]];
            if (s.empty() || s.back() != e) s.push_back(e);
        }
    // levels per (adof, element) and the enriched ids
    std::vector<std::map<size_t, int>> level(p.N);
    std::vector<int32_t> base(p.N, 0);
    int32_t next = 0;
    for (int32_t a = 0; a < p.N; ++a) {
        const auto& s = support[a];
        int n_lev = 0;
        for (size_t e : s) {
            if (level[a].count(e)) continue;
            // flood the component of e within the support
            std::vector<size_t> stack{e};
            level[a][e] = n_lev;
            while (!stack.empty()) {
                const size_t x = stack.back();
                stack.pop_back();
                for (size_t y : nbr[x]) {
                    if (!std::binary_search(s.begin(), s.end(), y) || level[a].count(y)) continue;
                    level[a][y] = n_lev;
                    stack.push_back(y);
                }
            }
            ++n_lev;
        }
        base[a] = next;
        next += n_lev;
    }
    for (size_t e = 0; e < n_el; ++e)
        for (int32_t q = p.t_ptr[e * nb]; q < p.t_ptr[(e + 1) * nb]; ++q) {
            const int32_t a = p.t_adof[q];
            p.t_adof[q] = base[a] + level[a].at(e);
        }
    p.N = next;
}

void ghost_facets(const OracleProblem& p, std::vector<int32_t>& out);

void build(OracleProblem& p, const mgpu_problem_spec& s) {
    p.spec = s;
    p.dim = s.dim;
    p.nb = s.dim == 3 ? 8 : 4;
    const int lag = s.lagrange_order <= 1 ? 1 : s.lagrange_order;
    if (lag > 2 || (lag == 2 && s.dim != 3)) throw std::runtime_error("lagrange_order 2 (HEX27) is 3D only");
    p.nb_ip = lag == 2 ? 27 : p.nb;
    p.nf = s.iwg == MGPU_IWG_STRUC_LINEAR_BULK ? s.dim : 1;
    p.nm = s.iwg == MGPU_IWG_STRUC_LINEAR_BULK ? 2 : 1;
    for (int d = 0; d < 3; ++d) {
        p.n[d] = d < s.dim ? s.n[d] : 1;
        p.lo[d] = d < s.dim ? s.lo[d] : 0.0;
        p.h[d] = d < s.dim ? (s.hi[d] - s.lo[d]) / (double)s.n[d] : 1.0;
    }
    p.ls.kind = s.ls_kind;
    p.ls.dim = s.dim;
    if (s.ls_kind == MGPU_LS_SPHERES) p.ls.spheres.assign(s.spheres, s.spheres + 4 * s.n_spheres);
    p.spec.spheres = nullptr;
    p.ls.gs = s.gyroid_scale;
    for (int d = 0; d < 4; ++d) p.ls.plane[d] = s.plane[d];
    p.n_cells = (int64_t)p.n[0] * p.n[1] * p.n[2];
    p.n_nodes = s.dim == 3 ? (int64_t)(p.n[0] + 1) * (p.n[1] + 1) * (p.n[2] + 1) : (int64_t)(p.n[0] + 1) * (p.n[1] + 1);
    p.node_phi.resize(p.n_nodes);
    for (int64_t nid = 0; nid < p.n_nodes; ++nid) {
        double x[3];
        node_x(p, nid, x);
        p.node_phi[nid] = p.ls.eval(x);
    }
    std::map<int64_t, std::vector<double>> cut;
    decompose(p, cut);
    const int nv = p.dim + 1;
    // THB: the intersected level-0 cells are refined; their level-1 cells are cut on the level-1
    // lattice (all of whose intersected cells are decomposed, as for level 0)
    if (s.refine_levels < 0 || s.refine_levels > 1) throw std::runtime_error("refine_levels must be 0 or 1");
    const bool thb = s.refine_levels == 1;
    if (thb && (s.dim != 3 || s.bspline_order != 2)) throw std::runtime_error("THB needs 3D and bspline_order 2");
    OracleProblem fine;
    std::map<int64_t, std::vector<double>> cut1;
    if (thb) {
        p.refined.assign(p.n_cells, 0);
        for (int64_t c = 0; c < p.n_cells; ++c) p.refined[c] = intersected(p, c);
        fine.spec = p.spec;
        fine.ls = p.ls;
        fine.dim = 3;
        fine.nb = p.nb;
        fine.nb_ip = p.nb_ip;
        for (int d = 0; d < 3; ++d) {
            fine.n[d] = 2 * p.n[d];
            fine.lo[d] = p.lo[d];
            fine.h[d] = 0.5 * p.h[d];
        }
        fine.n_cells = (int64_t)fine.n[0] * fine.n[1] * fine.n[2];
        fine.n_nodes = (int64_t)(fine.n[0] + 1) * (fine.n[1] + 1) * (fine.n[2] + 1);
        fine.node_phi.resize(fine.n_nodes);
        for (int64_t nid = 0; nid < fine.n_nodes; ++nid) {
            double x[3];
            node_x(fine, nid, x);
            fine.node_phi[nid] = fine.ls.eval(x);
        }
        decompose(fine, cut1);
    }
    p.ig_ptr.push_back(0);
    std::vector<char> used(p.n_nodes, 0);
    // one element: level, its cell on that level's lattice, its level-0 cell, its sub-cells
    auto add_element = [&](int level, int64_t lc, int64_t c0, const std::vector<double>* sub) {
        const OracleProblem& L = level == 0 ? p : fine;
        p.elem_cell.push_back((int32_t)c0);
        p.elem_level.push_back((int8_t)level);
        p.elem_lcell.push_back(lc);
        const uint64_t sid = level == 0 ? (uint64_t)lc : (uint64_t)(p.n_cells + lc);
        if (s.iwg == MGPU_IWG_SPATIALDIFF_BULK) {
            p.material.push_back(s.k_sigma == 0.0 ? std::exp(s.k_mu) : std::exp(s.k_mu + s.k_sigma * gauss_draw(s.seed, sid)));
        } else {
            p.material.push_back(s.E);
            p.material.push_back(s.nu);
        }
        if (sub) {
            p.ig_xi.insert(p.ig_xi.end(), sub->begin(), sub->end());
            ++p.n_cut;
        }
        p.ig_ptr.push_back((int32_t)(p.ig_xi.size() / (nv * p.dim)));
        const int i = (int)(lc % L.n[0]);
        const int j = (int)((lc / L.n[0]) % L.n[1]);
        const int k = p.dim == 3 ? (int)(lc / ((int64_t)L.n[0] * L.n[1])) : 0;
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < p.dim; ++d) p.ip_box.push_back(L.lo[d] + (double)ijk[d] * L.h[d]);
        for (int d = 0; d < p.dim; ++d) p.ip_box.push_back(L.h[d]);
    };
    for (int64_t c = 0; c < p.n_cells; ++c) {
        if (thb && p.refined[c]) {
            const int64_t i = c % p.n[0], j = (c / p.n[0]) % p.n[1], k = c / ((int64_t)p.n[0] * p.n[1]);
            for (int cz = 0; cz < 2; ++cz)
                for (int cy = 0; cy < 2; ++cy)
                    for (int cx = 0; cx < 2; ++cx) {
                        const int64_t fc = (2 * i + cx) + (int64_t)fine.n[0] * ((2 * j + cy) + (int64_t)fine.n[1] * (2 * k + cz));
                        auto it = cut1.find(fc);
                        int64_t nodes[8];
                        cell_nodes(fine, fc, nodes);
                        if (it == cut1.end()) {
                            if (region(fine.node_phi[nodes[0]]) != s.material_phase) continue;
                            add_element(1, fc, c, nullptr);
                        } else if (!it->second.empty()) {
                            add_element(1, fc, c, &it->second);
                        }
                    }
            continue;
        }
        auto it = cut.find(c);
        int64_t nodes[8];
        cell_nodes(p, c, nodes);
        if (it == cut.end()) {
            if (region(p.node_phi[nodes[0]]) != s.material_phase) continue;
        } else if (it->second.empty()) {
            continue;
        }
        if (s.enrich && it != cut.end()) {
            // one element per subphase, its sub-cells in their original order
            const size_t stride = (size_t)nv * p.dim;
            for (const auto& g : subphase_groups(p.dim, it->second)) {
                std::vector<double> part;
                for (int q : g) part.insert(part.end(), it->second.begin() + q * stride, it->second.begin() + (q + 1) * stride);
                add_element(0, c, c, &part);
            }
        } else {
            add_element(0, c, c, it == cut.end() ? nullptr : &it->second);
        }
        for (int v = 0; v < p.nb; ++v) used[nodes[v]] = 1;
    }
    const int order = s.bspline_order <= 1 ? 1 : s.bspline_order;
    if (order == 1 && p.nb_ip != 27) {
        std::vector<int32_t> adof(p.n_nodes, -1);
        for (int64_t nid = 0; nid < p.n_nodes; ++nid)
            if (used[nid]) adof[nid] = p.N++;
        for (size_t e = 0; e < p.elem_cell.size(); ++e) {
            int64_t nodes[8];
            cell_nodes(p, p.elem_cell[e], nodes);
            for (int v = 0; v < p.nb; ++v) {
                p.t_ptr.push_back((int32_t)p.t_adof.size());
                p.t_adof.push_back(adof[nodes[v]]);
                p.t_weight.push_back(1.0);
            }
        }
    } else if (order == 1) {
        // HEX27 Lagrange nodes on the half-step grid: used nodes in ascending index, x fastest
        const int64_t gx = 2 * (int64_t)p.n[0] + 1, gy = 2 * (int64_t)p.n[1] + 1;
        auto key = [&](const double* x) {
            return std::llround(2.0 * x[0]) + gx * (std::llround(2.0 * x[1]) + gy * std::llround(2.0 * x[2]));
        };
        std::map<int64_t, int32_t> adof;
        double x[3];
        for (size_t e = 0; e < p.elem_cell.size(); ++e)
            for (int v = 0; v < 27; ++v) {
                basis_pos(p, e, v, x);
                adof[key(x)] = 0;
            }
        for (auto& kv : adof) kv.second = p.N++;
        for (size_t e = 0; e < p.elem_cell.size(); ++e)
            for (int v = 0; v < 27; ++v) {
                basis_pos(p, e, v, x);
                p.t_ptr.push_back((int32_t)p.t_adof.size());
                p.t_adof.push_back(adof.at(key(x)));
                p.t_weight.push_back(1.0);
            }
    } else if (thb) {
        thb_rows(p);
    } else {
        bspline2_rows(p);
    }
    p.t_ptr.push_back((int32_t)p.t_adof.size());
    if (s.enrich) {
        if (thb) throw std::runtime_error("enrichment is implemented for level-0 lattices");
        enrich_rows(p);
    }
    if (s.ghost) {
        if (thb || p.nb_ip == 27 || s.enrich) throw std::runtime_error("ghost facets need HEX8 / QUAD4 level-0 sets");
        std::vector<int32_t> g;
        ghost_facets(p, g);
        for (size_t k = 0; k < g.size(); k += 3) {
            p.couplings.push_back(g[k]);
            p.couplings.push_back(g[k + 1]);
        }
    }
}

// element topology: per type sorted unique adof ids (global)
void element_topology(const mgpu_set_desc& d, int nf, int64_t e, std::vector<int32_t>& ids) {
    std::vector<int32_t> u;
    for (int b = 0; b < d.n_bases; ++b)
        for (int q = d.t_ptr[e * d.n_bases + b]; q < d.t_ptr[e * d.n_bases + b + 1]; ++q) u.push_back(d.t_adof[q]);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    ids.clear();
    for (int f = 0; f < nf; ++f)
        for (int32_t a : u) ids.push_back(a + f * d.n_adofs_per_type);
}

void build_graph(const mgpu_set_desc& d, int nf, std::vector<int32_t>& row_ptr, std::vector<int32_t>& col_idx) {
    const int64_t nrows = (int64_t)nf * d.n_adofs_per_type;
    std::vector<std::vector<int32_t>> rows(nrows);
    std::vector<int32_t> ids;
    for (int64_t e = 0; e < d.n_elements; ++e) {
        element_topology(d, nf, e, ids);
        for (int32_t r : ids) rows[r].insert(rows[r].end(), ids.begin(), ids.end());
    }
    // double-sided clusters (ghost facets) couple their two elements
    std::vector<int32_t> other;
    for (int32_t k = 0; d.couplings && k < d.n_couplings; ++k) {
        element_topology(d, nf, d.couplings[2 * k], ids);
        element_topology(d, nf, d.couplings[2 * k + 1], other);
        ids.insert(ids.end(), other.begin(), other.end());
        for (int32_t r : ids) rows[r].insert(rows[r].end(), ids.begin(), ids.end());
    }
    row_ptr.assign(nrows + 1, 0);
    col_idx.clear();
    for (int64_t r = 0; r < nrows; ++r) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        col_idx.insert(col_idx.end(), v.begin(), v.end());
        row_ptr[r + 1] = (int32_t)col_idx.size();
    }
}

// ------------------------------------------------------------------------------------------------
// element K_e, R_e and scatter (reference algorithm, dense small matrices)
// ------------------------------------------------------------------------------------------------
struct Scratch {
    std::vector<double> Kp, Rp, up, T, Ke, Re;
    std::vector<int32_t> ids;
};

void element_assemble(const mgpu_set_desc& d, int64_t e, const double* u, const int32_t* row_ptr,
                      const int32_t* col_idx, double* vals, double* rhs, Scratch& s, int64_t* n_ip) {
    const int D = d.spatial_dim, nb = d.n_bases;
    const bool elast = d.iwg == MGPU_IWG_STRUC_LINEAR_BULK;
    const int nf = elast ? D : 1;
    const int np = nf * nb;
    const int32_t N = d.n_adofs_per_type;
    element_topology(d, nf, e, s.ids);
    const int na = (int)s.ids.size();
    const int na_t = na / nf;
    // dense T (np x na), block diagonal per dof type
    s.T.assign((size_t)np * na, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int q = d.t_ptr[e * nb + b]; q < d.t_ptr[e * nb + b + 1]; ++q) {
            const int col = (int)(std::lower_bound(s.ids.begin(), s.ids.begin() + na_t, d.t_adof[q]) - s.ids.begin());
            for (int f = 0; f < nf; ++f) s.T[(size_t)(f * nb + b) * na + f * na_t + col] += d.t_weight[q];
        }
    // pdof values u_p = T u_adof
    s.up.assign(np, 0.0);
    for (int i = 0; i < np; ++i)
        for (int a = 0; a < na; ++a) s.up[i] += s.T[(size_t)i * na + a] * u[s.ids[a]];
    s.Kp.assign((size_t)np * np, 0.0);
    s.Rp.assign(np, 0.0);
    // IP cell geometry: vertices of the box
    const double* box = d.ip_box + (size_t)e * 2 * D;
    double X[27][3];
    for (int v = 0; v < nb; ++v) {
        double xv[3];
        basis_xi(D, nb, v, xv);
        for (int q = 0; q < D; ++q) X[v][q] = box[q] + 0.5 * (xv[q] + 1.0) * box[D + q];
    }
    // rules: HEX_2x2x2 / QUAD_2x2 and TET_11 / TRI_7 under linear IP cells; HEX_3x3x3 and TET_35
    // under HEX27 (cl_FEM_Set.cpp:3022-3091)
    struct Rules {
        double h3[27][3], h3w[27], t35[35][3], t35w[35];
        Rules() { hex3_rule(h3, h3w); tet35_rule(t35, t35w); }
    };
    static const Rules rules;
    const bool quad = nb == 27;
    // material
    const int ns = elast ? (D == 3 ? 6 : 3) : D;
    double Dm[6][6] = {{0}};
    double kcond = 0.0;
    if (elast) {
        const double E = d.material[2 * e], nu = d.material[2 * e + 1];
        if (D == 3) {
            const double pre = E / (1.0 + nu) / (1.0 - 2.0 * nu);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Dm[i][j] = i == j ? pre * (1.0 - nu) : pre * nu;
            for (int i = 3; i < 6; ++i) Dm[i][i] = pre * (1.0 - 2.0 * nu) / 2.0;
        } else {
            const double pre = E / (1.0 - std::pow(nu, 2));
            Dm[0][0] = pre; Dm[1][1] = pre; Dm[0][1] = pre * nu; Dm[1][0] = pre * nu;
            Dm[2][2] = pre * 0.5 * (1.0 - nu);
        }
    } else {
        kcond = d.material[e];
    }
    // integration domains
    const int ig0 = d.ig_ptr[e], ig1 = d.ig_ptr[e + 1];
    const bool cut = ig1 > ig0;
    const int n_dom = cut ? ig1 - ig0 : 1;
    const int nv = D + 1;
    std::vector<double> B((size_t)ns * np), DB((size_t)ns * np);
    for (int dom = 0; dom < n_dom; ++dom) {
        int nq;
        if (!cut) nq = quad ? 27 : (D == 3 ? 8 : 4);
        else nq = quad ? 35 : (D == 3 ? 11 : 7);
        const double* Xh = cut ? d.ig_xi + (size_t)(ig0 + dom) * nv * D : nullptr;
        // IG physical vertex coordinates (interpolated in the IP cell)
        double Xig[4][3];
        if (cut)
            for (int v = 0; v < nv; ++v) {
                double Nv[27];
                basis_N(D, nb, Xh + v * D, Nv);
                for (int q = 0; q < D; ++q) {
                    Xig[v][q] = 0.0;
                    for (int a = 0; a < nb; ++a) Xig[v][q] += Nv[a] * X[a][q];
                }
            }
        for (int iq = 0; iq < nq; ++iq) {
            double xip[3] = {0, 0, 0};
            double w, detJ;
            if (!cut) {
                for (int q = 0; q < D; ++q) xip[q] = quad ? rules.h3[iq][q] : (D == 3 ? kHex2[iq][q] : kQuad2[iq][q]);
                w = quad ? rules.h3w[iq] : 1.0;
                // IG cell = IP cell: detJ of the hex / quad from physical coordinates
                double dN[3][27];
                basis_dN(D, nb, xip, dN);
                double Jt[3][3] = {{0}};
                for (int r = 0; r < D; ++r)
                    for (int q = 0; q < D; ++q)
                        for (int a = 0; a < nb; ++a) Jt[r][q] += dN[r][a] * X[a][q];
                detJ = D == 3 ? Jt[0][0] * (Jt[1][1] * Jt[2][2] - Jt[2][1] * Jt[1][2]) -
                                    Jt[0][1] * (Jt[1][0] * Jt[2][2] - Jt[1][2] * Jt[2][0]) +
                                    Jt[0][2] * (Jt[1][0] * Jt[2][1] - Jt[1][1] * Jt[2][0])
                              : Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0];
            } else {
                double xg[3];
                double Ng[4];
                if (D == 3) {
                    for (int q = 0; q < 3; ++q) xg[q] = quad ? rules.t35[iq][q] : kTet11P[iq][q];
                    Ng[0] = xg[0]; Ng[1] = xg[1]; Ng[2] = 1.0 - xg[0] - xg[1] - xg[2]; Ng[3] = xg[2];
                    w = quad ? rules.t35w[iq] : tet11_weight(iq);
                    // dN/dxi of TET4: rows (1,0,-1,0), (0,1,-1,0), (0,0,-1,1)
                    const double dNt[3][4] = {{1, 0, -1, 0}, {0, 1, -1, 0}, {0, 0, -1, 1}};
                    double Jt[3][3] = {{0}};
                    for (int r = 0; r < 3; ++r)
                        for (int q = 0; q < 3; ++q)
                            for (int v = 0; v < 4; ++v) Jt[r][q] += dNt[r][v] * Xig[v][q];
                    detJ = (Jt[0][0] * (Jt[1][1] * Jt[2][2] - Jt[2][1] * Jt[1][2]) -
                            Jt[0][1] * (Jt[1][0] * Jt[2][2] - Jt[1][2] * Jt[2][0]) +
                            Jt[0][2] * (Jt[1][0] * Jt[2][1] - Jt[1][1] * Jt[2][0])) / 6.0;
                } else {
                    for (int q = 0; q < 2; ++q) xg[q] = kTri7P[iq][q];
                    Ng[0] = xg[0]; Ng[1] = xg[1]; Ng[2] = 1.0 - xg[0] - xg[1];
                    w = kTri7W[iq];
                    const double dNt[2][3] = {{1, 0, -1}, {0, 1, -1}};
                    double Jt[2][2] = {{0}};
                    for (int r = 0; r < 2; ++r)
                        for (int q = 0; q < 2; ++q)
                            for (int v = 0; v < 3; ++v) Jt[r][q] += dNt[r][v] * Xig[v][q];
                    detJ = (Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0]) / 2.0;
                }
                for (int q = 0; q < D; ++q) {
                    xip[q] = 0.0;
                    for (int v = 0; v < nv; ++v) xip[q] += Ng[v] * Xh[v * D + q];
                }
            }
            if (detJ < -1.0e-6) continue;
            const double ws = w * detJ;
            if (n_ip) ++*n_ip;
            // IP cell Jacobian at xi_IP and its rectangular (diagonal) inverse
            double dN[3][27];
            basis_dN(D, nb, xip, dN);
            double Jt[3][3] = {{0}};
            for (int r = 0; r < D; ++r)
                for (int q = 0; q < D; ++q)
                    for (int a = 0; a < nb; ++a) Jt[r][q] += dN[r][a] * X[a][q];
            double dNdx[3][27];
            for (int r = 0; r < D; ++r)
                for (int a = 0; a < nb; ++a) dNdx[r][a] = dN[r][a] / Jt[r][r];
            // B (test strain) and flux
            std::fill(B.begin(), B.end(), 0.0);
            if (elast) {
                if (D == 3) {
                    for (int a = 0; a < nb; ++a) {
                        B[0 * np + a] = dNdx[0][a];
                        B[4 * np + a] = dNdx[2][a];
                        B[5 * np + a] = dNdx[1][a];
                        B[1 * np + nb + a] = dNdx[1][a];
                        B[3 * np + nb + a] = dNdx[2][a];
                        B[5 * np + nb + a] = dNdx[0][a];
                        B[2 * np + 2 * nb + a] = dNdx[2][a];
                        B[3 * np + 2 * nb + a] = dNdx[1][a];
                        B[4 * np + 2 * nb + a] = dNdx[0][a];
                    }
                } else {
                    for (int a = 0; a < nb; ++a) {
                        B[0 * np + a] = dNdx[0][a];
                        B[2 * np + a] = dNdx[1][a];
                        B[1 * np + nb + a] = dNdx[1][a];
                        B[2 * np + nb + a] = dNdx[0][a];
                    }
                }
                // D B
                for (int i = 0; i < ns; ++i)
                    for (int c = 0; c < np; ++c) {
                        double v = 0.0;
                        for (int k = 0; k < ns; ++k) v += Dm[i][k] * B[k * np + c];
                        DB[i * np + c] = v;
                    }
            } else {
                for (int i = 0; i < D; ++i)
                    for (int a = 0; a < nb; ++a) {
                        B[i * np + a] = dNdx[i][a];
                        DB[i * np + a] = kcond * dNdx[i][a];
                    }
            }
            // flux = D B u_p
            double flux[6] = {0, 0, 0, 0, 0, 0};
            for (int i = 0; i < ns; ++i)
                for (int c = 0; c < np; ++c) flux[i] += DB[i * np + c] * s.up[c];
            for (int r = 0; r < np; ++r) {
                double rv = 0.0;
                for (int i = 0; i < ns; ++i) rv += B[i * np + r] * flux[i];
                s.Rp[r] += ws * rv;
                for (int c = 0; c < np; ++c) {
                    double kv = 0.0;
                    for (int i = 0; i < ns; ++i) kv += B[i * np + r] * DB[i * np + c];
                    s.Kp[(size_t)r * np + c] += ws * kv;
                }
            }
        }
    }
    // K_e = T^T K_p T, R_e = T^T R_p
    s.Ke.assign((size_t)na * na, 0.0);
    s.Re.assign(na, 0.0);
    std::vector<double> KT((size_t)np * na, 0.0);
    for (int i = 0; i < np; ++i)
        for (int a = 0; a < na; ++a) {
            double v = 0.0;
            for (int j = 0; j < np; ++j) v += s.Kp[(size_t)i * np + j] * s.T[(size_t)j * na + a];
            KT[(size_t)i * na + a] = v;
        }
    for (int a = 0; a < na; ++a) {
        for (int b = 0; b < na; ++b) {
            double v = 0.0;
            for (int i = 0; i < np; ++i) v += s.T[(size_t)i * na + a] * KT[(size_t)i * na + b];
            s.Ke[(size_t)a * na + b] = v;
        }
        double r = 0.0;
        for (int i = 0; i < np; ++i) r += s.T[(size_t)i * na + a] * s.Rp[i];
        s.Re[a] = r;
    }
    (void)N;
    // scatter: K[ids[a], ids[b]] += K_e[a][b] (row search), R[ids[a]] += R_e[a]
    for (int a = 0; a < na; ++a) {
        const int32_t row = s.ids[a];
        const int32_t* c0 = col_idx + row_ptr[row];
        const int32_t* c1 = col_idx + row_ptr[row + 1];
        for (int b = 0; b < na; ++b) {
            const int32_t* it = std::lower_bound(c0, c1, s.ids[b]);
            if (it == c1 || *it != s.ids[b]) throw std::runtime_error("column not in graph");
            vals[it - col_idx] += s.Ke[(size_t)a * na + b];
        }
        rhs[row] += s.Re[a];
    }
}

// ------------------------------------------------------------------------------------------------
// side sets (Element_Sideset::compute_jacobian_and_residual, cl_FEM_Element_Sideset.cpp:245-330,
// with IWG_*_Dirichlet(_Nitsche) / IWG_*_Neumann, cl_FEM_IWG_Isotropic_Struc_Linear_Dirichlet.cpp
// :52-240, cl_FEM_IWG_Diffusion_Dirichlet_Nitsche.cpp, cl_FEM_IWG_Isotropic_Struc_Linear_Neumann.cpp
// :38-85; SP_Dirichlet_Nitsche, cl_FEM_SP_Dirichlet_Nitsche.cpp:75-115; side cluster length
// measure, cl_FEM_Cluster.cpp:858-919; flatten_normal, cl_FEM_CM_Struc_Linear.cpp:836-861)
// ------------------------------------------------------------------------------------------------
const double kBar3P[3] = {-0.7745966692414834, 0.000000000000000, 0.7745966692414834};
const double kBar3Wt[3] = {5.555555555555556e-01, 8.888888888888888e-01, 5.555555555555556e-01};

// side rule points: shape values of the facet vertices (phi[q][k]), weights, and the physical
// Jacobian measure at each point
int side_points(int D, int nvs, const double (*Xf)[3], double (*phi)[4], double* w, double* detJ) {
    auto cross_norm = [](const double* a, const double* b) {
        const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
        return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    };
    if (nvs == 3) {   // TRI_7 on a tet face: space Jacobian rows x1 - x3, x2 - x3, detJ = |cross| / 2
        for (int q = 0; q < 7; ++q) {
            phi[q][0] = kTri7P[q][0];
            phi[q][1] = kTri7P[q][1];
            phi[q][2] = 1.0 - kTri7P[q][0] - kTri7P[q][1];
            w[q] = kTri7W[q];
            double a[3], b[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = Xf[0][k] - Xf[2][k];
                b[k] = Xf[1][k] - Xf[2][k];
            }
            detJ[q] = cross_norm(a, b) / 2.0;
        }
        return 7;
    }
    if (nvs == 4) {   // QUAD_2x2 on a hex face: bilinear map, detJ = |dX/ds x dX/dt|
        for (int q = 0; q < 4; ++q) {
            const double s = kQuad2[q][0], t = kQuad2[q][1];
            const double Nq[4] = {0.25 * (1 - s) * (1 - t), 0.25 * (1 + s) * (1 - t), 0.25 * (1 + s) * (1 + t),
                                  0.25 * (1 - s) * (1 + t)};
            const double ds[4] = {-0.25 * (1 - t), 0.25 * (1 - t), 0.25 * (1 + t), -0.25 * (1 + t)};
            const double dt[4] = {-0.25 * (1 - s), -0.25 * (1 + s), 0.25 * (1 + s), 0.25 * (1 - s)};
            double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k) {
                phi[q][k] = Nq[k];
                for (int c = 0; c < 3; ++c) {
                    a[c] += ds[k] * Xf[k][c];
                    b[c] += dt[k] * Xf[k][c];
                }
            }
            w[q] = 1.0;
            detJ[q] = cross_norm(a, b);
        }
        return 4;
    }
    for (int q = 0; q < 3; ++q) {   // BAR_3 on a 2D side: detJ = |dX/ds|
        phi[q][0] = 0.5 * (1.0 - kBar3P[q]);
        phi[q][1] = 0.5 * (1.0 + kBar3P[q]);
        w[q] = kBar3Wt[q];
        double l2 = 0.0;
        for (int c = 0; c < D; ++c) l2 += 0.25 * (Xf[1][c] - Xf[0][c]) * (Xf[1][c] - Xf[0][c]);
        detJ[q] = std::sqrt(l2);
    }
    return 3;
}

// Side set on descriptor d: facets (element, vertices in IP parametric coordinates, normal),
// added into vals / rhs with dense N, B, D, n_flat and T per facet.
void side_assemble(const mgpu_set_desc& d, int iwg, int nvs, int64_t nfc, const int32_t* fe, const double* fx,
                   const double* fnrm, double alpha, const double* value, const double* u, const int32_t* row_ptr,
                   const int32_t* col_idx, double* vals, double* rhs) {
    const int D = d.spatial_dim, nb = d.n_bases;
    const bool elast = d.iwg == MGPU_IWG_STRUC_LINEAR_BULK;
    const int nf = elast ? D : 1, np = nf * nb, ns = elast ? (D == 3 ? 6 : 3) : D;
    const bool neumann = iwg == MGPU_IWG_SPATIALDIFF_NEUMANN || iwg == MGPU_IWG_STRUC_LINEAR_NEUMANN;
    const double beta = (iwg == MGPU_IWG_SPATIALDIFF_DIRICHLET_SYMMETRIC_NITSCHE || iwg == MGPU_IWG_STRUC_LINEAR_DIRICHLET_SYMMETRIC_NITSCHE) ? -1.0 : 1.0;
    // cluster measure per element (sum of the side measures of its facets)
    std::map<int32_t, double> cluster;
    auto facet_phys = [&](int64_t k, double (*Xf)[3]) {
        const int32_t e = fe[k];
        const double* box = d.ip_box + (size_t)e * 2 * D;
        for (int v = 0; v < nvs; ++v)
            for (int c = 0; c < 3; ++c)
                Xf[v][c] = c < D ? box[c] + 0.5 * (fx[((size_t)k * nvs + v) * D + c] + 1.0) * box[D + c] : 0.0;
    };
    for (int64_t k = 0; k < nfc; ++k) {
        double Xf[4][3], phi[7][4], w[7], dJ[7];
        facet_phys(k, Xf);
        const int nq = side_points(D, nvs, Xf, phi, w, dJ);
        double m = 0.0;
        for (int q = 0; q < nq; ++q) m += w[q] * dJ[q];
        cluster[fe[k]] += m;
    }
    Scratch s;
    for (int64_t k = 0; k < nfc; ++k) {
        const int32_t e = fe[k];
        element_topology(d, nf, e, s.ids);
        const int na = (int)s.ids.size(), na_t = na / nf;
        s.T.assign((size_t)np * na, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int q = d.t_ptr[e * nb + b]; q < d.t_ptr[e * nb + b + 1]; ++q) {
                const int col = (int)(std::lower_bound(s.ids.begin(), s.ids.begin() + na_t, d.t_adof[q]) - s.ids.begin());
                for (int f = 0; f < nf; ++f) s.T[(size_t)(f * nb + b) * na + f * na_t + col] += d.t_weight[q];
            }
        s.up.assign(np, 0.0);
        for (int i = 0; i < np; ++i)
            for (int a = 0; a < na; ++a) s.up[i] += s.T[(size_t)i * na + a] * u[s.ids[a]];
        s.Kp.assign((size_t)np * np, 0.0);
        s.Rp.assign(np, 0.0);
        // material and the Nitsche parameter of the side cluster
        double Dm[6][6] = {{0}}, kcond = 0.0, Emod = 0.0;
        if (elast) {
            const double E = d.material[2 * e], nu = d.material[2 * e + 1];
            Emod = E;
            if (D == 3) {
                const double pre = E / (1.0 + nu) / (1.0 - 2.0 * nu);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Dm[i][j] = i == j ? pre * (1.0 - nu) : pre * nu;
                for (int i = 3; i < 6; ++i) Dm[i][i] = pre * (1.0 - 2.0 * nu) / 2.0;
            } else {
                const double pre = E / (1.0 - nu * nu);
                Dm[0][0] = pre; Dm[1][1] = pre; Dm[0][1] = pre * nu; Dm[1][0] = pre * nu;
                Dm[2][2] = pre * 0.5 * (1.0 - nu);
            }
        } else {
            kcond = d.material[e];
        }
        const double A = cluster[e];
        const double hlen = D == 3 ? std::pow(4.0 * A / 3.141592653589793, 0.5) : A;
        const double gamma = alpha * (elast ? Emod : kcond) / hlen;
        // flattened normal (n_f x n_s): traction = n_flat sigma
        const double* nv = fnrm + (size_t)k * D;
        double nfl[3][6] = {{0}};
        if (elast && D == 3) {
            nfl[0][0] = nv[0]; nfl[1][1] = nv[1]; nfl[2][2] = nv[2];
            nfl[0][4] = nv[2]; nfl[0][5] = nv[1]; nfl[1][3] = nv[2];
            nfl[1][5] = nv[0]; nfl[2][3] = nv[1]; nfl[2][4] = nv[0];
        } else if (elast) {
            nfl[0][0] = nv[0]; nfl[0][2] = nv[1]; nfl[1][1] = nv[1]; nfl[1][2] = nv[0];
        } else {
            for (int c = 0; c < D; ++c) nfl[0][c] = nv[c];
        }
        // IP cell vertices
        const double* box = d.ip_box + (size_t)e * 2 * D;
        double X[27][3];
        for (int v = 0; v < nb; ++v) {
            double xv[3];
            basis_xi(D, nb, v, xv);
            for (int c = 0; c < D; ++c) X[v][c] = box[c] + 0.5 * (xv[c] + 1.0) * box[D + c];
        }
        double Xf[4][3], phi[7][4], w[7], dJ[7];
        facet_phys(k, Xf);
        const int nq = side_points(D, nvs, Xf, phi, w, dJ);
        std::vector<double> Nm((size_t)nf * np), B((size_t)ns * np), Tt((size_t)nf * np);
        for (int q = 0; q < nq; ++q) {
            if (dJ[q] < -1.0e-6) continue;
            const double ws = w[q] * dJ[q];
            // point in the IP cell (parametric), N and dN/dx there
            double xip[3] = {0, 0, 0};
            for (int c = 0; c < D; ++c)
                for (int v = 0; v < nvs; ++v) xip[c] += phi[q][v] * fx[((size_t)k * nvs + v) * D + c];
            double Nb[27], dN[3][27];
            basis_N(D, nb, xip, Nb);
            basis_dN(D, nb, xip, dN);
            double Jt[3][3] = {{0}};
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c)
                    for (int a = 0; a < nb; ++a) Jt[r][c] += dN[r][a] * X[a][c];
            double dNdx[3][27];
            for (int r = 0; r < D; ++r)
                for (int a = 0; a < nb; ++a) dNdx[r][a] = dN[r][a] / Jt[r][r];
            std::fill(Nm.begin(), Nm.end(), 0.0);
            std::fill(B.begin(), B.end(), 0.0);
            for (int f = 0; f < nf; ++f)
                for (int a = 0; a < nb; ++a) Nm[(size_t)f * np + f * nb + a] = Nb[a];
            if (elast && D == 3) {
                for (int a = 0; a < nb; ++a) {
                    B[0 * np + a] = dNdx[0][a]; B[4 * np + a] = dNdx[2][a]; B[5 * np + a] = dNdx[1][a];
                    B[1 * np + nb + a] = dNdx[1][a]; B[3 * np + nb + a] = dNdx[2][a]; B[5 * np + nb + a] = dNdx[0][a];
                    B[2 * np + 2 * nb + a] = dNdx[2][a]; B[3 * np + 2 * nb + a] = dNdx[1][a];
                    B[4 * np + 2 * nb + a] = dNdx[0][a];
                }
            } else if (elast) {
                for (int a = 0; a < nb; ++a) {
                    B[0 * np + a] = dNdx[0][a]; B[2 * np + a] = dNdx[1][a];
                    B[1 * np + nb + a] = dNdx[1][a]; B[2 * np + nb + a] = dNdx[0][a];
                }
            } else {
                for (int i = 0; i < D; ++i)
                    for (int a = 0; a < nb; ++a) B[(size_t)i * np + a] = dNdx[i][a];
            }
            // T = n_flat D B (diffusion: k n^T B)
            for (int f = 0; f < nf; ++f)
                for (int c = 0; c < np; ++c) {
                    double v = 0.0;
                    for (int i = 0; i < ns; ++i) {
                        double db = 0.0;
                        if (elast)
                            for (int j = 0; j < ns; ++j) db += Dm[i][j] * B[(size_t)j * np + c];
                        else
                            db = kcond * B[(size_t)i * np + c];
                        v += nfl[f][i] * db;
                    }
                    Tt[(size_t)f * np + c] = v;
                }
            // jump u - u_D and traction at the point
            double uh[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
            for (int f = 0; f < nf; ++f)
                for (int c = 0; c < np; ++c) {
                    uh[f] += Nm[(size_t)f * np + c] * s.up[c];
                    tr[f] += Tt[(size_t)f * np + c] * s.up[c];
                }
            for (int r = 0; r < np; ++r) {
                if (neumann) {
                    double v = 0.0;
                    for (int f = 0; f < nf; ++f) v += Nm[(size_t)f * np + r] * value[f];
                    s.Rp[r] -= ws * v;
                    continue;
                }
                double rv = 0.0;
                for (int f = 0; f < nf; ++f) {
                    const double jump = uh[f] - value[f];
                    rv += -Nm[(size_t)f * np + r] * tr[f] + beta * Tt[(size_t)f * np + r] * jump +
                          gamma * Nm[(size_t)f * np + r] * jump;
                }
                s.Rp[r] += ws * rv;
                for (int c = 0; c < np; ++c) {
                    double kv = 0.0;
                    for (int f = 0; f < nf; ++f)
                        kv += -Nm[(size_t)f * np + r] * Tt[(size_t)f * np + c] + beta * Tt[(size_t)f * np + r] * Nm[(size_t)f * np + c] +
                              gamma * Nm[(size_t)f * np + r] * Nm[(size_t)f * np + c];
                    s.Kp[(size_t)r * np + c] += ws * kv;
                }
            }
        }
        // K_e = T^T K_p T, R_e = T^T R_p, scatter by row search
        std::vector<double> KT((size_t)np * na, 0.0);
        for (int i = 0; i < np; ++i)
            for (int a = 0; a < na; ++a) {
                double v = 0.0;
                for (int j = 0; j < np; ++j) v += s.Kp[(size_t)i * np + j] * s.T[(size_t)j * na + a];
                KT[(size_t)i * na + a] = v;
            }
        for (int a = 0; a < na; ++a) {
            const int32_t row = s.ids[a];
            const int32_t* c0 = col_idx + row_ptr[row];
            const int32_t* c1 = col_idx + row_ptr[row + 1];
            for (int b = 0; b < na; ++b) {
                double v = 0.0;
                for (int i = 0; i < np; ++i) v += s.T[(size_t)i * na + a] * KT[(size_t)i * na + b];
                const int32_t* it = std::lower_bound(c0, c1, s.ids[b]);
                if (it == c1 || *it != s.ids[b]) throw std::runtime_error("column not in graph");
                vals[it - col_idx] += v;
            }
            double r = 0.0;
            for (int i = 0; i < np; ++i) r += s.T[(size_t)i * na + a] * s.Rp[i];
            rhs[row] += r;
        }
    }
}

// Facets of the oracle's own problem: domain boundary faces of material cells and interface faces
// of material sub-cells (on no cell face, shared with no other material sub-cell of the cell).
void problem_facets(const OracleProblem& p, int where, int nvs, std::vector<int32_t>& fe, std::vector<double>& fx,
                    std::vector<double>& fn) {
    const int D = p.dim, nv = D + 1;
    const int bd = where > 0 ? (where - 1) / 2 : -1;
    const double bs = where > 0 ? ((where - 1) % 2 ? 1.0 : -1.0) : 0.0;
    const bool uncut_faces = nvs == 4 || (D == 2 && where > 0);
    const bool cut_faces = nvs != 4;
    auto on_face = [&](const std::vector<const double*>& vs, int dd, double sg) {
        for (const double* v : vs)
            if (std::fabs(v[dd] - sg) > 1e-12) return false;
        return true;
    };
    for (size_t e = 0; e < p.elem_cell.size(); ++e) {
        const int64_t c = p.elem_cell[e];
        const int64_t ijk[3] = {c % p.n[0], (c / p.n[0]) % p.n[1], D == 3 ? c / ((int64_t)p.n[0] * p.n[1]) : 0};
        const bool at_b = where > 0 && (bs < 0 ? ijk[bd] == 0 : ijk[bd] == p.n[bd] - 1);
        const double* box = p.ip_box.data() + e * 2 * D;
        const int s0 = p.ig_ptr[e], s1 = p.ig_ptr[e + 1];
        if (s0 == s1) {
            if (!uncut_faces || !at_b) continue;
            int o[2], m = 0;
            for (int dd = 0; dd < D; ++dd)
                if (dd != bd) o[m++] = dd;
            const double cyc[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
            fe.push_back((int32_t)e);
            for (int k = 0; k < (D == 3 ? 4 : 2); ++k) {
                double v[3] = {0, 0, 0};
                v[bd] = bs;
                if (D == 3) { v[o[0]] = cyc[k][0]; v[o[1]] = cyc[k][1]; }
                else v[o[0]] = k == 0 ? -1.0 : 1.0;
                for (int dd = 0; dd < D; ++dd) fx.push_back(v[dd]);
            }
            for (int dd = 0; dd < D; ++dd) fn.push_back(dd == bd ? bs : 0.0);
            continue;
        }
        if (!cut_faces || (where > 0 && !at_b)) continue;
        const double* X0 = p.ig_xi.data() + (size_t)s0 * nv * D;
        const int nsub = s1 - s0;
        // sorted vertex list of face (sub-cell s, omitted vertex k)
        auto face_sorted = [&](int sc, int k) {
            std::vector<std::vector<double>> vs;
            for (int q = 0; q < nv; ++q)
                if (q != k) vs.push_back(std::vector<double>(X0 + (sc * nv + q) * D, X0 + (sc * nv + q) * D + D));
            std::sort(vs.begin(), vs.end());
            return vs;
        };
        for (int sc = 0; sc < nsub; ++sc)
            for (int k = 0; k < nv; ++k) {
                std::vector<const double*> vs;
                for (int q = 0; q < nv; ++q)
                    if (q != k) vs.push_back(X0 + (sc * nv + q) * D);
                double nrm[3] = {0, 0, 0};
                if (where > 0) {
                    if (!on_face(vs, bd, bs)) continue;
                    nrm[bd] = bs;
                } else {
                    bool cf = false;
                    for (int dd = 0; dd < D; ++dd) cf = cf || on_face(vs, dd, -1.0) || on_face(vs, dd, 1.0);
                    if (cf) continue;
                    const auto me = face_sorted(sc, k);
                    bool shared = false;
                    for (int s2 = 0; s2 < nsub && !shared; ++s2)
                        for (int k2 = 0; k2 < nv && !shared; ++k2)
                            if (!(s2 == sc && k2 == k) && face_sorted(s2, k2) == me) shared = true;
                    if (shared) continue;
                    // outward normal of the sub-cell (physical coordinates)
                    double x[3][3] = {{0}}, xo[3] = {0, 0, 0};
                    for (int a = 0; a < (int)vs.size(); ++a)
                        for (int dd = 0; dd < D; ++dd) x[a][dd] = box[dd] + 0.5 * (vs[a][dd] + 1.0) * box[D + dd];
                    for (int dd = 0; dd < D; ++dd) xo[dd] = box[dd] + 0.5 * (X0[(sc * nv + k) * D + dd] + 1.0) * box[D + dd];
                    if (D == 3) {
                        const double a1[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
                        const double a2[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
                        nrm[0] = a1[1] * a2[2] - a1[2] * a2[1];
                        nrm[1] = a1[2] * a2[0] - a1[0] * a2[2];
                        nrm[2] = a1[0] * a2[1] - a1[1] * a2[0];
                    } else {
                        nrm[0] = x[1][1] - x[0][1];
                        nrm[1] = x[0][0] - x[1][0];
                    }
                    double dot = 0.0, len = 0.0;
                    for (int dd = 0; dd < D; ++dd) {
                        dot += nrm[dd] * (xo[dd] - x[0][dd]);
                        len += nrm[dd] * nrm[dd];
                    }
                    len = std::sqrt(len);
                    if (!(len > 0.0)) continue;
                    for (int dd = 0; dd < D; ++dd) nrm[dd] = (dot > 0.0 ? -nrm[dd] : nrm[dd]) / len;
                }
                fe.push_back((int32_t)e);
                for (const double* v : vs)
                    for (int dd = 0; dd < D; ++dd) fx.push_back(v[dd]);
                for (int dd = 0; dd < D; ++dd) fn.push_back(nrm[dd]);
            }
    }
}

thread_local std::string g_err;

// Ghost facets (XTK ghost stabilisation): every pair of face-neighbouring material elements of the
// level-0 lattice, at least one cut, whose shared face has a corner node strictly in the material;
// leader the element of the lower cell; (leader, follower, axis) triples in leader order.
void ghost_facets(const OracleProblem& p, std::vector<int32_t>& out) {
    out.clear();
    std::map<int64_t, size_t> elem_of;
    for (size_t e = 0; e < p.elem_cell.size(); ++e) elem_of[p.elem_cell[e]] = e;
    for (size_t e = 0; e < p.elem_cell.size(); ++e) {
        const int64_t c = p.elem_cell[e];
        int64_t nodes[8];
        cell_nodes(p, c, nodes);
        const int64_t ijk[3] = {c % p.n[0], (c / p.n[0]) % p.n[1], c / ((int64_t)p.n[0] * p.n[1])};
        for (int d = 0; d < p.dim; ++d) {
            if (ijk[d] + 1 >= p.n[d]) continue;
            const int64_t cn = c + (d == 0 ? 1 : d == 1 ? p.n[0] : (int64_t)p.n[0] * p.n[1]);
            auto it = elem_of.find(cn);
            if (it == elem_of.end()) continue;
            const size_t en = it->second;
            const bool cut = p.ig_ptr[e + 1] > p.ig_ptr[e] || p.ig_ptr[en + 1] > p.ig_ptr[en];
            if (!cut) continue;
            // the face x_d = upper of cell c: its corner nodes are those of the neighbour's lower face
            int64_t nn[8];
            cell_nodes(p, cn, nn);
            bool material = false;
            for (int v = 0; v < p.nb && !material; ++v)
                for (int w = 0; w < p.nb && !material; ++w)
                    if (nodes[v] == nn[w]) material = region(p.node_phi[nodes[v]]) == p.spec.material_phase;
            if (!material) continue;
            out.push_back((int32_t)e);
            out.push_back((int32_t)en);
            out.push_back(d);
        }
    }
}

// Ghost facets on descriptor d (IWG_Diffusion_Virtual_Work_Ghost / IWG_Isotropic_Struc_Linear_
// Virtual_Work_Ghost, order 1, SP_Ghost_Virtual_Work gamma = alpha h): per facet the leader's and
// the follower's B at each face point (the follower's parametric point from the physical one),
//   K_XY += w* gamma s_X s_Y (n_flat B_X)^T (n_flat D_Y B_Y),  s_L = 1, s_F = -1
// (diffusion: (n^T B_X)^T k_Y n^T B_Y), then T^T K T over the union of the two topologies, R = K u.
void ghost_assemble(const mgpu_set_desc& d, int64_t nfc, const int32_t* fl, const int32_t* ff, const double* fx,
                    const double* fnrm, double alpha, const double* u, const int32_t* row_ptr,
                    const int32_t* col_idx, double* vals, double* rhs) {
    const int D = d.spatial_dim, nb = d.n_bases;
    const int nvs = D == 3 ? 4 : 2;
    const bool elast = d.iwg == MGPU_IWG_STRUC_LINEAR_BULK;
    const int nf = elast ? D : 1, np = nf * nb, ns = elast ? (D == 3 ? 6 : 3) : D;
    for (int64_t k = 0; k < nfc; ++k) {
        const int32_t eX[2] = {fl[k], ff[k]};
        std::vector<int32_t> idL, idF, U;
        element_topology(d, nf, eX[0], idL);
        element_topology(d, nf, eX[1], idF);
        std::set_union(idL.begin(), idL.end(), idF.begin(), idF.end(), std::back_inserter(U));
        const int nu = (int)U.size();
        // T of both sides against U: rows (X, f, b)
        std::vector<double> T((size_t)2 * np * nu, 0.0);
        for (int X = 0; X < 2; ++X)
            for (int b = 0; b < nb; ++b)
                for (int q = d.t_ptr[eX[X] * nb + b]; q < d.t_ptr[eX[X] * nb + b + 1]; ++q)
                    for (int f = 0; f < nf; ++f) {
                        const int32_t id = d.t_adof[q] + f * d.n_adofs_per_type;
                        const int col = (int)(std::lower_bound(U.begin(), U.end(), id) - U.begin());
                        T[(size_t)(X * np + f * nb + b) * nu + col] += d.t_weight[q];
                    }
        // materials
        double Dm[2][6][6] = {}, kc[2] = {0.0, 0.0};
        for (int X = 0; X < 2; ++X) {
            const int32_t e = eX[X];
            if (elast) {
                const double E = d.material[2 * e], nu_ = d.material[2 * e + 1];
                if (D == 3) {
                    const double pre = E / (1.0 + nu_) / (1.0 - 2.0 * nu_);
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) Dm[X][i][j] = i == j ? pre * (1.0 - nu_) : pre * nu_;
                    for (int i = 3; i < 6; ++i) Dm[X][i][i] = pre * (1.0 - 2.0 * nu_) / 2.0;
                } else {
                    const double pre = E / (1.0 - nu_ * nu_);
                    Dm[X][0][0] = pre; Dm[X][1][1] = pre; Dm[X][0][1] = pre * nu_; Dm[X][1][0] = pre * nu_;
                    Dm[X][2][2] = pre * 0.5 * (1.0 - nu_);
                }
            } else {
                kc[X] = d.material[e];
            }
        }
        const double* nv = fnrm + (size_t)k * D;
        double nfl[3][6] = {{0}};
        if (elast && D == 3) {
            nfl[0][0] = nv[0]; nfl[1][1] = nv[1]; nfl[2][2] = nv[2];
            nfl[0][4] = nv[2]; nfl[0][5] = nv[1]; nfl[1][3] = nv[2];
            nfl[1][5] = nv[0]; nfl[2][3] = nv[1]; nfl[2][4] = nv[0];
        } else if (elast) {
            nfl[0][0] = nv[0]; nfl[0][2] = nv[1]; nfl[1][1] = nv[1]; nfl[1][2] = nv[0];
        } else {
            for (int c = 0; c < D; ++c) nfl[0][c] = nv[c];
        }
        // face points and the cluster's length measure
        const double* boxL = d.ip_box + (size_t)eX[0] * 2 * D;
        double Xf[4][3], phi[7][4], w[7], dJ[7];
        for (int v = 0; v < nvs; ++v)
            for (int c = 0; c < 3; ++c)
                Xf[v][c] = c < D ? boxL[c] + 0.5 * (fx[((size_t)k * nvs + v) * D + c] + 1.0) * boxL[D + c] : 0.0;
        const int nq = side_points(D, nvs, Xf, phi, w, dJ);
        double A = 0.0;
        for (int q = 0; q < nq; ++q) A += w[q] * dJ[q];
        const double gamma = alpha * (D == 3 ? std::pow(4.0 * A / 3.141592653589793, 0.5) : A);
        std::vector<double> Kp((size_t)4 * np * np, 0.0);
        std::vector<double> B[2], tq[2], Tq[2];
        for (int q = 0; q < nq; ++q) {
            const double ws = w[q] * dJ[q];
            double x[3] = {0, 0, 0};
            for (int c = 0; c < D; ++c)
                for (int v = 0; v < nvs; ++v) x[c] += phi[q][v] * Xf[v][c];
            for (int X = 0; X < 2; ++X) {
                const double* box = d.ip_box + (size_t)eX[X] * 2 * D;
                double xip[3] = {0, 0, 0};
                for (int c = 0; c < D; ++c) xip[c] = 2.0 * (x[c] - box[c]) / box[D + c] - 1.0;
                double dN[3][27];
                basis_dN(D, nb, xip, dN);
                double dNdx[3][27];
                for (int r = 0; r < D; ++r)
                    for (int a = 0; a < nb; ++a) dNdx[r][a] = dN[r][a] * 2.0 / box[D + r];
                B[X].assign((size_t)ns * np, 0.0);
                if (elast && D == 3) {
                    for (int a = 0; a < nb; ++a) {
                        B[X][0 * np + a] = dNdx[0][a]; B[X][4 * np + a] = dNdx[2][a]; B[X][5 * np + a] = dNdx[1][a];
                        B[X][1 * np + nb + a] = dNdx[1][a]; B[X][3 * np + nb + a] = dNdx[2][a]; B[X][5 * np + nb + a] = dNdx[0][a];
                        B[X][2 * np + 2 * nb + a] = dNdx[2][a]; B[X][3 * np + 2 * nb + a] = dNdx[1][a];
                        B[X][4 * np + 2 * nb + a] = dNdx[0][a];
                    }
                } else if (elast) {
                    for (int a = 0; a < nb; ++a) {
                        B[X][0 * np + a] = dNdx[0][a]; B[X][2 * np + a] = dNdx[1][a];
                        B[X][1 * np + nb + a] = dNdx[1][a]; B[X][2 * np + nb + a] = dNdx[0][a];
                    }
                } else {
                    for (int i = 0; i < D; ++i)
                        for (int a = 0; a < nb; ++a) B[X][(size_t)i * np + a] = dNdx[i][a];
                }
                // test traction n_flat B and traction n_flat D B (diffusion: n^T B, k n^T B)
                tq[X].assign((size_t)nf * np, 0.0);
                Tq[X].assign((size_t)nf * np, 0.0);
                for (int f = 0; f < nf; ++f)
                    for (int c = 0; c < np; ++c) {
                        double t = 0.0, tt = 0.0;
                        for (int i = 0; i < ns; ++i) {
                            t += nfl[f][i] * B[X][(size_t)i * np + c];
                            double db = 0.0;
                            if (elast)
                                for (int j = 0; j < ns; ++j) db += Dm[X][i][j] * B[X][(size_t)j * np + c];
                            else
                                db = kc[X] * B[X][(size_t)i * np + c];
                            tt += nfl[f][i] * db;
                        }
                        tq[X][(size_t)f * np + c] = t;
                        Tq[X][(size_t)f * np + c] = tt;
                    }
            }
            for (int X = 0; X < 2; ++X)
                for (int Y = 0; Y < 2; ++Y) {
                    const double sg = (X == Y ? 1.0 : -1.0) * ws * gamma;
                    for (int i = 0; i < np; ++i)
                        for (int j = 0; j < np; ++j) {
                            double v = 0.0;
                            for (int f = 0; f < nf; ++f) v += tq[X][(size_t)f * np + i] * Tq[Y][(size_t)f * np + j];
                            Kp[(size_t)(X * np + i) * 2 * np + Y * np + j] += sg * v;
                        }
                }
        }
        // K_e = T^T K_p T, R_e = K_e u_U, added into the CSR values / R
        std::vector<double> KT((size_t)2 * np * nu, 0.0), Ke((size_t)nu * nu, 0.0);
        for (int i = 0; i < 2 * np; ++i)
            for (int b = 0; b < nu; ++b) {
                double v = 0.0;
                for (int j = 0; j < 2 * np; ++j) v += Kp[(size_t)i * 2 * np + j] * T[(size_t)j * nu + b];
                KT[(size_t)i * nu + b] = v;
            }
        for (int a = 0; a < nu; ++a)
            for (int b = 0; b < nu; ++b) {
                double v = 0.0;
                for (int i = 0; i < 2 * np; ++i) v += T[(size_t)i * nu + a] * KT[(size_t)i * nu + b];
                Ke[(size_t)a * nu + b] = v;
            }
        for (int a = 0; a < nu; ++a) {
            const int32_t r = U[a];
            double rr = 0.0;
            for (int b = 0; b < nu; ++b) {
                rr += Ke[(size_t)a * nu + b] * u[U[b]];
                const int32_t* cb = col_idx + row_ptr[r];
                const int32_t* ce = col_idx + row_ptr[r + 1];
                const int32_t* it = std::lower_bound(cb, ce, U[b]);
                if (it == ce || *it != U[b]) throw std::runtime_error("ghost entry outside the graph");
                vals[it - col_idx] += Ke[(size_t)a * nu + b];
            }
            rhs[r] += rr;
        }
    }
}

mgpu_set_desc desc_of(const OracleProblem& p) {
    mgpu_set_desc d;
    std::memset(&d, 0, sizeof(d));
    d.spatial_dim = p.dim;
    d.iwg = p.spec.iwg;
    d.ip_geometry = p.nb_ip == 27 ? MGPU_GEOM_HEX27 : (p.dim == 3 ? MGPU_GEOM_HEX8 : MGPU_GEOM_QUAD4);
    d.ig_geometry = p.dim == 3 ? MGPU_GEOM_TET4 : MGPU_GEOM_TRI3;
    d.n_elements = (int32_t)p.elem_cell.size();
    d.n_bases = p.nb_ip;
    d.n_adofs_per_type = p.N;
    d.ip_box = p.ip_box.data();
    d.material = p.material.data();
    d.t_ptr = p.t_ptr.data();
    d.t_adof = p.t_adof.data();
    d.t_weight = p.t_weight.data();
    d.ig_ptr = p.ig_ptr.data();
    d.ig_xi = p.ig_xi.data();
    d.elem_cell = p.elem_cell.data();
    d.n_couplings = (int32_t)(p.couplings.size() / 2);
    d.couplings = p.couplings.empty() ? nullptr : p.couplings.data();
    return d;
}

}  // namespace

// ================================================================================================
// C API (ctypes)
// ================================================================================================
extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void* oracle_build(const mgpu_problem_spec* s) {
    try {
        auto p = std::make_unique<OracleProblem>();
        build(*p, *s);
        mgpu_set_desc d = desc_of(*p);
        build_graph(d, p->nf, p->row_ptr, p->col_idx);
        return p.release();
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void oracle_destroy(void* h) { delete static_cast<OracleProblem*>(h); }

// counts: [0] elements, [1] cut elements, [2] IG sub-cells, [3] adofs per type, [4] rows, [5] nnz
void oracle_counts(void* h, int64_t* c) {
    const OracleProblem& p = *static_cast<OracleProblem*>(h);
    c[0] = (int64_t)p.elem_cell.size();
    c[1] = p.n_cut;
    c[2] = p.ig_ptr.back();
    c[3] = p.N;
    c[4] = (int64_t)p.row_ptr.size() - 1;
    c[5] = (int64_t)p.col_idx.size();
    c[6] = (int64_t)p.t_adof.size();
}

// copies of the set-up arrays (sizes from oracle_counts)
void oracle_get_setup(void* h, int32_t* elem_cell, int32_t* ig_ptr, double* ig_xi, double* material,
                      int32_t* t_adof, int32_t* row_ptr, int32_t* col_idx, int32_t* t_ptr, double* t_weight) {
    const OracleProblem& p = *static_cast<OracleProblem*>(h);
    if (t_ptr) std::copy(p.t_ptr.begin(), p.t_ptr.end(), t_ptr);
    if (t_weight) std::copy(p.t_weight.begin(), p.t_weight.end(), t_weight);
    if (elem_cell) std::copy(p.elem_cell.begin(), p.elem_cell.end(), elem_cell);
    if (ig_ptr) std::copy(p.ig_ptr.begin(), p.ig_ptr.end(), ig_ptr);
    if (ig_xi) std::copy(p.ig_xi.begin(), p.ig_xi.end(), ig_xi);
    if (material) std::copy(p.material.begin(), p.material.end(), material);
    if (t_adof) std::copy(p.t_adof.begin(), p.t_adof.end(), t_adof);
    if (row_ptr) std::copy(p.row_ptr.begin(), p.row_ptr.end(), row_ptr);
    if (col_idx) std::copy(p.col_idx.begin(), p.col_idx.end(), col_idx);
}

// IP boxes of the elements: [n_elements][2 dim], lower corner then extent.
void oracle_get_ip_box(void* h, double* ip_box) {
    const OracleProblem& p = *static_cast<OracleProblem*>(h);
    std::copy(p.ip_box.begin(), p.ip_box.end(), ip_box);
}

// Assemble the oracle's own problem.  vals[nnz], rhs[rows] are zeroed then summed.
int oracle_assemble(void* h, const double* u, double* vals, double* rhs) {
    try {
        const OracleProblem& p = *static_cast<OracleProblem*>(h);
        mgpu_set_desc d = desc_of(p);
        std::fill(vals, vals + p.col_idx.size(), 0.0);
        std::fill(rhs, rhs + p.row_ptr.size() - 1, 0.0);
        Scratch s;
        for (int64_t e = 0; e < d.n_elements; ++e)
            element_assemble(d, e, u, p.row_ptr.data(), p.col_idx.data(), vals, rhs, s, nullptr);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// Reference algorithm on an external descriptor and graph, elements [e0, e1), with n_threads
// workers over a static element partition; each worker scatters into private arrays that are summed
// row-wise at the end (the stand-in for Epetra's GlobalAssemble).  Returns the number of IPs.
int64_t oracle_assemble_desc(const mgpu_set_desc* d, const int32_t* row_ptr, const int32_t* col_idx,
                             int32_t n_rows, const double* u, double* vals, double* rhs, int64_t e0,
                             int64_t e1, int32_t n_threads) {
    try {
        const int64_t nnz = row_ptr[n_rows];
        std::fill(vals, vals + nnz, 0.0);
        std::fill(rhs, rhs + n_rows, 0.0);
        if (n_threads <= 1) {
            Scratch s;
            int64_t nip = 0;
            for (int64_t e = e0; e < e1; ++e) element_assemble(*d, e, u, row_ptr, col_idx, vals, rhs, s, &nip);
            return nip;
        }
        std::vector<std::vector<double>> pv(n_threads), pr(n_threads);
        std::vector<int64_t> nips(n_threads, 0);
        std::vector<std::string> errs(n_threads);
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; ++t)
            th.emplace_back([&, t]() {
                try {
                    pv[t].assign(nnz, 0.0);
                    pr[t].assign(n_rows, 0.0);
                    Scratch s;
                    const int64_t a = e0 + (e1 - e0) * t / n_threads, b = e0 + (e1 - e0) * (t + 1) / n_threads;
                    for (int64_t e = a; e < b; ++e)
                        element_assemble(*d, e, u, row_ptr, col_idx, pv[t].data(), pr[t].data(), s, &nips[t]);
                } catch (const std::exception& ex) {
                    errs[t] = ex.what();
                }
            });
        for (auto& x : th) x.join();
        for (auto& e : errs)
            if (!e.empty()) throw std::runtime_error(e);
        for (int t = 0; t < n_threads; ++t) {
            for (int64_t i = 0; i < nnz; ++i) vals[i] += pv[t][i];
            for (int64_t i = 0; i < n_rows; ++i) rhs[i] += pr[t][i];
        }
        int64_t nip = 0;
        for (auto v : nips) nip += v;
        return nip;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// building blocks exposed for the golden-vector / closed-form tests
// ------------------------------------------------------------------------------------------------
extern "C" {

// Lagrange QUAD4 (dim 2) / HEX8 (dim 3) shape functions at n points xi[n][dim]:
// N[n][nb], dN[n][dim][nb].
void oracle_shape(int dim, int n, const double* xi, double* N, double* dN) {
    const int nb = dim == 3 ? 8 : 4;
    for (int k = 0; k < n; ++k) {
        double xx[3] = {xi[k * dim], xi[k * dim + 1], dim == 3 ? xi[k * dim + 2] : 0.0};
        if (N) shape_N(dim, xx, N + (size_t)k * nb);
        if (dN) {
            double d[3][8];
            shape_dN(dim, xx, d);
            for (int i = 0; i < dim; ++i)
                for (int a = 0; a < nb; ++a) dN[((size_t)k * dim + i) * nb + a] = d[i][a];
        }
    }
}

// HEX27 shape functions at n points xi[n][3]: N[n][27], dN[n][3][27].
void oracle_shape_hex27(int n, const double* xi, double* N, double* dN) {
    for (int k = 0; k < n; ++k) {
        if (N) hex27_N(xi + 3 * k, N + (size_t)k * 27);
        if (dN) {
            double d[3][27];
            hex27_dN(xi + 3 * k, d);
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 27; ++a) dN[((size_t)k * 3 + i) * 27 + a] = d[i][a];
        }
    }
}

// Quadrature rule: 0 QUAD_2x2, 1 HEX_2x2x2, 2 TRI_7, 3 TET_11, 4 HEX_3x3x3, 5 TET_35.  Writes
// points [n][dim] and weights [n]; returns n.
// Node hierarchy of one IG cell (nh_cell): vertex global ids g[dim+1], level-set values f[dim+1],
// coordinates x[dim+1][3]; writes the sub-cells of phase `phase` as [n][dim+1][3] into out (at most
// max_cells) and returns n, or -1 on error.
int oracle_nh_cell(int dim, const uint32_t* g, const double* f, const double* x, int phase, double* out,
                   int max_cells) {
    try {
        std::array<double, 3> xa[4];
        for (int q = 0; q <= dim; ++q) xa[q] = {x[3 * q], x[3 * q + 1], x[3 * q + 2]};
        std::vector<double> sub;
        nh_cell(dim, g, f, xa, phase, sub);
        const int nv = dim + 1;
        const int n = (int)(sub.size() / (nv * dim));
        if (n > max_cells) throw std::runtime_error("too many sub-cells");
        for (int c = 0; c < n; ++c)
            for (int q = 0; q < nv; ++q)
                for (int d = 0; d < 3; ++d) out[(c * nv + q) * 3 + d] = d < dim ? sub[(c * nv + q) * dim + d] : 0.0;
        return n;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Facets of the oracle's problem (problem_facets): returns their count n; elem[n], xi[n][nvs][dim],
// normal[n][dim] are written when the arrays are given (max facets), or -1 on error.
int64_t oracle_facets(void* h, int where, int nvs, int32_t* elem, double* xi, double* normal, int64_t max) {
    try {
        const OracleProblem& p = *static_cast<OracleProblem*>(h);
        std::vector<int32_t> fe;
        std::vector<double> fx, fn;
        problem_facets(p, where, nvs, fe, fx, fn);
        const int64_t n = (int64_t)fe.size();
        if (elem && n <= max) {
            std::copy(fe.begin(), fe.end(), elem);
            std::copy(fx.begin(), fx.end(), xi);
            std::copy(fn.begin(), fn.end(), normal);
        }
        return n;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Side set on an external descriptor and graph (side_assemble), ADDED into vals / rhs.
int oracle_side_assemble(const mgpu_set_desc* d, const int32_t* row_ptr, const int32_t* col_idx, int iwg, int nvs,
                         int64_t n_facets, const int32_t* elem, const double* xi, const double* normal, double alpha,
                         const double* value, const double* u, double* vals, double* rhs) {
    try {
        side_assemble(*d, iwg, nvs, n_facets, elem, xi, normal, alpha, value, u, row_ptr, col_idx, vals, rhs);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

int oracle_ghost_assemble(const mgpu_set_desc* d, const int32_t* row_ptr, const int32_t* col_idx, int64_t n_facets,
                          const int32_t* leader, const int32_t* follower, const double* xi, const double* normal,
                          double alpha, const double* u, double* vals, double* rhs) {
    try {
        ghost_assemble(*d, n_facets, leader, follower, xi, normal, alpha, u, row_ptr, col_idx, vals, rhs);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// ghost facets of an oracle problem: (leader, follower, axis) triples; returns the count (at most
// `max` written)
int64_t oracle_ghost_facets(void* h, int32_t* trip, int64_t max) {
    const OracleProblem& p = *static_cast<OracleProblem*>(h);
    std::vector<int32_t> g;
    ghost_facets(p, g);
    const int64_t n = (int64_t)(g.size() / 3);
    for (int64_t k = 0; k < std::min(n, max); ++k)
        for (int c = 0; c < 3; ++c) trip[3 * k + c] = g[3 * k + c];
    return n;
}

int oracle_rule(int rule, double* xi, double* w) {
    switch (rule) {
        case 4: {
            double x[27][3];
            hex3_rule(x, w);
            for (int q = 0; q < 27; ++q)
                for (int d = 0; d < 3; ++d) xi[3 * q + d] = x[q][d];
            return 27;
        }
        case 5: {
            double x[35][3];
            tet35_rule(x, w);
            for (int q = 0; q < 35; ++q)
                for (int d = 0; d < 3; ++d) xi[3 * q + d] = x[q][d];
            return 35;
        }
        case 0:
            for (int q = 0; q < 4; ++q) { xi[2 * q] = kQuad2[q][0]; xi[2 * q + 1] = kQuad2[q][1]; w[q] = 1.0; }
            return 4;
        case 1:
            for (int q = 0; q < 8; ++q) { for (int d = 0; d < 3; ++d) xi[3 * q + d] = kHex2[q][d]; w[q] = 1.0; }
            return 8;
        case 2:
            for (int q = 0; q < 7; ++q) { xi[2 * q] = kTri7P[q][0]; xi[2 * q + 1] = kTri7P[q][1]; w[q] = kTri7W[q]; }
            return 7;
        case 3:
            for (int q = 0; q < 11; ++q) { for (int d = 0; d < 3; ++d) xi[3 * q + d] = kTet11P[q][d]; w[q] = tet11_weight(q); }
            return 11;
        default:
            return 0;
    }
}

// Dense element matrices of element e of the oracle's own problem (reference flow, before scatter):
// ids[na], Ke[na][na], Re[na] for adof vector u.  Returns na.
int oracle_element(void* h, int64_t e, const double* u, int32_t* ids, double* Ke, double* Re) {
    try {
        const OracleProblem& p = *static_cast<OracleProblem*>(h);
        mgpu_set_desc d = desc_of(p);
        Scratch s;
        // scatter into a private one-element graph
        std::vector<int32_t> tids;
        element_topology(d, p.nf, e, tids);
        const int na = (int)tids.size();
        int32_t nrows = p.nf * p.N;
        std::vector<int32_t> rp(nrows + 1, 0), ci;
        for (int32_t r = 0; r < nrows; ++r) {
            rp[r + 1] = rp[r];
            if (std::binary_search(tids.begin(), tids.end(), r)) {
                ci.insert(ci.end(), tids.begin(), tids.end());
                rp[r + 1] += na;
            }
        }
        std::vector<double> vals(ci.size(), 0.0), rhs(nrows, 0.0);
        element_assemble(d, e, u, rp.data(), ci.data(), vals.data(), rhs.data(), s, nullptr);
        for (int a = 0; a < na; ++a) {
            ids[a] = tids[a];
            Re[a] = s.Re[a];
            for (int b = 0; b < na; ++b) Ke[a * na + b] = s.Ke[(size_t)a * na + b];
        }
        return na;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

}  // extern "C"

extern "C" {
// Reference algorithm over an explicit element list (a sample of a workload), n_threads workers.
int64_t oracle_assemble_list(const mgpu_set_desc* d, const int32_t* row_ptr, const int32_t* col_idx,
                             int32_t n_rows, const double* u, double* vals, double* rhs, const int64_t* elems,
                             int64_t n_elems, int32_t n_threads) {
    try {
        const int64_t nnz = row_ptr[n_rows];
        std::fill(vals, vals + nnz, 0.0);
        std::fill(rhs, rhs + n_rows, 0.0);
        if (n_threads < 1) n_threads = 1;
        std::vector<std::vector<double>> pv(n_threads), pr(n_threads);
        std::vector<int64_t> nips(n_threads, 0);
        std::vector<std::string> errs(n_threads);
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; ++t)
            th.emplace_back([&, t]() {
                try {
                    double* v = vals;
                    double* r = rhs;
                    if (n_threads > 1) {
                        pv[t].assign(nnz, 0.0);
                        pr[t].assign(n_rows, 0.0);
                        v = pv[t].data();
                        r = pr[t].data();
                    }
                    Scratch s;
                    const int64_t a = n_elems * t / n_threads, b = n_elems * (t + 1) / n_threads;
                    for (int64_t i = a; i < b; ++i) element_assemble(*d, elems[i], u, row_ptr, col_idx, v, r, s, &nips[t]);
                } catch (const std::exception& ex) {
                    errs[t] = ex.what();
                }
            });
        for (auto& x : th) x.join();
        for (auto& e : errs)
            if (!e.empty()) throw std::runtime_error(e);
        if (n_threads > 1) {
            std::vector<std::thread> red;
            for (int t = 0; t < n_threads; ++t)
                red.emplace_back([&, t]() {
                    const int64_t a = nnz * t / n_threads, b = nnz * (t + 1) / n_threads;
                    for (int q = 0; q < n_threads; ++q)
                        for (int64_t i = a; i < b; ++i) vals[i] += pv[q][i];
                    const int64_t ra = (int64_t)n_rows * t / n_threads, rb = (int64_t)n_rows * (t + 1) / n_threads;
                    for (int q = 0; q < n_threads; ++q)
                        for (int64_t i = ra; i < rb; ++i) rhs[i] += pr[q][i];
                });
            for (auto& x : red) x.join();
        }
        int64_t nip = 0;
        for (auto v : nips) nip += v;
        return nip;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}
}  // extern "C"
