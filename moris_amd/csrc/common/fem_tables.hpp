// fem_tables.hpp -- quadrature rules, cut templates and element conventions shared by the host
// set-up and the HIP kernels.  Tables whose values must equal the reference's (so that K agrees to
// the last bits with MORIS on the same cut) are restated here in this project's own layout; each
// block cites the reference file it follows.
#pragma once

#ifdef __HIPCC__
#define MGPU_TABLE static __constant__ const
#define MGPU_HD __host__ __device__
#else
#define MGPU_TABLE static const
#define MGPU_HD
#endif

namespace mgpu {

// ---------------------------------------------------------------------------------------------
// Gauss and simplex rules.  Points are stored point-major [n][dim]; the literal digits follow
// projects/MTK/src/integration/cl_MTK_Integration_Coeffs_*.hpp (the reference keeps 15 digits,
// e.g. 0.577350269189626 instead of 1/sqrt(3); the same literals are kept here on purpose).
// ---------------------------------------------------------------------------------------------

// QUAD_2x2  (cl_MTK_Integration_Coeffs_Quad_2x2.hpp)
constexpr int kQuad2Points = 4;
MGPU_TABLE double kQuad2Xi[4][2] = {
    {-0.577350269189626, -0.577350269189626},
    { 0.577350269189626, -0.577350269189626},
    { 0.577350269189626,  0.577350269189626},
    {-0.577350269189626,  0.577350269189626}};
MGPU_TABLE double kQuad2W[4] = {1.0, 1.0, 1.0, 1.0};

// HEX_2x2x2  (cl_MTK_Integration_Coeffs_Hex_2x2x2.hpp)
constexpr int kHex2Points = 8;
MGPU_TABLE double kHex2Xi[8][3] = {
    {-0.577350269189626, -0.577350269189626, -0.577350269189626},
    { 0.577350269189626, -0.577350269189626, -0.577350269189626},
    { 0.577350269189626,  0.577350269189626, -0.577350269189626},
    {-0.577350269189626,  0.577350269189626, -0.577350269189626},
    {-0.577350269189626, -0.577350269189626,  0.577350269189626},
    { 0.577350269189626, -0.577350269189626,  0.577350269189626},
    { 0.577350269189626,  0.577350269189626,  0.577350269189626},
    {-0.577350269189626,  0.577350269189626,  0.577350269189626}};
MGPU_TABLE double kHex2W[8] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};

// HEX_3x3x3  (cl_MTK_Integration_Coeffs_Hex_3x3x3.hpp): points in HEX27 node order per layer.
constexpr int kHex3Points = 27;
#define MGPU_G3 0.774596669241483
#define MGPU_W1 0.171467764060357
#define MGPU_W2 0.274348422496571
#define MGPU_W3 0.438957475994513
#define MGPU_W4 0.702331961591221
MGPU_TABLE double kHex3Xi[27][3] = {
    {-MGPU_G3, -MGPU_G3, -MGPU_G3}, { MGPU_G3, -MGPU_G3, -MGPU_G3}, { MGPU_G3,  MGPU_G3, -MGPU_G3},
    {-MGPU_G3,  MGPU_G3, -MGPU_G3}, {     0.0, -MGPU_G3, -MGPU_G3}, { MGPU_G3,      0.0, -MGPU_G3},
    {     0.0,  MGPU_G3, -MGPU_G3}, {-MGPU_G3,      0.0, -MGPU_G3}, {     0.0,      0.0, -MGPU_G3},
    {-MGPU_G3, -MGPU_G3,      0.0}, { MGPU_G3, -MGPU_G3,      0.0}, { MGPU_G3,  MGPU_G3,      0.0},
    {-MGPU_G3,  MGPU_G3,      0.0}, {     0.0, -MGPU_G3,      0.0}, { MGPU_G3,      0.0,      0.0},
    {     0.0,  MGPU_G3,      0.0}, {-MGPU_G3,      0.0,      0.0}, {     0.0,      0.0,      0.0},
    {-MGPU_G3, -MGPU_G3,  MGPU_G3}, { MGPU_G3, -MGPU_G3,  MGPU_G3}, { MGPU_G3,  MGPU_G3,  MGPU_G3},
    {-MGPU_G3,  MGPU_G3,  MGPU_G3}, {     0.0, -MGPU_G3,  MGPU_G3}, { MGPU_G3,      0.0,  MGPU_G3},
    {     0.0,  MGPU_G3,  MGPU_G3}, {-MGPU_G3,      0.0,  MGPU_G3}, {     0.0,      0.0,  MGPU_G3}};
MGPU_TABLE double kHex3W[27] = {
    MGPU_W1, MGPU_W1, MGPU_W1, MGPU_W1, MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W3,
    MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W3, MGPU_W3, MGPU_W3, MGPU_W3, MGPU_W4,
    MGPU_W1, MGPU_W1, MGPU_W1, MGPU_W1, MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W2, MGPU_W3};
#undef MGPU_G3
#undef MGPU_W1
#undef MGPU_W2
#undef MGPU_W3
#undef MGPU_W4

// TRI_7  (cl_MTK_Integration_Coeffs_Tri_7.hpp): (xi1, xi2); weights sum to 1, detJ = area.
constexpr int kTri7Points = 7;
MGPU_TABLE double kTri7Xi[7][2] = {
    {0.333333333333333, 0.333333333333333},
    {0.797426985353087, 0.101286507323456},
    {0.101286507323456, 0.797426985353087},
    {0.101286507323456, 0.101286507323456},
    {0.059715871789770, 0.470142064105115},
    {0.470142064105115, 0.059715871789770},
    {0.470142064105115, 0.470142064105115}};
MGPU_TABLE double kTri7W[7] = {
    0.225000000000000, 0.125939180544827, 0.125939180544827, 0.125939180544827,
    0.132394152788506, 0.132394152788506, 0.132394152788506};

// TRI_3  (cl_MTK_Integration_Coeffs_Tri_3.hpp)
constexpr int kTri3Points = 3;
MGPU_TABLE double kTri3Xi[3][2] = {
    {0.666666666666667, 0.166666666666667},
    {0.166666666666667, 0.666666666666667},
    {0.166666666666667, 0.166666666666667}};
MGPU_TABLE double kTri3W[3] = {0.333333333333333, 0.333333333333333, 0.333333333333333};

// TET_11  (cl_MTK_Integration_Coeffs_Tet_11.hpp): (xi1, xi2, xi3); the reference stores the
// weights divided by 6 and multiplies by 6.0 when it hands them out, which is restated literally.
constexpr int kTet11Points = 11;
MGPU_TABLE double kTet11Xi[11][3] = {
    {0.7857142857142860, 0.0714285714285714, 0.0714285714285714},
    {0.0714285714285714, 0.7857142857142860, 0.0714285714285714},
    {0.0714285714285714, 0.0714285714285714, 0.0714285714285714},
    {0.0714285714285714, 0.0714285714285714, 0.7857142857142860},
    {0.3994035761667990, 0.3994035761667990, 0.1005964238332010},
    {0.3994035761667990, 0.1005964238332010, 0.1005964238332010},
    {0.3994035761667990, 0.1005964238332010, 0.3994035761667990},
    {0.1005964238332010, 0.3994035761667990, 0.1005964238332010},
    {0.1005964238332010, 0.3994035761667990, 0.3994035761667990},
    {0.1005964238332010, 0.1005964238332010, 0.3994035761667990},
    {0.2500000000000000, 0.2500000000000000, 0.2500000000000000}};
MGPU_TABLE double kTet11W[11] = {
    6.0 * 0.0076222222222222, 6.0 * 0.0076222222222222, 6.0 * 0.0076222222222222,
    6.0 * 0.0076222222222222, 6.0 * 0.0248888888888889, 6.0 * 0.0248888888888889,
    6.0 * 0.0248888888888889, 6.0 * 0.0248888888888889, 6.0 * 0.0248888888888889,
    6.0 * 0.0248888888888889, 6.0 * -0.0131555555555556};

// ---------------------------------------------------------------------------------------------
// Lagrange parametric node coordinates (vertex order of MTK's interpolation functions).
// HEX8: cl_MTK_Interpolation_Function_Lagrange_Hex8.hpp get_param_coords;
// QUAD4: counter-clockwise from (-1,-1).
// ---------------------------------------------------------------------------------------------
MGPU_TABLE double kHex8XiHat[8][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
    {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
MGPU_TABLE double kQuad4XiHat[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
// Lattice offsets of the IP-cell vertices (same order).
MGPU_TABLE int kHex8Off[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
MGPU_TABLE int kQuad4Off[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

// ---------------------------------------------------------------------------------------------
// XTK regular subdivision (cl_XTK_Regular_Subdivision_Interface.hpp).
// HEX8 -> 24 TET4: vertices 0-7 the hex corners, 8-13 the face midpoints (faces 0..5), 14 the
// centre; QUAD4 -> 4 TRI3: vertices 0-3 the corners, 4 the centre.
// ---------------------------------------------------------------------------------------------
MGPU_TABLE double kHexFaceXi[6][3] = {
    {0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
    {-1.0, 0.0, 0.0}, {0.0, 0.0, -1.0}, {0.0, 0.0, 1.0}};
MGPU_TABLE int kHex24Tets[24][4] = {
    {0, 8, 1, 14},  {1, 8, 5, 14},  {4, 5, 8, 14},  {0, 4, 8, 14},
    {1, 9, 2, 14},  {2, 9, 6, 14},  {5, 6, 9, 14},  {1, 5, 9, 14},
    {2, 10, 3, 14}, {2, 6, 10, 14}, {6, 7, 10, 14}, {3, 10, 7, 14},
    {0, 3, 11, 14}, {3, 7, 11, 14}, {4, 11, 7, 14}, {0, 11, 4, 14},
    {0, 1, 12, 14}, {1, 2, 12, 14}, {2, 3, 12, 14}, {0, 12, 3, 14},
    {4, 13, 5, 14}, {5, 13, 6, 14}, {6, 13, 7, 14}, {4, 7, 13, 14}};
MGPU_TABLE int kQuad4Tris[4][3] = {{0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4}};

// Edge ordinals (mtk Cell_Info_Tet4 / Cell_Info_Tri3 get_node_to_edge_map).
MGPU_TABLE int kTetEdges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
MGPU_TABLE int kTriEdges[3][2] = {{0, 1}, {1, 2}, {2, 0}};

// ---------------------------------------------------------------------------------------------
// Node-hierarchy templates (cl_XTK_Node_Hierarchy_Template_Library.cpp).  The reference lists
// 96 tet permutation ids; they fall into 8 connectivity groups.  Ordinals index the sorted node
// list built by Node_Hierarchy_Interface::sort_nodes_3d (0-3 tet vertices, 4.. intersection
// vertices by ascending id).  Group 0/1: three intersected edges (4 tets); 2-7: four (6 tets).
// ---------------------------------------------------------------------------------------------
MGPU_TABLE int kNhTet3Cells[2][4][4] = {
    {{0, 4, 5, 6}, {4, 1, 5, 6}, {1, 2, 5, 6}, {1, 3, 2, 6}},
    {{4, 0, 5, 6}, {1, 4, 5, 6}, {2, 1, 5, 6}, {3, 1, 2, 6}}};
MGPU_TABLE int kNhTet4Cells[6][6][4] = {
    {{0, 3, 7, 6}, {4, 3, 6, 7}, {4, 3, 7, 5}, {1, 2, 7, 5}, {2, 4, 7, 5}, {4, 2, 7, 6}},
    {{3, 0, 7, 6}, {3, 4, 6, 7}, {3, 4, 7, 5}, {2, 1, 7, 5}, {4, 2, 7, 5}, {2, 4, 7, 6}},
    {{4, 0, 6, 7}, {0, 3, 6, 7}, {3, 5, 6, 7}, {5, 1, 6, 7}, {1, 2, 6, 7}, {2, 4, 6, 7}},
    {{0, 4, 6, 7}, {3, 0, 6, 7}, {5, 3, 6, 7}, {1, 5, 6, 7}, {2, 1, 6, 7}, {4, 2, 6, 7}},
    {{4, 0, 5, 7}, {0, 3, 5, 7}, {5, 3, 6, 7}, {1, 2, 6, 7}, {2, 4, 5, 7}, {2, 5, 6, 7}},
    {{0, 4, 5, 7}, {3, 0, 5, 7}, {3, 5, 6, 7}, {2, 1, 6, 7}, {4, 2, 5, 7}, {5, 2, 6, 7}}};
// Permutation id -> group.  Three edges: id = low + 10*mid + 100*high (edge ordinals sorted by
// intersection-vertex id); four edges: low + 10*midlow + 100*midhigh + 1000*high.
MGPU_TABLE int kNhPerm3[24][2] = {
    {320, 0}, {32, 0},  {203, 0}, {251, 0}, {512, 0}, {125, 0}, {140, 0}, {401, 0},
    {14, 0},  {453, 0}, {534, 0}, {345, 0}, {230, 1}, {302, 1}, {23, 1},  {521, 1},
    {152, 1}, {215, 1}, {410, 1}, {41, 1},  {104, 1}, {543, 1}, {354, 1}, {435, 1}};
MGPU_TABLE int kNhPerm4[72][2] = {
    {5420, 0}, {5240, 0}, {425, 0},  {3501, 0}, {1503, 0}, {3051, 0}, {1053, 0}, {4312, 0},
    {2314, 0}, {4132, 0}, {2134, 0}, {245, 0},  {4502, 1}, {4052, 1}, {1243, 1}, {2504, 1},
    {2054, 1}, {5310, 1}, {5130, 1}, {135, 1},  {315, 1},  {3421, 1}, {3241, 1}, {1423, 1},
    {4250, 2}, {2450, 2}, {4205, 2}, {2405, 2}, {5031, 2}, {5013, 2}, {531, 2},  {513, 2},
    {3124, 2}, {1342, 2}, {1324, 2}, {3142, 2}, {5024, 3}, {524, 3},  {5042, 3}, {542, 3},
    {3150, 3}, {1350, 3}, {3105, 3}, {1305, 3}, {4231, 3}, {2431, 3}, {4213, 3}, {2413, 3},
    {4520, 4}, {2540, 4}, {4025, 4}, {2045, 4}, {5301, 4}, {5103, 4}, {351, 4},  {153, 4},
    {3412, 4}, {3214, 4}, {1432, 4}, {1234, 4}, {5402, 5}, {452, 5},  {5204, 5}, {254, 5},
    {3510, 5}, {1530, 5}, {3015, 5}, {1035, 5}, {4321, 5}, {2341, 5}, {4123, 5}, {2143, 5}};
// Sub-cells with vertices on the isocontour (|phi| <= 1e-12): edges touching such a vertex are not
// intersected.  One intersected edge: id 10000 + edge ordinal, nodes [tet vertices, P];
// two: 10000 + 100*(low+1) + 10*(high+1), nodes [tet vertices, P_low, P_high].
MGPU_TABLE int kNhTet1Cells[6][2][4] = {
    {{0, 4, 2, 3}, {4, 1, 2, 3}}, {{0, 1, 4, 3}, {0, 4, 2, 3}}, {{0, 1, 4, 3}, {4, 1, 2, 3}},
    {{0, 1, 2, 4}, {4, 1, 2, 3}}, {{0, 4, 2, 3}, {0, 1, 2, 4}}, {{0, 1, 4, 3}, {0, 1, 2, 4}}};
MGPU_TABLE int kNhTet2Perm[24] = {10250, 10520, 10260, 10620, 10560, 10650, 10640, 10460,
                                  10230, 10320, 10130, 10310, 10120, 10210, 10140, 10410,
                                  10150, 10510, 10450, 10540, 10360, 10630, 10340, 10430};
MGPU_TABLE int kNhTet2Cells[24][3][4] = {
    {{0, 1, 4, 5}, {0, 4, 2, 5}, {0, 5, 2, 3}}, {{0, 1, 5, 4}, {0, 4, 5, 3}, {0, 5, 2, 3}},
    {{0, 1, 4, 5}, {0, 1, 5, 3}, {0, 4, 2, 5}}, {{0, 1, 5, 3}, {0, 5, 4, 3}, {0, 5, 2, 4}},
    {{0, 1, 5, 4}, {0, 1, 2, 5}, {0, 4, 5, 3}}, {{0, 1, 2, 5}, {0, 5, 2, 4}, {0, 5, 4, 3}},
    {{0, 1, 2, 5}, {5, 1, 4, 3}, {1, 2, 5, 4}}, {{0, 1, 2, 5}, {0, 1, 5, 4}, {4, 1, 5, 3}},
    {{5, 4, 2, 3}, {1, 4, 5, 3}, {0, 1, 5, 3}}, {{5, 2, 4, 3}, {0, 5, 4, 3}, {0, 1, 5, 3}},
    {{5, 1, 2, 3}, {5, 4, 1, 3}, {0, 4, 5, 3}}, {{5, 1, 2, 3}, {4, 5, 2, 3}, {0, 5, 4, 3}},
    {{0, 5, 2, 3}, {0, 4, 5, 3}, {4, 1, 5, 3}}, {{0, 5, 2, 3}, {5, 4, 2, 3}, {5, 1, 4, 3}},
    {{0, 4, 2, 5}, {4, 1, 2, 5}, {5, 1, 2, 3}}, {{0, 5, 2, 4}, {4, 5, 2, 3}, {5, 1, 2, 3}},
    {{0, 4, 2, 5}, {4, 1, 2, 5}, {0, 5, 2, 3}}, {{5, 1, 2, 4}, {5, 4, 2, 3}, {0, 5, 2, 3}},
    {{0, 5, 2, 4}, {0, 1, 2, 5}, {4, 5, 2, 3}}, {{0, 1, 2, 5}, {5, 1, 2, 4}, {5, 4, 2, 3}},
    {{0, 1, 5, 3}, {0, 1, 4, 5}, {1, 2, 4, 5}}, {{0, 1, 5, 3}, {1, 4, 5, 3}, {1, 2, 5, 4}},
    {{0, 1, 4, 5}, {4, 1, 2, 5}, {1, 2, 5, 3}}, {{0, 1, 5, 4}, {4, 1, 5, 3}, {1, 2, 5, 3}}};
// 2D, one intersected edge (a vertex on the isocontour): ids 10, 11, 12 = 10 + edge ordinal,
// nodes [tri vertices, P].
MGPU_TABLE int kNhTri1Cells[3][2][3] = {
    {{0, 3, 2}, {3, 1, 2}}, {{0, 1, 3}, {0, 3, 2}}, {{0, 1, 3}, {1, 2, 3}}};
// 2D templates 1, 2, 3 (two intersected edges; ordinals: 0-2 tri vertices, 3 the intersection on
// the lower edge ordinal, 4 on the higher).
MGPU_TABLE int kNhTri2Cells[3][3][3] = {
    {{0, 3, 2}, {3, 1, 4}, {3, 4, 2}},
    {{0, 3, 4}, {3, 2, 4}, {3, 1, 2}},
    {{0, 1, 3}, {0, 3, 4}, {3, 2, 4}}};

// Level-set vertex classification tolerance (fn_PRM_GEN_Parameters.hpp:258-260).
constexpr double kIsocontourTolerance = 1e-12;
// IPs with detJ below this are skipped (cl_FEM_Geometry_Interpolator.cpp:28).
constexpr double kDetJLowerLimit = -1.0e-6;

}  // namespace mgpu
