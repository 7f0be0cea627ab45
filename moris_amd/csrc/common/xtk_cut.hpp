// xtk_cut.hpp -- per-cell restatement of XTK's conformal decomposition of one intersected
// background cell: regular subdivision (HEX8 -> 24 TET4, QUAD4 -> 4 TRI3) followed by the
// node-hierarchy templates on every intersected sub-cell.
//
//   regular subdivision   projects/XTK/src/cl_XTK_Regular_Subdivision_Interface.hpp:110-360
//   intersected edges     projects/XTK/src/cl_XTK_Node_Hierarchy_Interface.cpp:160-287
//   edge coordinate       projects/GEN/src/cl_GEN_Interpolation.hpp:107-132 (linear branch of
//                         cl_GEN_Level_Set_Geometry.cpp:438-450); |xi| <= 1 test
//                         projects/GEN/src/PDV/cl_GEN_Intersection_Node.cpp:78-81
//   intersection node     cl_GEN_Intersection_Node.cpp:33 (0.5(1-xi) a + 0.5(1+xi) b)
//   template selection    cl_XTK_Node_Hierarchy_Interface.cpp:672-733 (sort_nodes_2d/3d)
//   templates             cl_XTK_Node_Hierarchy_Template_Library.cpp:29-836
//
// The function is self-contained (fixed-size arrays, no allocation) so the same code can run in a
// GPU set-up kernel.  Vertex "ids" used to order intersection vertices are the global keys of the
// parent edges (lower global vertex id in the high 32 bits); the reference orders them by the
// global ids XTK hands out in creation order, which this key reproduces up to that ordering.
#pragma once
#include <cstdint>
#include <cmath>
#include "fem_tables.hpp"

namespace mgpu {

constexpr int kMaxSubTets = 24 * 6;
constexpr int kMaxSubTris = 4 * 3;

// Region of a level-set value relative to threshold 0 (Level_Set_Geometry::determine_geometric_region,
// cl_GEN_Level_Set_Geometry.cpp:511-528): 0 negative, 1 positive, 2 interface.
MGPU_HD inline int ls_region(double phi) {
    if (fabs(phi) <= kIsocontourTolerance) return 2;
    return phi < 0.0 ? 0 : 1;
}

// Intersection of the edge (a, b): a is the endpoint with the lower global id.
MGPU_HD inline double edge_local_coordinate(double phi_a, double phi_b) {
    return (2.0 * 0.0 - phi_b - phi_a) / (phi_b - phi_a);
}

template <int D>
MGPU_HD inline void edge_point(const double* xa, const double* xb, double xi, double* out) {
    const double s1 = 0.5 * (1.0 - xi);
    const double s2 = 0.5 * (1.0 + xi);
    for (int d = 0; d < D; ++d) out[d] = s1 * xa[d] + s2 * xb[d];
}

MGPU_HD inline int tet_shared_vertex(int e1, int e2) {
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            if (kTetEdges[e1][i] == kTetEdges[e2][j]) return kTetEdges[e1][i];
    return -1;
}

MGPU_HD inline int tet_other_vertex(int edge, int v) {
    return kTetEdges[edge][0] == v ? kTetEdges[edge][1] : kTetEdges[edge][0];
}

// Vertex of the 15-vertex regular-subdivision table of one hex cell.
struct HexSubdivVertex {
    double xi[3];     // parametric coordinate wrt the background (IP) cell
    double phi;       // level-set value
    uint32_t gid;     // global vertex id (lattice node, face node or cell-centre node)
};

// Status codes of the cut routines.
constexpr int kCutOk = 0;
constexpr int kCutBadTemplate = 2;       // permutation not in the template library

// Decompose one intersected hex cell and return the sub-tets of `phase` (0: phi<0, 1: phi>0).
// out_xi: [kMaxSubTets][4][3] parametric vertex coordinates; *n_out: number written.
// Sub-tets are emitted in regular-subdivision order; a cut sub-tet is replaced by its template
// cells in template order.  An edge with an endpoint on the isocontour is not intersected
// (Node_Hierarchy_Interface::determine_intersected_edges_and_make_requests, :214-222); the phase of
// a sub-cell is the region of its first parent vertex off the isocontour.
MGPU_HD inline int cut_hex8(const HexSubdivVertex v[15], int phase, double (*out_xi)[4][3],
                            int* n_out) {
    *n_out = 0;
    for (int t = 0; t < 24; ++t) {
        const int* tv = kHex24Tets[t];
        int n_int = 0;
        int edge_ord[6];
        uint64_t edge_key[6];
        double edge_xi[6][3];
        for (int e = 0; e < 6; ++e) {
            const HexSubdivVertex& p = v[tv[kTetEdges[e][0]]];
            const HexSubdivVertex& q = v[tv[kTetEdges[e][1]]];
            const int rp = ls_region(p.phi), rq = ls_region(q.phi);
            if (rp == 2 || rq == 2 || rp == rq) continue;
            const HexSubdivVertex& a = p.gid < q.gid ? p : q;
            const HexSubdivVertex& b = p.gid < q.gid ? q : p;
            const double loc = edge_local_coordinate(a.phi, b.phi);
            if (!(fabs(loc) <= 1.0)) continue;
            edge_ord[n_int] = e;
            edge_key[n_int] = ((uint64_t)a.gid << 32) | (uint64_t)b.gid;
            edge_point<3>(a.xi, b.xi, loc, edge_xi[n_int]);
            ++n_int;
        }
        if (n_int == 0) {
            int ph = -1;
            for (int k = 0; k < 4 && ph < 0; ++k)
                if (ls_region(v[tv[k]].phi) != 2) ph = ls_region(v[tv[k]].phi);
            if (ph == phase) {
                for (int k = 0; k < 4; ++k)
                    for (int d = 0; d < 3; ++d) out_xi[*n_out][k][d] = v[tv[k]].xi[d];
                ++(*n_out);
            }
            continue;
        }
        // ascending by intersection-vertex id (stable; keys are distinct)
        int order[4] = {0, 1, 2, 3};
        for (int i = 1; i < n_int; ++i) {
            int cur = order[i];
            int j = i - 1;
            while (j >= 0 && edge_key[order[j]] > edge_key[cur]) { order[j + 1] = order[j]; --j; }
            order[j + 1] = cur;
        }
        int node_ord[4] = {0, 1, 2, 3};
        int perm;
        const int* cells;
        int n_cells;
        if (n_int == 1) {
            perm = 10000 + edge_ord[0];
            cells = &kNhTet1Cells[edge_ord[0]][0][0];
            n_cells = 2;
        } else if (n_int == 2) {
            perm = 10000 + 100 * (edge_ord[order[0]] + 1) + 10 * (edge_ord[order[1]] + 1);
            int idx = -1;
            for (int i = 0; i < 24; ++i) if (kNhTet2Perm[i] == perm) idx = i;
            if (idx < 0) return kCutBadTemplate;
            cells = &kNhTet2Cells[idx][0][0];
            n_cells = 3;
        } else if (n_int == 3) {
            const int L = edge_ord[order[0]], M = edge_ord[order[1]], H = edge_ord[order[2]];
            const int n0 = tet_shared_vertex(L, M);
            node_ord[0] = n0;
            node_ord[1] = tet_other_vertex(L, n0);
            node_ord[2] = tet_other_vertex(M, n0);
            node_ord[3] = tet_other_vertex(H, n0);
            perm = L + 10 * M + 100 * H;
            int group = -1;
            for (int i = 0; i < 24; ++i) if (kNhPerm3[i][0] == perm) group = kNhPerm3[i][1];
            if (group < 0 || n0 < 0) return kCutBadTemplate;
            cells = &kNhTet3Cells[group][0][0];
            n_cells = 4;
        } else if (n_int == 4) {
            const int L = edge_ord[order[0]], ML = edge_ord[order[1]];
            const int MH = edge_ord[order[2]], H = edge_ord[order[3]];
            perm = L + 10 * ML + 100 * MH + 1000 * H;
            const int hl = tet_shared_vertex(L, H), hmh = tet_shared_vertex(MH, H);
            const int hml = tet_shared_vertex(ML, H), lml = tet_shared_vertex(L, ML);
            const int lmh = tet_shared_vertex(L, MH), mlmh = tet_shared_vertex(ML, MH);
            if (hl < 0) {
                node_ord[0] = hmh; node_ord[1] = hml; node_ord[2] = lmh; node_ord[3] = lml;
            } else if (hmh < 0) {
                node_ord[0] = hl; node_ord[1] = hml; node_ord[2] = lmh; node_ord[3] = mlmh;
            } else if (hml < 0) {
                node_ord[0] = hl; node_ord[1] = hmh; node_ord[2] = lml; node_ord[3] = mlmh;
            } else {
                return kCutBadTemplate;
            }
            int group = -1;
            for (int i = 0; i < 72; ++i) if (kNhPerm4[i][0] == perm) group = kNhPerm4[i][1];
            if (group < 0) return kCutBadTemplate;
            cells = &kNhTet4Cells[group][0][0];
            n_cells = 6;
        } else {
            return kCutBadTemplate;
        }
        for (int c = 0; c < n_cells; ++c) {
            const int* cv = cells + 4 * c;
            int ph = -1;
            for (int k = 0; k < 4 && ph < 0; ++k)
                if (cv[k] < 4) {
                    const int r = ls_region(v[tv[node_ord[cv[k]]]].phi);
                    if (r != 2) ph = r;
                }
            if (ph != phase) continue;
            for (int k = 0; k < 4; ++k) {
                const double* src = cv[k] < 4 ? v[tv[node_ord[cv[k]]]].xi : edge_xi[order[cv[k] - 4]];
                for (int d = 0; d < 3; ++d) out_xi[*n_out][k][d] = src[d];
            }
            ++(*n_out);
        }
    }
    return kCutOk;
}

// Vertex of the 5-vertex regular-subdivision table of one quad cell.
struct QuadSubdivVertex {
    double xi[2];
    double phi;
    uint32_t gid;
};

// Decompose one intersected quad cell; sub-tris of `phase` into out_xi[kMaxSubTris][3][2].
MGPU_HD inline int cut_quad4(const QuadSubdivVertex v[5], int phase, double (*out_xi)[3][2],
                             int* n_out) {
    *n_out = 0;
    for (int t = 0; t < 4; ++t) {
        const int* tv = kQuad4Tris[t];
        int n_int = 0;
        int edge_ord[3];
        double edge_xi[3][2];
        for (int e = 0; e < 3; ++e) {
            const QuadSubdivVertex& p = v[tv[kTriEdges[e][0]]];
            const QuadSubdivVertex& q = v[tv[kTriEdges[e][1]]];
            const int rp = ls_region(p.phi), rq = ls_region(q.phi);
            if (rp == 2 || rq == 2 || rp == rq) continue;
            const QuadSubdivVertex& a = p.gid < q.gid ? p : q;
            const QuadSubdivVertex& b = p.gid < q.gid ? q : p;
            const double loc = edge_local_coordinate(a.phi, b.phi);
            if (!(fabs(loc) <= 1.0)) continue;
            edge_ord[n_int] = e;
            edge_point<2>(a.xi, b.xi, loc, edge_xi[n_int]);
            ++n_int;
        }
        if (n_int == 0) {
            int ph = -1;
            for (int k = 0; k < 3 && ph < 0; ++k)
                if (ls_region(v[tv[k]].phi) != 2) ph = ls_region(v[tv[k]].phi);
            if (ph == phase) {
                for (int k = 0; k < 3; ++k)
                    for (int d = 0; d < 2; ++d) out_xi[*n_out][k][d] = v[tv[k]].xi[d];
                ++(*n_out);
            }
            continue;
        }
        // sort_nodes_2d orders intersections by edge ordinal; they are collected ascending
        const int* cells;
        int n_cells;
        if (n_int == 1) {
            cells = &kNhTri1Cells[edge_ord[0]][0][0];   // template 10 + ordinal
            n_cells = 2;
        } else if (n_int == 2) {
            const int perm = edge_ord[0] + edge_ord[1];   // template 1, 2 or 3
            if (perm < 1 || perm > 3) return kCutBadTemplate;
            cells = &kNhTri2Cells[perm - 1][0][0];
            n_cells = 3;
        } else {
            return kCutBadTemplate;
        }
        for (int c = 0; c < n_cells; ++c) {
            const int* cv = cells + 3 * c;
            int ph = -1;
            for (int k = 0; k < 3 && ph < 0; ++k)
                if (cv[k] < 3) {
                    const int r = ls_region(v[tv[cv[k]]].phi);
                    if (r != 2) ph = r;
                }
            if (ph != phase) continue;
            for (int k = 0; k < 3; ++k) {
                const double* src = cv[k] < 3 ? v[tv[cv[k]]].xi : edge_xi[cv[k] - 3];
                for (int d = 0; d < 2; ++d) out_xi[*n_out][k][d] = src[d];
            }
            ++(*n_out);
        }
    }
    return kCutOk;
}

}  // namespace mgpu
