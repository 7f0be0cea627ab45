// problem.cpp -- host set-up of one workload: level-0 HMR lattice, analytic GEN level set, XTK cut
// of every intersected cell, MSI adof numbering and the element-major set descriptor consumed by
// mgpu_set_create.  This is the input side of the hot path; nothing here is timed.
//
//   intersected background cells   Geometry_Engine::is_intersected, cl_GEN_Geometry_Engine.cpp:278-300
//   cut of one cell                common/xtk_cut.hpp
//   phase of a cell                material = phase 1 (phi > 0) by default; blocks *_n_p1 (uncut)
//                                  and *_c_p1 (cut), UT_MDL_XTK_HMR_Linear_Struc_2D.cpp:355-359
//   adof ids                       Dof_Manager::set_owned_adofs_ids_by_type,
//                                  cl_MSI_Dof_Manager.cpp:703-745: by dof type, then ascending
//                                  basis index over the bases some element uses
//   T rows                         p = 1: one entry of weight 1 per Lagrange node; p = 2: the
//                                  lattice's quadratic B-splines at the node (HMR level 0,
//                                  T_Matrix_Base::evaluate, cl_HMR_T_Matrix_Base.cpp:306-457)
//   HEX27 IP cells                 lagrange_order 2: 27 nodes per cell on the half-step grid
//                                  (cl_MTK_Interpolation_Function_Lagrange_Hex27.hpp node order);
//                                  the cut is that of the cell's box, as for HEX8
#include <algorithm>
#include <memory>
#include <numeric>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../common/fem_tables.hpp"
#include "../common/levelset.hpp"
#include "../common/xtk_cut.hpp"
#include "enrich.hpp"
#include "internal.hpp"

using namespace mgpu;

struct mgpu_problem {
    mgpu_problem_spec spec{};
    std::vector<double> spheres;
    LevelSet ls;
    int dim = 3;
    int n[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0};
    double h[3] = {1, 1, 1};
    int nb = 8;      // vertices of the IP cell's box (the cut geometry)
    int nb_ip = 8;   // IP-cell bases: nb, or 27 for HEX27
    int64_t n_cells = 0, n_nodes = 0, n_faces = 0;
    int64_t n_cut = 0, n_void = 0;

    std::vector<double> node_phi;
    std::vector<int32_t> cell_elem;      // first element of a level-0 cell, -1 if none
    std::vector<int32_t> cell_nelem;     // elements of a level-0 cell (subphases when enriched)
    std::vector<uint8_t> elem_corner;    // enrichment: strictly-material corners per element
    std::vector<int32_t> ghost;          // ghost facets: (leader, follower, axis) triples
    std::vector<int32_t> couplings;      // spec.ghost: (leader, follower) pairs for the descriptor

    std::vector<int32_t> elem_cell;
    std::vector<int8_t> elem_level;      // 0, or 1 for a THB level-1 cell
    std::vector<int64_t> elem_lcell;     // cell index on the element's level lattice
    std::vector<char> refined;           // THB: level-0 cells split into 8 level-1 cells
    std::vector<double> ip_box, material, vertex_phi;
    std::vector<int32_t> t_ptr, t_adof;
    std::vector<double> t_weight;
    std::vector<int32_t> ig_ptr;
    std::vector<double> ig_xi;
    int32_t n_adofs_per_type = 0;
};

namespace {

int64_t node_index(const mgpu_problem& p, int i, int j, int k) {
    return (int64_t)i + (int64_t)(p.n[0] + 1) * ((int64_t)j + (int64_t)(p.n[1] + 1) * k);
}

void node_coords(const mgpu_problem& p, int i, int j, int k, double* x) {
    const int ijk[3] = {i, j, k};
    for (int d = 0; d < p.dim; ++d) x[d] = p.lo[d] + (double)ijk[d] * p.h[d];
    if (p.dim == 2) x[2] = 0.0;
}

// Trilinear / bilinear Lagrange basis at xi (vertex order of kHex8XiHat / kQuad4XiHat).
void lagrange_N(int dim, const double* xi, double* N) {
    if (dim == 3) {
        for (int k = 0; k < 8; ++k)
            N[k] = 0.125 * (1.0 + xi[0] * kHex8XiHat[k][0]) * (1.0 + xi[1] * kHex8XiHat[k][1]) *
                   (1.0 + xi[2] * kHex8XiHat[k][2]);
    } else {
        for (int k = 0; k < 4; ++k)
            N[k] = 0.25 * (1.0 + xi[0] * kQuad4XiHat[k][0]) * (1.0 + xi[1] * kQuad4XiHat[k][1]);
    }
}

// Physical coordinates of the point with parametric coordinate xi in cell (i,j,k): N(xi) . X.
void cell_point(const mgpu_problem& p, int i, int j, int k, const double* xi, double* x) {
    double N[8];
    lagrange_N(p.dim, xi, N);
    const int nb = p.dim == 3 ? 8 : 4;
    for (int d = 0; d < 3; ++d) x[d] = 0.0;
    for (int v = 0; v < nb; ++v) {
        double xv[3];
        if (p.dim == 3)
            node_coords(p, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2], xv);
        else
            node_coords(p, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0, xv);
        for (int d = 0; d < p.dim; ++d) x[d] += N[v] * xv[d];
    }
}

// Background cell intersected: its lattice vertices do not all lie in one region
// (Geometry_Engine::is_intersected, cl_GEN_Geometry_Engine.cpp:278-300).
bool cell_intersected(const mgpu_problem& p, int i, int j, int k) {
    const int nb = p.dim == 3 ? 8 : 4;
    int r0 = -1;
    for (int v = 0; v < nb; ++v) {
        const int64_t nid = p.dim == 3 ? node_index(p, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2])
                                       : node_index(p, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0);
        const int r = ls_region(p.node_phi[(size_t)nid]);
        if (v == 0) r0 = r;
        else if (r != r0) return true;
    }
    return false;
}

struct CellOut {
    int32_t cell;
    bool cut;
    std::vector<double> xi;   // sub-cell vertex parametric coords
    int n_sub;
    // enrichment: sub-cells per subphase (xi grouped by subphase) and each subphase's
    // strictly-material corners
    std::vector<int32_t> sp_count;
    std::vector<uint8_t> sp_corner;
};

// Process a contiguous range of cells; material cells appended in order.
int process_cells(const mgpu_problem& p, int64_t c0, int64_t c1, std::vector<CellOut>& out,
                  int64_t& n_void, std::string& err) {
    const int phase = p.spec.material_phase;
    static thread_local double sub3[kMaxSubTets][4][3];
    static thread_local double sub2[kMaxSubTris][3][2];
    for (int64_t c = c0; c < c1; ++c) {
        const int i = (int)(c % p.n[0]);
        const int j = (int)((c / p.n[0]) % p.n[1]);
        const int k = p.dim == 3 ? (int)(c / ((int64_t)p.n[0] * p.n[1])) : 0;
        const int nb = p.nb;
        int regions[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        bool any_diff = false;
        for (int v = 0; v < nb; ++v) {
            const int64_t nid = p.dim == 3
                                    ? node_index(p, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2])
                                    : node_index(p, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0);
            regions[v] = ls_region(p.node_phi[nid]);
            if (regions[v] != regions[0]) any_diff = true;
        }
        if (!any_diff) {
            if (regions[0] == phase) out.push_back(CellOut{(int32_t)c, false, {}, 0});
            else ++n_void;
            continue;
        }
        int n_sub = 0;
        int rc;
        CellOut co{(int32_t)c, true, {}, 0};
        if (p.dim == 3) {
            HexSubdivVertex hv[15];
            for (int v = 0; v < 8; ++v) {
                for (int d = 0; d < 3; ++d) hv[v].xi[d] = kHex8XiHat[v][d];
                const int64_t nid = node_index(p, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2]);
                hv[v].phi = p.node_phi[nid];
                hv[v].gid = (uint32_t)nid;
            }
            // face nodes: coordinates interpolated in the lowest-index intersected cell sharing
            // the face (the cell that first requests the node when cells are decomposed in order)
            const int64_t nxf = (int64_t)(p.n[0] + 1) * p.n[1] * p.n[2];
            const int64_t nyf = (int64_t)p.n[0] * (p.n[1] + 1) * p.n[2];
            for (int f = 0; f < 6; ++f) {
                HexSubdivVertex& fv = hv[8 + f];
                for (int d = 0; d < 3; ++d) fv.xi[d] = kHexFaceXi[f][d];
                int pi = i, pj = j, pk = k, pf = f;
                int64_t fid;
                switch (f) {
                    case 0: if (j > 0 && cell_intersected(p, i, j - 1, k)) { pj = j - 1; pf = 2; }
                            fid = nxf + i + (int64_t)p.n[0] * (j + (int64_t)(p.n[1] + 1) * k); break;
                    case 1: fid = (i + 1) + (int64_t)(p.n[0] + 1) * (j + (int64_t)p.n[1] * k); break;
                    case 2: fid = nxf + i + (int64_t)p.n[0] * ((j + 1) + (int64_t)(p.n[1] + 1) * k); break;
                    case 3: if (i > 0 && cell_intersected(p, i - 1, j, k)) { pi = i - 1; pf = 1; }
                            fid = i + (int64_t)(p.n[0] + 1) * (j + (int64_t)p.n[1] * k); break;
                    case 4: if (k > 0 && cell_intersected(p, i, j, k - 1)) { pk = k - 1; pf = 5; }
                            fid = nxf + nyf + i + (int64_t)p.n[0] * (j + (int64_t)p.n[1] * k); break;
                    default: fid = nxf + nyf + i + (int64_t)p.n[0] * (j + (int64_t)p.n[1] * (k + 1)); break;
                }
                double x[3];
                cell_point(p, pi, pj, pk, kHexFaceXi[pf], x);
                fv.phi = p.ls(x);
                fv.gid = (uint32_t)(p.n_nodes + fid);
            }
            HexSubdivVertex& cv = hv[14];
            const double zero[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < 3; ++d) cv.xi[d] = 0.0;
            double x[3];
            cell_point(p, i, j, k, zero, x);
            cv.phi = p.ls(x);
            cv.gid = (uint32_t)(p.n_nodes + p.n_faces + c);
            rc = cut_hex8(hv, phase, sub3, &n_sub);
            if (rc == kCutOk) co.xi.assign(&sub3[0][0][0], &sub3[0][0][0] + (size_t)n_sub * 12);
        } else {
            QuadSubdivVertex qv[5];
            for (int v = 0; v < 4; ++v) {
                for (int d = 0; d < 2; ++d) qv[v].xi[d] = kQuad4XiHat[v][d];
                const int64_t nid = node_index(p, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0);
                qv[v].phi = p.node_phi[nid];
                qv[v].gid = (uint32_t)nid;
            }
            const double zero[2] = {0.0, 0.0};
            double x[3];
            cell_point(p, i, j, 0, zero, x);
            qv[4].xi[0] = qv[4].xi[1] = 0.0;
            qv[4].phi = p.ls(x);
            qv[4].gid = (uint32_t)(p.n_nodes + c);
            rc = cut_quad4(qv, phase, sub2, &n_sub);
            if (rc == kCutOk) co.xi.assign(&sub2[0][0][0], &sub2[0][0][0] + (size_t)n_sub * 6);
        }
        if (rc != kCutOk) {
            err = "node-hierarchy template not found in cell " + std::to_string(c);
            return MGPU_ERR_GEOMETRY;
        }
        if (n_sub == 0) { ++n_void; continue; }
        co.n_sub = n_sub;
        if (p.spec.enrich) split_subphases(p.dim, n_sub, co.xi.data(), regions, phase, co.sp_count, co.sp_corner);
        out.push_back(std::move(co));
    }
    return MGPU_OK;
}

// Ghost facets of a level-0 problem (XTK's ghost stabilisation facets between a cut cell's material
// and its neighbours): faces between two material elements, at least one of them cut, whose
// shared face holds a lattice node strictly in the material (so that the material of both sides
// meets there); leader the lower cell, elements ascending, then axes ascending.
void ghost_facets(const mgpu_problem& p, std::vector<int32_t>& out) {
    out.clear();
    const int dim = p.dim;
    const int phase = p.spec.material_phase;
    const size_t n_el = p.elem_cell.size();
    for (size_t e = 0; e < n_el; ++e) {
        const int64_t c = p.elem_cell[e];
        const int ijk[3] = {(int)(c % p.n[0]), (int)((c / p.n[0]) % p.n[1]),
                            dim == 3 ? (int)(c / ((int64_t)p.n[0] * p.n[1])) : 0};
        const bool cut_e = p.ig_ptr[e + 1] > p.ig_ptr[e];
        for (int d = 0; d < dim; ++d) {
            if (ijk[d] + 1 >= p.n[d]) continue;
            const int64_t cn = c + (d == 0 ? 1 : d == 1 ? p.n[0] : (int64_t)p.n[0] * p.n[1]);
            const int32_t en = p.cell_elem[(size_t)cn];
            if (en < 0) continue;
            const bool cut_n = p.ig_ptr[(size_t)en + 1] > p.ig_ptr[(size_t)en];
            if (!cut_e && !cut_n) continue;
            bool material = false;
            const int nv = dim == 3 ? 8 : 4;
            for (int v = 0; v < nv && !material; ++v) {
                const int* off = dim == 3 ? kHex8Off[v] : kQuad4Off[v];
                if (off[d] != 1) continue;
                const int64_t nid = dim == 3 ? node_index(p, ijk[0] + off[0], ijk[1] + off[1], ijk[2] + off[2])
                                             : node_index(p, ijk[0] + off[0], ijk[1] + off[1], 0);
                material = ls_region(p.node_phi[(size_t)nid]) == phase;
            }
            if (!material) continue;
            out.push_back((int32_t)e);
            out.push_back(en);
            out.push_back(d);
        }
    }
}

int n_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 4;
    return (int)std::min(t, 32u);
}

}  // namespace

extern "C" mgpu_status mgpu_problem_build(const mgpu_problem_spec* spec, mgpu_problem** out) {
    MGPU_GUARD_BEGIN
    if (!spec || !out) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (spec->dim != 2 && spec->dim != 3) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "dim must be 2 or 3");
    if (spec->iwg != MGPU_IWG_SPATIALDIFF_BULK && spec->iwg != MGPU_IWG_STRUC_LINEAR_BULK)
        return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "iwg not supported");
    for (int d = 0; d < spec->dim; ++d)
        if (spec->n[d] < 1 || !(spec->hi[d] > spec->lo[d]))
            return mgpu::set_error(MGPU_ERR_INVALID_ARG, "bad grid extent");
    if (spec->ls_kind == MGPU_LS_SPHERES && (spec->n_spheres < 1 || !spec->spheres))
        return mgpu::set_error(MGPU_ERR_INVALID_ARG, "sphere level set without spheres");

    auto p = std::make_unique<mgpu_problem>();
    p->spec = *spec;
    p->dim = spec->dim;
    for (int d = 0; d < 3; ++d) {
        p->n[d] = d < p->dim ? spec->n[d] : 1;
        p->lo[d] = d < p->dim ? spec->lo[d] : 0.0;
        p->h[d] = d < p->dim ? (spec->hi[d] - spec->lo[d]) / (double)spec->n[d] : 1.0;
    }
    if (spec->ls_kind == MGPU_LS_SPHERES) {
        p->spheres.assign(spec->spheres, spec->spheres + 4 * (size_t)spec->n_spheres);
        p->spec.spheres = p->spheres.data();
    } else {
        p->spec.spheres = nullptr;
    }
    p->ls.kind = spec->ls_kind;
    p->ls.dim = p->dim;
    p->ls.n_spheres = spec->n_spheres;
    p->ls.spheres = p->spheres.data();
    p->ls.gyroid_scale = spec->gyroid_scale;
    for (int d = 0; d < 4; ++d) p->ls.plane[d] = spec->plane[d];
    p->nb = p->dim == 3 ? 8 : 4;
    const int lag = spec->lagrange_order <= 1 ? 1 : spec->lagrange_order;
    if (lag > 2 || (lag == 2 && p->dim != 3))
        return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "lagrange_order 2 (HEX27) is implemented in 3D only");
    p->nb_ip = lag == 2 ? 27 : p->nb;
    p->n_cells = (int64_t)p->n[0] * p->n[1] * p->n[2];
    p->n_nodes = p->dim == 3 ? (int64_t)(p->n[0] + 1) * (p->n[1] + 1) * (p->n[2] + 1)
                             : (int64_t)(p->n[0] + 1) * (p->n[1] + 1);
    p->n_faces = p->dim == 3 ? (int64_t)(p->n[0] + 1) * p->n[1] * p->n[2] +
                                   (int64_t)p->n[0] * (p->n[1] + 1) * p->n[2] +
                                   (int64_t)p->n[0] * p->n[1] * (p->n[2] + 1)
                             : 0;
    if (p->n_nodes + p->n_faces + p->n_cells >= (int64_t)UINT32_MAX || p->n_cells >= INT32_MAX)
        return mgpu::set_error(MGPU_ERR_INVALID_ARG, "grid too large for 32-bit ids");

    const int T = n_threads();
    // 1. level set at lattice nodes
    p->node_phi.resize((size_t)p->n_nodes);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t]() {
                const int64_t a = p->n_nodes * t / T, b = p->n_nodes * (t + 1) / T;
                for (int64_t nid = a; nid < b; ++nid) {
                    const int i = (int)(nid % (p->n[0] + 1));
                    const int j = (int)((nid / (p->n[0] + 1)) % (p->n[1] + 1));
                    const int k = p->dim == 3 ? (int)(nid / ((int64_t)(p->n[0] + 1) * (p->n[1] + 1))) : 0;
                    double x[3];
                    node_coords(*p, i, j, k, x);
                    p->node_phi[(size_t)nid] = p->ls(x);
                }
            });
        for (auto& t : th) t.join();
    }
    // THB: one refinement level on the intersected level-0 cells, cut on the level-1 lattice
    if (spec->refine_levels < 0 || spec->refine_levels > 1)
        return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "refine_levels must be 0 or 1");
    const bool thb = spec->refine_levels == 1;
    if (thb && (p->dim != 3 || spec->bspline_order != 2))
        return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "THB refinement is implemented for 3D sets with bspline_order 2");
    std::unique_ptr<mgpu_problem> fine;
    if (thb) {
        p->refined.assign((size_t)p->n_cells, 0);
        for (int64_t c = 0; c < p->n_cells; ++c)
            p->refined[(size_t)c] = cell_intersected(*p, (int)(c % p->n[0]), (int)((c / p->n[0]) % p->n[1]),
                                                     (int)(c / ((int64_t)p->n[0] * p->n[1])));
        fine = std::make_unique<mgpu_problem>();
        mgpu_problem& f = *fine;
        f.spec = p->spec;
        f.ls = p->ls;   // spheres stay owned by p
        f.dim = 3;
        f.nb = p->nb;
        f.nb_ip = p->nb_ip;
        for (int d = 0; d < 3; ++d) {
            f.n[d] = 2 * p->n[d];
            f.lo[d] = p->lo[d];
            f.h[d] = 0.5 * p->h[d];
        }
        f.n_cells = (int64_t)f.n[0] * f.n[1] * f.n[2];
        f.n_nodes = (int64_t)(f.n[0] + 1) * (f.n[1] + 1) * (f.n[2] + 1);
        f.n_faces = (int64_t)(f.n[0] + 1) * f.n[1] * f.n[2] + (int64_t)f.n[0] * (f.n[1] + 1) * f.n[2] +
                    (int64_t)f.n[0] * f.n[1] * (f.n[2] + 1);
        if (f.n_nodes + f.n_faces + f.n_cells >= (int64_t)UINT32_MAX)
            return mgpu::set_error(MGPU_ERR_INVALID_ARG, "level-1 grid too large for 32-bit ids");
        f.node_phi.resize((size_t)f.n_nodes);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t]() {
                const int64_t a = f.n_nodes * t / T, b = f.n_nodes * (t + 1) / T;
                for (int64_t nid = a; nid < b; ++nid) {
                    double x[3];
                    node_coords(f, (int)(nid % (f.n[0] + 1)), (int)((nid / (f.n[0] + 1)) % (f.n[1] + 1)),
                                (int)(nid / ((int64_t)(f.n[0] + 1) * (f.n[1] + 1))), x);
                    f.node_phi[(size_t)nid] = f.ls(x);
                }
            });
        for (auto& t : th) t.join();
    }
    // 2. cut every leaf cell (chunks of level-0 cells in lattice order; a refined cell contributes
    // its 8 level-1 cells, x fastest)
    struct LeafOut {
        int level;
        CellOut co;
    };
    std::vector<std::vector<LeafOut>> parts(T);
    std::vector<int64_t> voids(T, 0);
    std::vector<int> rcs(T, MGPU_OK);
    std::vector<std::string> errs(T);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t]() {
                const int64_t a = p->n_cells * t / T, b = p->n_cells * (t + 1) / T;
                std::vector<CellOut> tmp;
                if (!thb) {
                    rcs[t] = process_cells(*p, a, b, tmp, voids[t], errs[t]);
                    for (auto& co : tmp) parts[t].push_back(LeafOut{0, std::move(co)});
                    return;
                }
                for (int64_t c = a; c < b && rcs[t] == MGPU_OK; ++c) {
                    tmp.clear();
                    if (!p->refined[(size_t)c]) {
                        rcs[t] = process_cells(*p, c, c + 1, tmp, voids[t], errs[t]);
                        for (auto& co : tmp) parts[t].push_back(LeafOut{0, std::move(co)});
                        continue;
                    }
                    const int i = (int)(c % p->n[0]), j = (int)((c / p->n[0]) % p->n[1]),
                              k = (int)(c / ((int64_t)p->n[0] * p->n[1]));
                    for (int cz = 0; cz < 2 && rcs[t] == MGPU_OK; ++cz)
                        for (int cy = 0; cy < 2 && rcs[t] == MGPU_OK; ++cy)
                            for (int cx = 0; cx < 2 && rcs[t] == MGPU_OK; ++cx) {
                                const int64_t fc = (2 * i + cx) + (int64_t)fine->n[0] *
                                                   ((2 * j + cy) + (int64_t)fine->n[1] * (2 * k + cz));
                                rcs[t] = process_cells(*fine, fc, fc + 1, tmp, voids[t], errs[t]);
                            }
                    for (auto& co : tmp) parts[t].push_back(LeafOut{1, std::move(co)});
                }
            });
        for (auto& t : th) t.join();
    }
    for (int t = 0; t < T; ++t)
        if (rcs[t] != MGPU_OK) return mgpu::set_error(rcs[t], errs[t]);

    // 3. element arrays
    // with enrichment a cut cell contributes one element per material subphase
    const bool enrich = spec->enrich != 0;
    if (enrich && thb) return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "enrichment is implemented for level-0 lattices");
    auto n_elems_of = [&](const CellOut& co) { return enrich && co.cut ? (int64_t)co.sp_count.size() : (int64_t)1; };
    int64_t n_el = 0, n_ig = 0;
    for (int t = 0; t < T; ++t) {
        p->n_void += voids[t];
        for (const auto& lo : parts[t]) {
            n_el += n_elems_of(lo.co);
            n_ig += lo.co.n_sub;
        }
    }
    if (n_el >= INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many elements");
    const int d2 = 2 * p->dim;
    const int nv_ig = p->dim == 3 ? 4 : 3;
    p->cell_elem.assign((size_t)p->n_cells, -1);
    p->cell_nelem.assign((size_t)p->n_cells, 0);
    if (enrich) p->elem_corner.resize((size_t)n_el);
    p->elem_cell.resize((size_t)n_el);
    p->elem_level.resize((size_t)n_el);
    p->elem_lcell.resize((size_t)n_el);
    p->ip_box.resize((size_t)n_el * d2);
    p->vertex_phi.resize((size_t)n_el * p->nb_ip);
    p->ig_ptr.resize((size_t)n_el + 1);
    p->ig_xi.resize((size_t)n_ig * nv_ig * p->dim);
    const bool diff = spec->iwg == MGPU_IWG_SPATIALDIFF_BULK;
    p->material.resize((size_t)n_el * (diff ? 1 : 2));
    int64_t e = 0, ig = 0;
    for (int t = 0; t < T; ++t) {
        for (const auto& lo : parts[t]) {
            const CellOut& co = lo.co;
            const mgpu_problem& L = lo.level == 0 ? *p : *fine;
            const int64_t c = co.cell;
            const int i = (int)(c % L.n[0]);
            const int j = (int)((c / L.n[0]) % L.n[1]);
            const int k = p->dim == 3 ? (int)(c / ((int64_t)L.n[0] * L.n[1])) : 0;
            const int ijk[3] = {i, j, k};
            // level-0 cell of the element (its parent for a level-1 cell)
            const int64_t c0 = lo.level == 0 ? c : (i / 2) + (int64_t)p->n[0] * ((j / 2) + (int64_t)p->n[1] * (k / 2));
            if (lo.level == 0) {
                p->cell_elem[(size_t)c] = (int32_t)e;
                p->cell_nelem[(size_t)c] = (int32_t)n_elems_of(co);
            }
            const int64_t n_sp = n_elems_of(co);
            for (int64_t q = 0; q < n_sp; ++q, ++e) {
                if (enrich) p->elem_corner[(size_t)e] = co.cut ? co.sp_corner[(size_t)q] : (uint8_t)((1u << (1 << p->dim)) - 1u);
                p->elem_cell[(size_t)e] = (int32_t)c0;
                p->elem_level[(size_t)e] = (int8_t)lo.level;
                p->elem_lcell[(size_t)e] = c;
                for (int d = 0; d < p->dim; ++d) {
                    p->ip_box[(size_t)e * d2 + d] = L.lo[d] + (double)ijk[d] * L.h[d];
                    p->ip_box[(size_t)e * d2 + p->dim + d] = L.h[d];
                }
                if (p->nb_ip == 27) {
                    // level set at the 27 HEX27 nodes (the cut itself uses the 8 corner values)
                    for (int v = 0; v < 27; ++v) {
                        double x[3];
                        for (int d = 0; d < 3; ++d) x[d] = L.lo[d] + ((double)ijk[d] + 0.5 * kHex27Half[v][d]) * L.h[d];
                        p->vertex_phi[(size_t)e * 27 + v] = p->ls(x);
                    }
                } else {
                    for (int v = 0; v < p->nb; ++v) {
                        const int64_t nid = p->dim == 3 ? node_index(L, i + kHex8Off[v][0], j + kHex8Off[v][1], k + kHex8Off[v][2])
                                                        : node_index(L, i + kQuad4Off[v][0], j + kQuad4Off[v][1], 0);
                        p->vertex_phi[(size_t)e * p->nb + v] = L.node_phi[(size_t)nid];
                    }
                }
                if (diff) {
                    // per-element conductivity seeded by the cell id (level-1 cells: n_cells + fine id)
                    const uint64_t sid = lo.level == 0 ? (uint64_t)c : (uint64_t)(p->n_cells + c);
                    p->material[(size_t)e] = spec->k_sigma == 0.0
                                                 ? std::exp(spec->k_mu)
                                                 : std::exp(spec->k_mu + spec->k_sigma * normal_at(spec->seed, sid));
                } else {
                    p->material[(size_t)e * 2] = spec->E;
                    p->material[(size_t)e * 2 + 1] = spec->nu;
                }
                p->ig_ptr[(size_t)e] = (int32_t)ig;
                if (co.cut) {
                    // the subphase's sub-cells (all of them without enrichment)
                    const int64_t first = enrich ? std::accumulate(co.sp_count.begin(), co.sp_count.begin() + q, (int64_t)0) : 0;
                    const int64_t cnt = enrich ? co.sp_count[(size_t)q] : co.n_sub;
                    std::memcpy(&p->ig_xi[(size_t)ig * nv_ig * p->dim], co.xi.data() + (size_t)first * nv_ig * p->dim,
                                (size_t)cnt * nv_ig * p->dim * sizeof(double));
                    ig += cnt;
                    ++p->n_cut;
                }
            }
        }
    }
    p->ig_ptr[(size_t)n_el] = (int32_t)ig;

    // 4./5. adof numbering and T rows
    const int order = spec->bspline_order <= 1 ? 1 : spec->bspline_order;
    if (order > 2) return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "bspline_order must be 1 or 2");
    // position of basis v of the element in cell (i, j, k) on the half-step grid of the lattice
    auto basis_half = [&](int i, int j, int k, int v, int* g) {
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
            if (d >= p->dim) { g[d] = 0; continue; }
            const int off = p->nb_ip == 27 ? kHex27Half[v][d] : 2 * (p->dim == 3 ? kHex8Off[v][d] : kQuad4Off[v][d]);
            g[d] = 2 * ijk[d] + off;
        }
    };
    auto elem_ijk = [&](int64_t ee, int* ijk) {
        const int64_t c = p->elem_cell[(size_t)ee];
        ijk[0] = (int)(c % p->n[0]);
        ijk[1] = (int)((c / p->n[0]) % p->n[1]);
        ijk[2] = p->dim == 3 ? (int)(c / ((int64_t)p->n[0] * p->n[1])) : 0;
    };
    const size_t n_rows = (size_t)n_el * p->nb_ip;
    // enrichment: the enriched adof of every (element, basis of its cell)
    std::vector<int32_t> eadof;
    int32_t n_eadofs = 0;
    const int escheme = order == 2 ? kEnrichBspline2 : p->nb_ip == 27 ? kEnrichLagrange27 : kEnrichLagrange1;
    const int enloc = enrich_local_count(p->dim, escheme);
    if (enrich) {
        EnrichInput ein{};
        ein.dim = p->dim;
        ein.scheme = escheme;
        for (int d = 0; d < 3; ++d) ein.n[d] = p->n[d];
        ein.cell_first = p->cell_elem.data();
        ein.cell_count = p->cell_nelem.data();
        ein.elem_corner = p->elem_corner.data();
        ein.n_el = n_el;
        std::string eerr;
        if (!enrich_numbering(ein, eadof, &n_eadofs, &eerr)) return mgpu::set_error(MGPU_ERR_UNSUPPORTED, eerr);
    }
    // enriched adof of the basis at per-axis offsets off[] from the element's cell
    auto enriched = [&](int64_t ee, const int* off) {
        return eadof[(size_t)ee * enloc + enrich_local_index(p->dim, escheme, off)];
    };
    if (thb) {
        // Truncated hierarchical quadratic B-splines, two levels (T_Matrix_Base::
        // calculate_truncated_t_matrix, cl_HMR_T_Matrix_Base.cpp:127-258, applied at the Lagrange
        // nodes as in evaluate, :306-457).  Level-0 spline s (n + 2 per axis, support cells s - 2 .. s)
        // is active unless every cell of its support is refined; level-1 spline t (2n + 2 per axis,
        // support level-1 cells t - 2 .. t) is active if every cell of its support lies in a refined
        // level-0 cell.  On a level-1 cell an active level-0 spline is replaced by its truncation:
        // the sum of its inactive level-1 children t = 2 s - 2 + k (k = 0..3 per axis, two-scale
        // weights 1/4, 3/4, 3/4, 1/4).  On an unrefined cell every level-0 spline is active and no
        // level-1 spline is non-zero.  Adofs: used level-0 splines in ascending index, then used
        // level-1 splines in ascending index; each row sorted by adof.
        const int64_t n0[3] = {p->n[0], p->n[1], p->n[2]};
        const int64_t s0[3] = {n0[0] + 2, n0[1] + 2, n0[2] + 2};
        const int64_t s1[3] = {2 * n0[0] + 2, 2 * n0[1] + 2, 2 * n0[2] + 2};
        if (s1[0] * s1[1] * s1[2] >= INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many splines");
        auto refined0 = [&](int64_t i, int64_t j, int64_t k) { return p->refined[(size_t)(i + n0[0] * (j + n0[1] * k))] != 0; };
        std::vector<char> act0((size_t)(s0[0] * s0[1] * s0[2]), 0), act1((size_t)(s1[0] * s1[1] * s1[2]), 0);
        for (int64_t c = 0; c < s0[2]; ++c)
            for (int64_t b = 0; b < s0[1]; ++b)
                for (int64_t a = 0; a < s0[0]; ++a) {
                    bool any_unrefined = false;
                    for (int64_t k = std::max<int64_t>(0, c - 2); k <= std::min(n0[2] - 1, c) && !any_unrefined; ++k)
                        for (int64_t j = std::max<int64_t>(0, b - 2); j <= std::min(n0[1] - 1, b) && !any_unrefined; ++j)
                            for (int64_t i = std::max<int64_t>(0, a - 2); i <= std::min(n0[0] - 1, a); ++i)
                                if (!refined0(i, j, k)) { any_unrefined = true; break; }
                    act0[(size_t)(a + s0[0] * (b + s0[1] * c))] = any_unrefined;
                }
        for (int64_t c = 0; c < s1[2]; ++c)
            for (int64_t b = 0; b < s1[1]; ++b)
                for (int64_t a = 0; a < s1[0]; ++a) {
                    bool all_refined = true;
                    for (int64_t k = std::max<int64_t>(0, c - 2); k <= std::min(2 * n0[2] - 1, c) && all_refined; ++k)
                        for (int64_t j = std::max<int64_t>(0, b - 2); j <= std::min(2 * n0[1] - 1, b) && all_refined; ++j)
                            for (int64_t i = std::max<int64_t>(0, a - 2); i <= std::min(2 * n0[0] - 1, a); ++i)
                                if (!refined0(i / 2, j / 2, k / 2)) { all_refined = false; break; }
                    act1[(size_t)(a + s1[0] * (b + s1[1] * c))] = all_refined;
                }
        // per-axis splines of a lattice with its half-step coordinate g: (index, value)
        struct AxisRow { int n; int64_t s[3]; double w[3]; };
        auto axis_row = [](int64_t g, AxisRow& r) {
            const int64_t m = g / 2;
            if ((g & 1) == 0) {
                r.n = 2; r.s[0] = m; r.s[1] = m + 1; r.w[0] = r.w[1] = 0.5;
            } else {
                r.n = 3; r.s[0] = m; r.s[1] = m + 1; r.s[2] = m + 2; r.w[0] = 0.125; r.w[1] = 0.75; r.w[2] = 0.125;
            }
        };
        const double two_scale[4] = {0.25, 0.75, 0.75, 0.25};
        // row of basis v of element ee: (level, spline index, weight), |weight| > 1e-12
        struct Ent { int level; int64_t sp; double w; };
        auto row = [&](int64_t ee, int v, std::vector<Ent>& out) {
            out.clear();
            const int lvl = p->elem_level[(size_t)ee];
            const int64_t nl[3] = {lvl ? 2 * n0[0] : n0[0], lvl ? 2 * n0[1] : n0[1], lvl ? 2 * n0[2] : n0[2]};
            const int64_t c = p->elem_lcell[(size_t)ee];
            const int64_t ijk[3] = {c % nl[0], (c / nl[0]) % nl[1], c / (nl[0] * nl[1])};
            AxisRow ax[3];
            for (int d = 0; d < 3; ++d) {
                const int off = p->nb_ip == 27 ? kHex27Half[v][d] : 2 * kHex8Off[v][d];
                axis_row(2 * ijk[d] + off, ax[d]);
            }
            if (lvl == 0) {
                for (int zc = 0; zc < ax[2].n; ++zc)
                    for (int yc = 0; yc < ax[1].n; ++yc)
                        for (int xc = 0; xc < ax[0].n; ++xc)
                            out.push_back({0, ax[0].s[xc] + s0[0] * (ax[1].s[yc] + s0[1] * ax[2].s[zc]),
                                           ax[0].w[xc] * ax[1].w[yc] * ax[2].w[zc]});
                return;
            }
            // level-1 splines at the node
            for (int zc = 0; zc < ax[2].n; ++zc)
                for (int yc = 0; yc < ax[1].n; ++yc)
                    for (int xc = 0; xc < ax[0].n; ++xc) {
                        const int64_t t = ax[0].s[xc] + s1[0] * (ax[1].s[yc] + s1[1] * ax[2].s[zc]);
                        if (act1[(size_t)t]) out.push_back({1, t, ax[0].w[xc] * ax[1].w[yc] * ax[2].w[zc]});
                    }
            // truncated active level-0 splines of the parent cell
            const int64_t P[3] = {ijk[0] / 2, ijk[1] / 2, ijk[2] / 2};
            for (int64_t sc = P[2]; sc <= P[2] + 2; ++sc)
                for (int64_t sb = P[1]; sb <= P[1] + 2; ++sb)
                    for (int64_t sa = P[0]; sa <= P[0] + 2; ++sa) {
                        const int64_t s = sa + s0[0] * (sb + s0[1] * sc);
                        if (!act0[(size_t)s]) continue;
                        const int64_t sd[3] = {sa, sb, sc};
                        double val = 0.0;
                        for (int zc = 0; zc < ax[2].n; ++zc)
                            for (int yc = 0; yc < ax[1].n; ++yc)
                                for (int xc = 0; xc < ax[0].n; ++xc) {
                                    const int64_t td[3] = {ax[0].s[xc], ax[1].s[yc], ax[2].s[zc]};
                                    double w = 1.0;
                                    for (int d = 0; d < 3 && w != 0.0; ++d) {
                                        const int64_t kk = td[d] - (2 * sd[d] - 2);
                                        w = (kk >= 0 && kk <= 3) ? w * two_scale[kk] : 0.0;
                                    }
                                    if (w == 0.0) continue;
                                    if (act1[(size_t)(td[0] + s1[0] * (td[1] + s1[1] * td[2]))]) continue;
                                    val += w * (ax[0].w[xc] * ax[1].w[yc] * ax[2].w[zc]);
                                }
                        if (std::fabs(val) > 1e-12) out.push_back({0, s, val});
                    }
        };
        std::vector<int32_t> id0((size_t)(s0[0] * s0[1] * s0[2]), -1), id1((size_t)(s1[0] * s1[1] * s1[2]), -1);
        std::vector<Ent> r;
        size_t n_entries = 0;
        for (int64_t ee = 0; ee < n_el; ++ee)
            for (int v = 0; v < p->nb_ip; ++v) {
                row(ee, v, r);
                n_entries += r.size();
                for (const Ent& q : r) (q.level == 0 ? id0 : id1)[(size_t)q.sp] = 0;
            }
        int32_t next = 0;
        for (auto& a : id0)
            if (a == 0) a = next++;
        for (auto& a : id1)
            if (a == 0) a = next++;
        p->n_adofs_per_type = next;
        if (n_entries >= (size_t)INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many T entries");
        p->t_ptr.resize(n_rows + 1);
        p->t_adof.resize(n_entries);
        p->t_weight.resize(n_entries);
        size_t pos = 0;
        std::vector<std::pair<int32_t, double>> sorted;
        for (int64_t ee = 0; ee < n_el; ++ee)
            for (int v = 0; v < p->nb_ip; ++v) {
                row(ee, v, r);
                sorted.clear();
                for (const Ent& q : r) sorted.push_back({(q.level == 0 ? id0 : id1)[(size_t)q.sp], q.w});
                std::sort(sorted.begin(), sorted.end());
                p->t_ptr[(size_t)ee * p->nb_ip + v] = (int32_t)pos;
                for (const auto& q : sorted) {
                    p->t_adof[pos] = q.first;
                    p->t_weight[pos] = q.second;
                    ++pos;
                }
            }
        p->t_ptr[n_rows] = (int32_t)pos;
    } else if (order == 1 && p->nb_ip != 27) {
        // p = 1: one adof per used lattice node, ascending node index; T rows of weight 1
        std::vector<int32_t> node_adof((size_t)p->n_nodes, -1);
        for (int64_t ee = 0; ee < n_el; ++ee) {
            int ijk[3];
            elem_ijk(ee, ijk);
            for (int v = 0; v < p->nb; ++v) {
                const int64_t nid = p->dim == 3 ? node_index(*p, ijk[0] + kHex8Off[v][0], ijk[1] + kHex8Off[v][1], ijk[2] + kHex8Off[v][2])
                                                : node_index(*p, ijk[0] + kQuad4Off[v][0], ijk[1] + kQuad4Off[v][1], 0);
                node_adof[(size_t)nid] = 0;
            }
        }
        int32_t next = 0;
        for (int64_t nid = 0; nid < p->n_nodes; ++nid)
            if (node_adof[(size_t)nid] == 0) node_adof[(size_t)nid] = next++;
        p->n_adofs_per_type = enrich ? n_eadofs : next;
        p->t_ptr.resize(n_rows + 1);
        p->t_adof.resize(n_rows);
        p->t_weight.assign(n_rows, 1.0);
        for (int64_t ee = 0; ee < n_el; ++ee) {
            int ijk[3];
            elem_ijk(ee, ijk);
            for (int v = 0; v < p->nb; ++v) {
                const int* off = p->dim == 3 ? kHex8Off[v] : kQuad4Off[v];
                const int64_t nid = p->dim == 3 ? node_index(*p, ijk[0] + off[0], ijk[1] + off[1], ijk[2] + off[2])
                                                : node_index(*p, ijk[0] + off[0], ijk[1] + off[1], 0);
                const size_t r = (size_t)ee * p->nb + v;
                p->t_ptr[r] = (int32_t)r;
                const int off3[3] = {off[0], off[1], p->dim == 3 ? off[2] : 0};
                p->t_adof[r] = enrich ? enriched(ee, off3) : node_adof[(size_t)nid];
            }
        }
        p->t_ptr[n_rows] = (int32_t)n_rows;
    } else if (order == 1) {
        // HEX27 Lagrange mesh with nodal T rows: one adof per used node of the half-step grid,
        // ascending node index (x fastest), weight 1
        const int64_t gx = 2 * (int64_t)p->n[0] + 1, gy = 2 * (int64_t)p->n[1] + 1, gz = 2 * (int64_t)p->n[2] + 1;
        if (gx * gy * gz >= INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many HEX27 nodes");
        std::vector<int32_t> node_adof((size_t)(gx * gy * gz), -1);
        auto gid = [&](const int* g) { return (size_t)(g[0] + gx * (g[1] + gy * (int64_t)g[2])); };
        int g[3], ijk[3];
        for (int64_t ee = 0; ee < n_el; ++ee) {
            elem_ijk(ee, ijk);
            for (int v = 0; v < 27; ++v) {
                basis_half(ijk[0], ijk[1], ijk[2], v, g);
                node_adof[gid(g)] = 0;
            }
        }
        int32_t next = 0;
        for (auto& a : node_adof)
            if (a == 0) a = next++;
        p->n_adofs_per_type = enrich ? n_eadofs : next;
        p->t_ptr.resize(n_rows + 1);
        p->t_adof.resize(n_rows);
        p->t_weight.assign(n_rows, 1.0);
        for (int64_t ee = 0; ee < n_el; ++ee) {
            elem_ijk(ee, ijk);
            for (int v = 0; v < 27; ++v) {
                basis_half(ijk[0], ijk[1], ijk[2], v, g);
                const size_t r = (size_t)ee * 27 + v;
                p->t_ptr[r] = (int32_t)r;
                p->t_adof[r] = enrich ? enriched(ee, kHex27Half[v]) : node_adof[gid(g)];
            }
        }
        p->t_ptr[n_rows] = (int32_t)n_rows;
    } else {
        // p = 2: uniform quadratic B-splines on the lattice, n + 2 per axis (spline s covers the
        // cells s - 2 .. s).  Per axis, a lattice node m (even half-step coordinate 2m) carries
        // the splines m, m + 1 with value 1/2 each; the centre of cell m (odd coordinate 2m + 1)
        // carries the splines m, m + 1, m + 2 with 1/8, 3/4, 1/8.  A row is the tensor product,
        // x fastest.  Adofs: used splines in ascending spline index (x fastest).
        const int64_t sx = p->n[0] + 2, sy = p->n[1] + 2, sz = p->dim == 3 ? p->n[2] + 2 : 1;
        if (sx * sy * sz >= INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many splines");
        struct AxisRow { int n; int s[3]; double w[3]; };
        auto axis_row = [&](int d, int g, AxisRow& a) {
            if (d >= p->dim) { a.n = 1; a.s[0] = 0; a.w[0] = 1.0; return; }
            const int m = g / 2;
            if ((g & 1) == 0) {
                a.n = 2; a.s[0] = m; a.s[1] = m + 1; a.w[0] = a.w[1] = 0.5;
            } else {
                a.n = 3; a.s[0] = m; a.s[1] = m + 1; a.s[2] = m + 2; a.w[0] = 0.125; a.w[1] = 0.75; a.w[2] = 0.125;
            }
        };
        // row of basis v of element ee: (spline, weight) pairs
        auto row = [&](int64_t ee, int v, int64_t* sp, double* w) {
            int ijk[3], g[3];
            elem_ijk(ee, ijk);
            basis_half(ijk[0], ijk[1], ijk[2], v, g);
            AxisRow ax[3];
            for (int d = 0; d < 3; ++d) axis_row(d, g[d], ax[d]);
            int q = 0;
            for (int c = 0; c < ax[2].n; ++c)
                for (int b = 0; b < ax[1].n; ++b)
                    for (int a = 0; a < ax[0].n; ++a) {
                        sp[q] = ax[0].s[a] + sx * (ax[1].s[b] + sy * (int64_t)ax[2].s[c]);
                        w[q] = ax[0].w[a] * ax[1].w[b] * ax[2].w[c];
                        ++q;
                    }
            return q;
        };
        std::vector<int32_t> spline_adof((size_t)(sx * sy * sz), -1);
        int64_t sp[27];
        double wt[27];
        size_t n_entries = 0;
        for (int64_t ee = 0; ee < n_el; ++ee)
            for (int v = 0; v < p->nb_ip; ++v) {
                const int nq = row(ee, v, sp, wt);
                n_entries += (size_t)nq;
                for (int q = 0; q < nq; ++q) spline_adof[(size_t)sp[q]] = 0;
            }
        int32_t next = 0;
        for (auto& a : spline_adof)
            if (a == 0) a = next++;
        p->n_adofs_per_type = enrich ? n_eadofs : next;
        if (n_entries >= (size_t)INT32_MAX) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "too many T entries");
        p->t_ptr.resize(n_rows + 1);
        p->t_adof.resize(n_entries);
        p->t_weight.resize(n_entries);
        size_t pos = 0;
        for (int64_t ee = 0; ee < n_el; ++ee) {
            int ijk[3];
            elem_ijk(ee, ijk);
            for (int v = 0; v < p->nb_ip; ++v) {
                const int nq = row(ee, v, sp, wt);
                p->t_ptr[(size_t)ee * p->nb_ip + v] = (int32_t)pos;
                for (int q = 0; q < nq; ++q) {
                    if (enrich) {
                        // the spline's offsets from the element's cell (0..2 per axis)
                        const int64_t sidx[3] = {sp[q] % sx, (sp[q] / sx) % sy, sp[q] / (sx * sy)};
                        int off[3] = {0, 0, 0};
                        for (int d = 0; d < p->dim; ++d) off[d] = (int)(sidx[d] - ijk[d]);
                        p->t_adof[pos] = enriched(ee, off);
                    } else {
                        p->t_adof[pos] = spline_adof[(size_t)sp[q]];
                    }
                    p->t_weight[pos] = wt[q];
                    ++pos;
                }
            }
        }
        p->t_ptr[n_rows] = (int32_t)pos;
    }

    if (spec->ghost) {
        if (thb || p->nb_ip == 27 || enrich)
            return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "ghost facets are implemented for HEX8 / QUAD4 level-0 sets without enrichment");
        ghost_facets(*p, p->ghost);
        for (size_t k = 0; k < p->ghost.size(); k += 3) {
            p->couplings.push_back(p->ghost[k]);
            p->couplings.push_back(p->ghost[k + 1]);
        }
    }
    *out = p.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_problem_desc(const mgpu_problem* p, mgpu_set_desc* d) {
    if (!p || !d) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null argument");
    std::memset(d, 0, sizeof(*d));
    d->spatial_dim = p->dim;
    d->iwg = p->spec.iwg;
    d->ip_geometry = p->nb_ip == 27 ? MGPU_GEOM_HEX27 : (p->dim == 3 ? MGPU_GEOM_HEX8 : MGPU_GEOM_QUAD4);
    d->ig_geometry = p->dim == 3 ? MGPU_GEOM_TET4 : MGPU_GEOM_TRI3;
    d->n_elements = (int32_t)p->elem_cell.size();
    d->n_bases = p->nb_ip;
    d->n_adofs_per_type = p->n_adofs_per_type;
    d->ip_box = p->ip_box.data();
    d->material = p->material.data();
    d->t_ptr = p->t_ptr.data();
    d->t_adof = p->t_adof.data();
    d->t_weight = p->t_weight.data();
    d->ig_ptr = p->ig_ptr.data();
    d->ig_xi = p->ig_xi.data();
    for (int k = 0; k < 3; ++k) {
        d->lattice.n[k] = p->n[k];
        d->lattice.lo[k] = p->lo[k];
        d->lattice.h[k] = p->h[k];
    }
    // a THB set's elements are not the cells of one lattice: no lattice hint (generic path)
    if (!p->refined.empty()) d->lattice.n[0] = 0;
    d->elem_cell = p->elem_cell.data();
    d->vertex_phi = p->vertex_phi.data();
    d->n_couplings = (int32_t)(p->couplings.size() / 2);
    d->couplings = p->couplings.empty() ? nullptr : p->couplings.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_problem_counts(const mgpu_problem* p, int64_t counts[6]) {
    if (!p || !counts) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null argument");
    counts[0] = (int64_t)p->elem_cell.size();
    counts[1] = p->n_cut;
    counts[2] = p->ig_ptr.empty() ? 0 : p->ig_ptr.back();
    counts[3] = p->n_adofs_per_type;
    counts[4] = p->n_cells;
    counts[5] = p->n_void;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_problem_destroy(mgpu_problem* p) {
    delete p;
    return MGPU_OK;
}

// ------------------------------------------------------------------------------------------------
// Side facets of a problem (the reference's side sets of the XTK integration mesh: the domain
// boundary faces of the material cells, and the interface facets between the material sub-cells
// and the void).  A face of a material sub-cell is on the interface when it lies on no face of
// its cell and no other material sub-cell of that cell has the same vertices: the material
// sub-cells tile the cell's material region, so every other face is shared or on the cell's
// boundary.  Interface normals point out of the material sub-cell that owns the face.
// ------------------------------------------------------------------------------------------------
#include "../../../include/moris_gpu_sideset.h"

#include <array>
#include <cmath>
#include <map>

struct mgpu_facets {
    int32_t dim = 3, nv = 3;
    std::vector<int32_t> elem;
    std::vector<int32_t> follower;   // ghost facets only
    std::vector<double> xi;      // [n][nv][dim]
    std::vector<double> normal;  // [n][dim]
};

namespace {
constexpr double kFaceTol = 1e-12;   // a vertex lies on a cell face when |xi_d -+ 1| <= this

bool on_cell_face(const double* const* v, int nvf, int dim, int d, double s) {
    for (int k = 0; k < nvf; ++k)
        if (std::fabs(v[k][d] - s) > kFaceTol) return false;
    (void)dim;
    return true;
}
}  // namespace

extern "C" mgpu_status mgpu_problem_facets(const mgpu_problem* p, int32_t where, int32_t geom, mgpu_facets** out) {
    MGPU_GUARD_BEGIN
    if (!p || !out) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const int dim = p->dim;
    if (where == MGPU_SIDE_GHOST) {
        if (!(dim == 3 ? geom == MGPU_GEOM_QUAD4 : geom == MGPU_GEOM_LINE2))
            return mgpu::set_error(MGPU_ERR_INVALID_ARG, "ghost facets are whole cell faces: QUAD4 in 3D, LINE2 in 2D");
        if (!p->refined.empty() || p->nb_ip == 27 || p->spec.enrich)
            return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "ghost facets are implemented for HEX8 / QUAD4 level-0 sets without enrichment");
        std::vector<int32_t> g;
        if (p->spec.ghost) g = p->ghost;
        else ghost_facets(*p, g);
        auto F = std::make_unique<mgpu_facets>();
        F->dim = dim;
        F->nv = dim == 3 ? 4 : 2;
        const double cyc[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (size_t k = 0; k < g.size(); k += 3) {
            const int d = g[k + 2];
            F->elem.push_back(g[k]);
            F->follower.push_back(g[k + 1]);
            int ax[2] = {0, 0}, na = 0;
            for (int q = 0; q < dim; ++q)
                if (q != d) ax[na++] = q;
            for (int v = 0; v < F->nv; ++v) {
                double x[3] = {0, 0, 0};
                x[d] = 1.0;
                if (dim == 3) {
                    x[ax[0]] = cyc[v][0];
                    x[ax[1]] = cyc[v][1];
                } else {
                    x[ax[0]] = v == 0 ? -1.0 : 1.0;
                }
                for (int q = 0; q < dim; ++q) F->xi.push_back(x[q]);
            }
            for (int q = 0; q < dim; ++q) F->normal.push_back(q == d ? 1.0 : 0.0);
        }
        *out = F.release();
        return MGPU_OK;
    }
    if (where < MGPU_SIDE_INTERFACE || where > 2 * dim) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "bad facet location");
    const bool ok_geom = dim == 3 ? (geom == MGPU_GEOM_TRI3 || geom == MGPU_GEOM_QUAD4) : geom == MGPU_GEOM_LINE2;
    if (!ok_geom) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "side geometry must be TRI3 / QUAD4 in 3D, LINE2 in 2D");
    if (!p->refined.empty() && where != MGPU_SIDE_INTERFACE)
        return mgpu::set_error(MGPU_ERR_UNSUPPORTED, "boundary facets are implemented for level-0 lattices");
    auto F = std::make_unique<mgpu_facets>();
    F->dim = dim;
    const int nv_ig = dim + 1;                                 // sub-cell vertices
    F->nv = geom == MGPU_GEOM_QUAD4 ? 4 : dim;                 // facet vertices
    const int bd = where > 0 ? (where - 1) / 2 : -1;           // boundary axis
    const double bs = where > 0 ? ((where - 1) % 2 ? 1.0 : -1.0) : 0.0;
    const bool want_uncut = geom == MGPU_GEOM_QUAD4 || (dim == 2 && where != MGPU_SIDE_INTERFACE);
    const bool want_cut = geom != MGPU_GEOM_QUAD4;
    const size_t n_el = p->elem_cell.size();
    for (size_t e = 0; e < n_el; ++e) {
        const double* box = p->ip_box.data() + e * 2 * dim;
        auto phys = [&](const double* xi, double* x) {
            for (int d = 0; d < dim; ++d) x[d] = box[d] + 0.5 * (xi[d] + 1.0) * box[dim + d];
        };
        bool at_boundary = false;
        if (where > 0) {
            const int64_t c = p->elem_cell[e];
            const int64_t ijk[3] = {c % p->n[0], (c / p->n[0]) % p->n[1], dim == 3 ? c / ((int64_t)p->n[0] * p->n[1]) : 0};
            at_boundary = bs < 0 ? ijk[bd] == 0 : ijk[bd] == p->n[bd] - 1;
        }
        const int32_t s0 = p->ig_ptr[e], s1 = p->ig_ptr[e + 1];
        if (s0 == s1) {
            if (!(want_uncut && at_boundary)) continue;
            // the whole face (edge) of the box: the other axes in ascending order, cyclic vertices
            int ax[2] = {0, 0}, na = 0;
            for (int d = 0; d < dim; ++d)
                if (d != bd) ax[na++] = d;
            const double cyc[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
            F->elem.push_back((int32_t)e);
            for (int k = 0; k < F->nv; ++k) {
                double v[3] = {0, 0, 0};
                v[bd] = bs;
                if (dim == 3) {
                    v[ax[0]] = cyc[k][0];
                    v[ax[1]] = cyc[k][1];
                } else {
                    v[ax[0]] = k == 0 ? -1.0 : 1.0;
                }
                for (int d = 0; d < dim; ++d) F->xi.push_back(v[d]);
            }
            for (int d = 0; d < dim; ++d) F->normal.push_back(d == bd ? bs : 0.0);
            continue;
        }
        if (!want_cut) continue;
        if (where > 0 && !at_boundary) continue;
        // faces of the material sub-cells: count of each vertex set within the element
        std::map<std::array<double, 9>, int> count;
        auto face_key = [&](const double* X, int omit) {
            std::array<std::array<double, 3>, 3> vs{};
            int m = 0;
            for (int k = 0; k < nv_ig; ++k)
                if (k != omit) {
                    for (int d = 0; d < dim; ++d) vs[m][d] = X[k * dim + d];
                    ++m;
                }
            std::sort(vs.begin(), vs.begin() + m);
            std::array<double, 9> key{};
            for (int a = 0; a < m; ++a)
                for (int d = 0; d < 3; ++d) key[a * 3 + d] = vs[a][d];
            return key;
        };
        if (where == MGPU_SIDE_INTERFACE)
            for (int32_t s = s0; s < s1; ++s)
                for (int k = 0; k < nv_ig; ++k) ++count[face_key(p->ig_xi.data() + (size_t)s * nv_ig * dim, k)];
        for (int32_t s = s0; s < s1; ++s) {
            const double* X = p->ig_xi.data() + (size_t)s * nv_ig * dim;
            for (int k = 0; k < nv_ig; ++k) {
                const double* v[3];
                int m = 0;
                for (int q = 0; q < nv_ig; ++q)
                    if (q != k) v[m++] = X + q * dim;
                double nrm[3] = {0, 0, 0};
                if (where > 0) {
                    if (!on_cell_face(v, m, dim, bd, bs)) continue;
                    nrm[bd] = bs;
                } else {
                    bool on_face = false;
                    for (int d = 0; d < dim && !on_face; ++d)
                        on_face = on_cell_face(v, m, dim, d, -1.0) || on_cell_face(v, m, dim, d, 1.0);
                    if (on_face || count[face_key(X, k)] != 1) continue;
                    double x[3][3] = {}, xo[3] = {0, 0, 0};
                    for (int a = 0; a < m; ++a) phys(v[a], x[a]);
                    phys(X + k * dim, xo);
                    if (dim == 3) {
                        const double a1[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
                        const double a2[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
                        nrm[0] = a1[1] * a2[2] - a1[2] * a2[1];
                        nrm[1] = a1[2] * a2[0] - a1[0] * a2[2];
                        nrm[2] = a1[0] * a2[1] - a1[1] * a2[0];
                    } else {
                        nrm[0] = x[1][1] - x[0][1];
                        nrm[1] = -(x[1][0] - x[0][0]);
                    }
                    double dot = 0.0, len = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        dot += nrm[d] * (xo[d] - x[0][d]);
                        len += nrm[d] * nrm[d];
                    }
                    len = std::sqrt(len);
                    if (!(len > 0.0)) continue;   // degenerate face
                    const double sgn = dot > 0.0 ? -1.0 : 1.0;
                    for (int d = 0; d < dim; ++d) nrm[d] *= sgn / len;
                }
                F->elem.push_back((int32_t)e);
                for (int a = 0; a < m; ++a)
                    for (int d = 0; d < dim; ++d) F->xi.push_back(v[a][d]);
                for (int d = 0; d < dim; ++d) F->normal.push_back(nrm[d]);
            }
        }
    }
    *out = F.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_facets_get(const mgpu_facets* f, int32_t* n, const int32_t** elem, const double** xi,
                                       const double** normal) {
    if (!f) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null facets");
    if (n) *n = (int32_t)f->elem.size();
    if (elem) *elem = f->elem.data();
    if (xi) *xi = f->xi.data();
    if (normal) *normal = f->normal.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_facets_followers(const mgpu_facets* f, const int32_t** follower) {
    if (!f) return mgpu::set_error(MGPU_ERR_INVALID_ARG, "null facets");
    if (follower) *follower = f->follower.empty() ? nullptr : f->follower.data();
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_facets_destroy(mgpu_facets* f) {
    delete f;
    return MGPU_OK;
}
