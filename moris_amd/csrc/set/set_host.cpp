// set_host.cpp -- element topology, CSR pattern and lattice maps of an mgpu_set.
//
//   element topology   Equation_Object::get_equation_obj_dof_ids, cl_MSI_Equation_Object.cpp:505-554
//                      (create_my_list_of_adof_ids :199-291 -- sorted unique adof ids per dof type,
//                      concatenated over the requested types)
//   CSR pattern        Sparse_Matrix_EpetraFECrs::build_graph, cl_Sparse_Matrix_EpetraFECrs.cpp:195-240
//                      (union over elements of topology x topology), rows/cols ascending
#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "../common/fem_tables.hpp"
#include "set_internal.hpp"

namespace mgpu {

namespace {
int n_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 4;
    return (int)std::min(t, 32u);
}
template <class F>
void parallel_for(int64_t n, F f) {
    const int T = n_threads();
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t]() {
            const int64_t a = n * t / T, b = n * (t + 1) / T;
            for (int64_t i = a; i < b; ++i) f(i);
        });
    for (auto& x : th) x.join();
}
}  // namespace

// Generic path: element topology of any T rows (sorted unique adofs per type), element-local T
// columns, the CSR pattern as the union of the topologies (Sparse_Matrix_EpetraFECrs::build_graph,
// cl_Sparse_Matrix_EpetraFECrs.cpp:195-240) and, per CSR value, the list of element-matrix
// entries summed into it in ascending element order (the deterministic replacement of
// SumIntoGlobalValues, :103-133).  Large HEX27 sets take the scatter mode instead: elements in
// colour classes that share no adof, each adding its K_e into the values, one colour after the
// other (a fixed order of summation again), with per-element tables of column positions.
mgpu_status build_generic(const mgpu_set_desc& d, SetHost& h, std::string* err) {
    const int nb = h.nb, nf = h.nf;
    const int32_t N = h.n_adofs_type;
    h.na_t.assign((size_t)h.n_el, 0);
    h.t_col.assign((size_t)d.t_ptr[(size_t)h.n_el * nb], 0);
    // per element: sorted unique type-local adofs
    std::vector<std::vector<int32_t>> loc((size_t)h.n_el);
    std::atomic<bool> too_big{false};
    parallel_for(h.n_el, [&](int64_t e) {
        std::vector<int32_t>& u = loc[(size_t)e];
        for (int b = 0; b < nb; ++b)
            for (int32_t q = d.t_ptr[e * nb + b]; q < d.t_ptr[e * nb + b + 1]; ++q) u.push_back(d.t_adof[q]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if (u.size() > (size_t)kGenericMaxAdofsPerType) too_big.store(true);
        h.na_t[(size_t)e] = (int32_t)u.size();
        for (int b = 0; b < nb; ++b)
            for (int32_t q = d.t_ptr[e * nb + b]; q < d.t_ptr[e * nb + b + 1]; ++q)
                h.t_col[(size_t)q] = (int32_t)(std::lower_bound(u.begin(), u.end(), d.t_adof[q]) - u.begin());
    });
    if (too_big.load()) {
        *err = "an element has more than " + std::to_string(kGenericMaxAdofsPerType) + " adofs per dof type";
        return MGPU_ERR_UNSUPPORTED;
    }
    h.topo_ptr.assign((size_t)h.n_el + 1, 0);
    h.ke_off.assign((size_t)h.n_el + 1, 0);
    for (int32_t e = 0; e < h.n_el; ++e) {
        const int64_t na = (int64_t)nf * h.na_t[(size_t)e];
        h.topo_ptr[(size_t)e + 1] = h.topo_ptr[(size_t)e] + na;
        h.ke_off[(size_t)e + 1] = h.ke_off[(size_t)e] + na * na;
        h.na_t_max = std::max(h.na_t_max, h.na_t[(size_t)e]);
    }
    h.topo.resize((size_t)h.topo_ptr[(size_t)h.n_el]);
    parallel_for(h.n_el, [&](int64_t e) {
        int32_t* t = h.topo.data() + h.topo_ptr[(size_t)e];
        for (int f = 0; f < nf; ++f)
            for (int32_t a : loc[(size_t)e]) *t++ = a + f * N;
    });
    // adof -> elements
    std::vector<int64_t> a_ptr((size_t)N + 1, 0);
    for (int32_t e = 0; e < h.n_el; ++e)
        for (int32_t a : loc[(size_t)e]) a_ptr[(size_t)a + 1]++;
    for (int32_t a = 0; a < N; ++a) a_ptr[(size_t)a + 1] += a_ptr[(size_t)a];
    std::vector<int32_t> a_el((size_t)a_ptr[(size_t)N]);
    {
        std::vector<int64_t> fill(a_ptr.begin(), a_ptr.end() - 1);
        for (int32_t e = 0; e < h.n_el; ++e)
            for (int32_t a : loc[(size_t)e]) a_el[(size_t)fill[(size_t)a]++] = e;
    }
    // coupled element pairs (double-sided sets): each adof of one element of a pair is also a
    // neighbour of every adof of the other
    const int32_t ncp = d.couplings ? d.n_couplings : 0;
    std::vector<int64_t> c_ptr((size_t)N + 1, 0);
    std::vector<int32_t> c_el;
    if (ncp > 0) {
        for (int32_t k = 0; k < ncp; ++k)
            for (int s2 = 0; s2 < 2; ++s2)
                for (int32_t a : loc[(size_t)d.couplings[2 * k + s2]]) c_ptr[(size_t)a + 1]++;
        for (int32_t a = 0; a < N; ++a) c_ptr[(size_t)a + 1] += c_ptr[(size_t)a];
        c_el.resize((size_t)c_ptr[(size_t)N]);
        std::vector<int64_t> fill(c_ptr.begin(), c_ptr.end() - 1);
        for (int32_t k = 0; k < ncp; ++k)
            for (int s2 = 0; s2 < 2; ++s2)
                for (int32_t a : loc[(size_t)d.couplings[2 * k + s2]])
                    c_el[(size_t)fill[(size_t)a]++] = d.couplings[2 * k + 1 - s2];   // the partner element
    }
    // neighbours of each type-local adof
    std::vector<std::vector<int32_t>> nbr((size_t)N);
    parallel_for(N, [&](int64_t a) {
        std::vector<int32_t>& v = nbr[(size_t)a];
        for (int64_t q = a_ptr[(size_t)a]; q < a_ptr[(size_t)a + 1]; ++q) {
            const auto& u = loc[(size_t)a_el[(size_t)q]];
            v.insert(v.end(), u.begin(), u.end());
        }
        for (int64_t q = c_ptr[(size_t)a]; ncp > 0 && q < c_ptr[(size_t)a + 1]; ++q) {
            const auto& u = loc[(size_t)c_el[(size_t)q]];
            v.insert(v.end(), u.begin(), u.end());
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    });
    h.row_ptr.assign((size_t)h.n_rows + 1, 0);
    int64_t nnz = 0;
    for (int f = 0; f < nf; ++f)
        for (int32_t a = 0; a < N; ++a) {
            nnz += (int64_t)nf * (int64_t)nbr[(size_t)a].size();
            if (nnz >= INT32_MAX) { *err = "nnz exceeds int32"; return MGPU_ERR_INVALID_ARG; }
            h.row_ptr[(size_t)f * N + a + 1] = (int32_t)nnz;
        }
    h.col_idx.resize((size_t)nnz);
    h.nnz = (int64_t)nnz;
    parallel_for(N, [&](int64_t a) {
        const auto& v = nbr[(size_t)a];
        for (int f = 0; f < nf; ++f) {
            int32_t* dst = h.col_idx.data() + h.row_ptr[(size_t)f * N + a];
            for (int g = 0; g < nf; ++g)
                for (int32_t c : v) *dst++ = c + g * N;
        }
    });
    // cut elements
    h.elem_cut.assign((size_t)h.n_el, -1);
    h.cut_elems.clear();
    for (int32_t e = 0; e < h.n_el; ++e)
        if (d.ig_ptr[e + 1] > d.ig_ptr[e]) {
            h.elem_cut[(size_t)e] = (int32_t)h.cut_elems.size();
            h.cut_elems.push_back(e);
        }
    const int64_t n_src = h.ke_off[(size_t)h.n_el];
    h.scatter = nb == 27 && n_src > kGatherMaxEntries;
    if (nb == 27)
        if (const char* env = std::getenv("MGPU_GENERIC_SCATTER")) {
            if (env[0] == '1') h.scatter = true;
            if (env[0] == '0') h.scatter = false;
        }
    if (h.scatter) {
        // greedy colouring in element order: the smallest colour that no element sharing an
        // adof with e already holds
        std::vector<int32_t> colour((size_t)h.n_el, -1), seen;
        int32_t n_col = 0;
        for (int32_t e = 0; e < h.n_el; ++e) {
            for (int32_t a : loc[(size_t)e])
                for (int64_t q = a_ptr[(size_t)a]; q < a_ptr[(size_t)a + 1]; ++q) {
                    const int32_t c = colour[(size_t)a_el[(size_t)q]];
                    if (c >= 0) seen[(size_t)c] = e;
                }
            int32_t c = 0;
            while (c < n_col && seen[(size_t)c] == e) ++c;
            if (c == n_col) {
                ++n_col;
                seen.push_back(-1);
            }
            colour[(size_t)e] = c;
        }
        h.colour_ptr.assign((size_t)n_col + 1, 0);
        for (int32_t e = 0; e < h.n_el; ++e) h.colour_ptr[(size_t)colour[(size_t)e] + 1]++;
        for (int32_t c = 0; c < n_col; ++c) h.colour_ptr[(size_t)c + 1] += h.colour_ptr[(size_t)c];
        h.colour_el.resize((size_t)h.n_el);
        {
            std::vector<int32_t> fill(h.colour_ptr.begin(), h.colour_ptr.end() - 1);
            for (int32_t e = 0; e < h.n_el; ++e) h.colour_el[(size_t)fill[(size_t)colour[(size_t)e]]++] = e;
        }
        // within a colour the elements with most sub-cells first, so that a launch does not end
        // on one long element
        for (int32_t c = 0; c < n_col; ++c)
            std::stable_sort(h.colour_el.begin() + h.colour_ptr[(size_t)c], h.colour_el.begin() + h.colour_ptr[(size_t)c + 1],
                             [&](int32_t x, int32_t y) {
                                 return d.ig_ptr[x + 1] - d.ig_ptr[x] > d.ig_ptr[y + 1] - d.ig_ptr[y];
                             });
        // rank tables: both the element's adofs and each adof's neighbours are ascending, so
        // one merge per row finds every column
        for (int32_t a = 0; a < N; ++a)
            if (nbr[(size_t)a].size() > 65535) { *err = "an adof has more than 65535 neighbours"; return MGPU_ERR_UNSUPPORTED; }
        h.rank_off.assign((size_t)h.n_el + 1, 0);
        for (int32_t e = 0; e < h.n_el; ++e)
            h.rank_off[(size_t)e + 1] = h.rank_off[(size_t)e] + (int64_t)h.na_t[(size_t)e] * h.na_t[(size_t)e];
        h.rank.resize((size_t)h.rank_off[(size_t)h.n_el]);
        parallel_for(h.n_el, [&](int64_t e) {
            const auto& u = loc[(size_t)e];
            const int na = (int)u.size();
            uint16_t* rk = h.rank.data() + h.rank_off[(size_t)e];
            for (int x = 0; x < na; ++x) {
                const auto& v = nbr[(size_t)u[(size_t)x]];
                size_t p = 0;
                for (int y = 0; y < na; ++y) {
                    while (v[p] < u[(size_t)y]) ++p;
                    rk[x * na + y] = (uint16_t)p;
                }
            }
        });
        return MGPU_OK;
    }
    // gather lists: K_e entry (x, y) of element e goes to CSR value (topo[x], topo[y]).  Built per
    // CSR row, rows in parallel: the values of row f N + a take entries only from the elements
    // holding adof a (a_el, ascending), each element at most one entry per value (its topology has
    // no repeated adof), so visiting those elements in ascending order lists each value's entries
    // in ascending k = ke_off[e] + x na + y, the order of summation.  The element's topology and
    // the row's columns are both ascending: one merge per (row, element) finds every column.
    if (n_src >= ((int64_t)1 << 40)) { *err = "element matrices too large"; return MGPU_ERR_INVALID_ARG; }
    // visit(row, fn): fn(p, k) for every entry k of the row's values p, elements ascending
    auto visit = [&](int64_t row, auto&& fn) {
        const int32_t a = (int32_t)(row % N);
        const int f = (int)(row / N);
        const int32_t* cols = h.col_idx.data();
        const int64_t c0 = h.row_ptr[(size_t)row], c1 = h.row_ptr[(size_t)row + 1];
        for (int64_t q = a_ptr[(size_t)a]; q < a_ptr[(size_t)a + 1]; ++q) {
            const int32_t e = a_el[(size_t)q];
            const auto& u = loc[(size_t)e];
            const int64_t nat = (int64_t)u.size(), na = (int64_t)nf * nat;
            const int64_t x = f * nat + (int64_t)(std::lower_bound(u.begin(), u.end(), a) - u.begin());
            const int32_t* t = h.topo.data() + h.topo_ptr[(size_t)e];
            const int64_t k0 = h.ke_off[(size_t)e] + x * na;
            int64_t p = c0;
            for (int64_t y = 0; y < na; ++y) {
                while (p < c1 && cols[p] < t[y]) ++p;
                fn(p, k0 + y);
            }
        }
    };
    h.gather_ptr.assign((size_t)nnz + 1, 0);
    parallel_for(h.n_rows, [&](int64_t row) { visit(row, [&](int64_t p, int64_t) { h.gather_ptr[(size_t)p + 1]++; }); });
    // prefix sums: row totals in order, then each row's values from its base, rows in parallel
    std::vector<int64_t> row_base((size_t)h.n_rows + 1, 0);
    parallel_for(h.n_rows, [&](int64_t row) {
        int64_t s = 0;
        for (int64_t p = h.row_ptr[(size_t)row]; p < h.row_ptr[(size_t)row + 1]; ++p) s += h.gather_ptr[(size_t)p + 1];
        row_base[(size_t)row + 1] = s;
    });
    for (int32_t r = 0; r < h.n_rows; ++r) row_base[(size_t)r + 1] += row_base[(size_t)r];
    parallel_for(h.n_rows, [&](int64_t row) {
        int64_t s = row_base[(size_t)row];
        for (int64_t p = h.row_ptr[(size_t)row]; p < h.row_ptr[(size_t)row + 1]; ++p) {
            s += h.gather_ptr[(size_t)p + 1];
            h.gather_ptr[(size_t)p + 1] = s;
        }
    });
    h.gather_src.resize((size_t)n_src);
    parallel_for(h.n_rows, [&](int64_t row) {
        const int64_t c0 = h.row_ptr[(size_t)row], c1 = h.row_ptr[(size_t)row + 1];
        std::vector<int64_t> fill(h.gather_ptr.begin() + c0, h.gather_ptr.begin() + c1);
        visit(row, [&](int64_t p, int64_t k) { h.gather_src[(size_t)fill[(size_t)(p - c0)]++] = k; });
    });
    return MGPU_OK;
}

mgpu_status build_set_host(const mgpu_set_desc& d, SetHost& h, std::string* err) {
    if (d.spatial_dim != 2 && d.spatial_dim != 3) { *err = "spatial_dim must be 2 or 3"; return MGPU_ERR_INVALID_ARG; }
    if (d.iwg != MGPU_IWG_SPATIALDIFF_BULK && d.iwg != MGPU_IWG_STRUC_LINEAR_BULK) {
        *err = "only SPATIALDIFF_BULK and STRUC_LINEAR_BULK are implemented";
        return MGPU_ERR_UNSUPPORTED;
    }
    const int want_ip = d.spatial_dim == 3 ? MGPU_GEOM_HEX8 : MGPU_GEOM_QUAD4;
    const int want_ig = d.spatial_dim == 3 ? MGPU_GEOM_TET4 : MGPU_GEOM_TRI3;
    const bool hex27 = d.spatial_dim == 3 && d.ip_geometry == MGPU_GEOM_HEX27;
    if ((d.ip_geometry != want_ip && !hex27) || d.ig_geometry != want_ig) {
        *err = "ip/ig geometry must be HEX8/TET4 or HEX27/TET4 (3D), QUAD4/TRI3 (2D)";
        return MGPU_ERR_UNSUPPORTED;
    }
    const int nb_expect = hex27 ? 27 : (d.spatial_dim == 3 ? 8 : 4);
    if (d.n_bases != nb_expect) { *err = "n_bases must be 27 (HEX27), 8 (HEX8) or 4 (QUAD4)"; return MGPU_ERR_INVALID_ARG; }
    if (d.n_elements < 0 || d.n_adofs_per_type < 0) { *err = "negative sizes"; return MGPU_ERR_INVALID_ARG; }
    if (d.n_elements > 0 && (!d.ip_box || !d.material || !d.t_ptr || !d.t_adof || !d.t_weight || !d.ig_ptr)) {
        *err = "descriptor array missing";
        return MGPU_ERR_INVALID_ARG;
    }
    h.dim = d.spatial_dim;
    h.iwg = d.iwg;
    h.nf = d.iwg == MGPU_IWG_STRUC_LINEAR_BULK ? d.spatial_dim : 1;
    h.nb = d.n_bases;
    h.na = h.nf * h.nb;
    h.nm = d.iwg == MGPU_IWG_STRUC_LINEAR_BULK ? 2 : 1;
    h.n_el = d.n_elements;
    h.n_adofs_type = d.n_adofs_per_type;
    if ((int64_t)h.nf * h.n_adofs_type >= INT32_MAX) { *err = "too many adofs for int32 ids"; return MGPU_ERR_INVALID_ARG; }
    h.n_rows = h.nf * h.n_adofs_type;
    const int nb = h.nb, nf = h.nf, na = h.na;
    const int32_t N = h.n_adofs_type;

    // validate T rows and sub-cell offsets
    const int64_t n_trows = (int64_t)h.n_el * nb;
    if (d.t_ptr[0] != 0) { *err = "t_ptr[0] must be 0"; return MGPU_ERR_INVALID_ARG; }
    for (int64_t r = 0; r < n_trows; ++r) {
        if (d.t_ptr[r + 1] < d.t_ptr[r]) { *err = "t_ptr not monotone"; return MGPU_ERR_INVALID_ARG; }
        for (int32_t q = d.t_ptr[r]; q < d.t_ptr[r + 1]; ++q)
            if (d.t_adof[q] < 0 || d.t_adof[q] >= N) { *err = "t_adof out of range"; return MGPU_ERR_INVALID_ARG; }
    }
    if (d.ig_ptr[0] != 0) { *err = "ig_ptr[0] must be 0"; return MGPU_ERR_INVALID_ARG; }
    for (int32_t e = 0; e < h.n_el; ++e)
        if (d.ig_ptr[e + 1] < d.ig_ptr[e]) { *err = "ig_ptr not monotone"; return MGPU_ERR_INVALID_ARG; }
    if (d.ig_ptr[h.n_el] > 0 && !d.ig_xi) { *err = "ig_xi missing"; return MGPU_ERR_INVALID_ARG; }
    // material: k > 0 (CM_Diffusion_Linear_Isotropic), E > 0 and -1 < nu < 1/2
    // (CM_Struc_Linear_Isotropic); the kernels encode void and cut cells beside these values
    for (int32_t e = 0; e < h.n_el; ++e) {
        if (h.nm == 1) {
            const double k = d.material[e];
            if (!(std::isfinite(k) && k > 0.0)) { *err = "conductivity must be finite and positive"; return MGPU_ERR_INVALID_ARG; }
        } else {
            const double E = d.material[2 * (size_t)e], nu = d.material[2 * (size_t)e + 1];
            if (!(std::isfinite(E) && E > 0.0 && nu > -1.0 && nu < 0.5)) {
                *err = "elasticity needs finite E > 0 and -1 < nu < 0.5";
                return MGPU_ERR_INVALID_ARG;
            }
        }
    }
    // sub-cells the kernels skip (detJ < -1e-6, as in cl_FEM_Element_Bulk.cpp:291): counted once
    // here with the kernels' formula (rows X_r - X_2 of the vertices scaled by half the cell extent)
    {
        const int D = h.dim, nv = D + 1;
        const int npts = hex27 ? kTet35Points : (D == 3 ? kTet11Points : kTri7Points);
        for (int32_t e = 0; e < h.n_el; ++e) {
            const double* ext = d.ip_box + (size_t)e * 2 * D + D;
            for (int32_t s = d.ig_ptr[e]; s < d.ig_ptr[e + 1]; ++s) {
                const double* X = d.ig_xi + (size_t)s * nv * D;
                double detJ;
                if (D == 3) {
                    double J[3][3];
                    const int rr[3] = {0, 1, 3};
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) J[r][c] = (X[rr[r] * 3 + c] - X[2 * 3 + c]) * (0.5 * ext[c]);
                    detJ = (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / 6.0;
                } else {
                    double J[2][2];
                    for (int r = 0; r < 2; ++r)
                        for (int c = 0; c < 2; ++c) J[r][c] = (X[r * 2 + c] - X[2 * 2 + c]) * (0.5 * ext[c]);
                    detJ = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / 2.0;
                }
                if (detJ < kDetJLowerLimit) {
                    ++h.n_skip_subcells;
                    h.n_skip_points += npts;
                }
            }
        }
    }

    // lattice stencil path: nodal T rows (one adof of weight 1 per basis) on a lattice; anything
    // else (B-spline p > 1, THB rows, sets without a lattice) takes the generic element path
    bool nodal = true;
    for (int64_t r = 0; r < n_trows && nodal; ++r)
        nodal = (d.t_ptr[r + 1] - d.t_ptr[r] == 1) && d.t_weight[d.t_ptr[r]] == 1.0;
    const bool has_lattice = d.lattice.n[0] > 0 && d.elem_cell != nullptr;
    // the stencil also needs one element per cell and one adof per lattice node, ascending in node
    // order; enriched sets (several subphase elements in a cell, a node replicated per level) and
    // other numberings take the generic path
    bool one_per_node = true;
    if (nodal && has_lattice && !hex27) {
        const int64_t lx = d.lattice.n[0], ly = d.lattice.n[1], lz = d.spatial_dim == 3 ? d.lattice.n[2] : 1;
        const int64_t cells = lx * ly * lz;
        std::vector<char> seen((size_t)std::max<int64_t>(cells, 0), 0);
        std::vector<int32_t> nadof((size_t)(d.spatial_dim == 3 ? (lx + 1) * (ly + 1) * (lz + 1) : (lx + 1) * (ly + 1)), -1);
        for (int32_t e = 0; e < d.n_elements && one_per_node; ++e) {
            const int64_t c = d.elem_cell[e];
            if (c < 0 || c >= cells) break;   // reported by the stencil set-up
            if (seen[(size_t)c]) { one_per_node = false; break; }
            seen[(size_t)c] = 1;
            const int64_t i = c % lx, j = (c / lx) % ly, k = c / (lx * ly);
            for (int b = 0; b < nb; ++b) {
                const int* off = d.spatial_dim == 3 ? kHex8Off[b] : kQuad4Off[b];
                const int64_t nid = d.spatial_dim == 3 ? (i + off[0]) + (lx + 1) * ((j + off[1]) + (ly + 1) * (k + off[2]))
                                                       : (i + off[0]) + (lx + 1) * (j + off[1]);
                const int32_t a = d.t_adof[d.t_ptr[(int64_t)e * nb + b]];
                if (nadof[(size_t)nid] >= 0 && nadof[(size_t)nid] != a) { one_per_node = false; break; }
                nadof[(size_t)nid] = a;
            }
        }
        int32_t last = -1;
        for (size_t nid = 0; nid < nadof.size() && one_per_node; ++nid) {
            if (nadof[nid] < 0) continue;
            if (nadof[nid] <= last) one_per_node = false;
            last = nadof[nid];
        }
    }
    if (d.n_couplings < 0 || (d.n_couplings > 0 && !d.couplings)) {
        *err = "bad coupling arrays";
        return MGPU_ERR_INVALID_ARG;
    }
    for (int32_t k = 0; k < 2 * d.n_couplings; ++k)
        if (d.couplings[k] < 0 || d.couplings[k] >= h.n_el) {
            *err = "coupled element out of range";
            return MGPU_ERR_INVALID_ARG;
        }
    if (!nodal || !has_lattice || hex27 || !one_per_node || d.n_couplings > 0) {
        h.generic = true;
        const mgpu_status rc = build_generic(d, h, err);
        // the HEX27 kernel stages T and T^T K_p T blocks of 27 x na_t in LDS beside K_p
        if (rc == MGPU_OK && hex27 && h.na_t_max > kHex27MaxAdofsPerType) {
            *err = "a HEX27 element has more than " + std::to_string(kHex27MaxAdofsPerType) + " adofs per dof type";
            return MGPU_ERR_UNSUPPORTED;
        }
        return rc;
    }

    // element topology: sorted unique adofs per type; vertex -> sorted position
    h.topo.resize((size_t)h.n_el * na);
    h.vert_pos.resize((size_t)h.n_el * nb);
    std::atomic<bool> dup{false};
    parallel_for(h.n_el, [&](int64_t e) {
        int32_t ids[8];
        for (int b = 0; b < nb; ++b) ids[b] = d.t_adof[d.t_ptr[e * nb + b]];
        int32_t s[8];
        std::copy(ids, ids + nb, s);
        std::sort(s, s + nb);
        if (std::adjacent_find(s, s + nb) != s + nb) dup.store(true);
        for (int b = 0; b < nb; ++b)
            h.vert_pos[(size_t)e * nb + b] = (uint8_t)(std::lower_bound(s, s + nb, ids[b]) - s);
        for (int f = 0; f < nf; ++f)
            for (int b = 0; b < nb; ++b) h.topo[(size_t)e * na + f * nb + b] = s[b] + f * N;
    });
    if (dup.load()) { *err = "an element references one adof from two bases"; return MGPU_ERR_UNSUPPORTED; }
    h.topo_ptr.resize((size_t)h.n_el + 1);
    for (int64_t e = 0; e <= h.n_el; ++e) h.topo_ptr[(size_t)e] = e * na;

    // adof -> elements (counting sort)
    std::vector<int32_t> a_ptr((size_t)N + 1, 0), a_el;
    for (int64_t e = 0; e < h.n_el; ++e)
        for (int b = 0; b < nb; ++b) a_ptr[(size_t)h.topo[(size_t)e * na + b] + 1]++;
    for (int32_t a = 0; a < N; ++a) a_ptr[(size_t)a + 1] += a_ptr[(size_t)a];
    a_el.resize((size_t)a_ptr[(size_t)N]);
    {
        std::vector<int32_t> fill(a_ptr.begin(), a_ptr.end() - 1);
        for (int64_t e = 0; e < h.n_el; ++e)
            for (int b = 0; b < nb; ++b) a_el[(size_t)fill[(size_t)h.topo[(size_t)e * na + b]]++] = (int32_t)e;
    }
    // per type-local adof: sorted unique neighbour adofs (union of element topologies)
    std::vector<int32_t> deg((size_t)N, 0);
    parallel_for(N, [&](int64_t a) {
        int32_t buf[8 * 27];
        int n = 0;
        for (int32_t q = a_ptr[(size_t)a]; q < a_ptr[(size_t)a + 1]; ++q) {
            const int64_t e = a_el[(size_t)q];
            for (int b = 0; b < nb; ++b) {
                if (n < 8 * 27) buf[n++] = h.topo[(size_t)e * na + b];
            }
        }
        std::sort(buf, buf + n);
        deg[(size_t)a] = (int32_t)(std::unique(buf, buf + n) - buf);
    });
    // rows: type-major; row (f, a) has columns of every type g: g*N + nbr(a)
    h.row_ptr.resize((size_t)h.n_rows + 1);
    h.row_ptr[0] = 0;
    int64_t nnz = 0;
    for (int f = 0; f < nf; ++f)
        for (int32_t a = 0; a < N; ++a) {
            nnz += (int64_t)nf * deg[(size_t)a];
            if (nnz >= INT32_MAX) { *err = "nnz exceeds int32"; return MGPU_ERR_INVALID_ARG; }
            h.row_ptr[(size_t)f * N + a + 1] = (int32_t)nnz;
        }
    h.col_idx.resize((size_t)nnz);
    h.nnz = (int64_t)nnz;
    parallel_for(N, [&](int64_t a) {
        int32_t buf[8 * 27];
        int n = 0;
        for (int32_t q = a_ptr[(size_t)a]; q < a_ptr[(size_t)a + 1]; ++q) {
            const int64_t e = a_el[(size_t)q];
            for (int b = 0; b < nb; ++b)
                if (n < 8 * 27) buf[n++] = h.topo[(size_t)e * na + b];
        }
        std::sort(buf, buf + n);
        n = (int)(std::unique(buf, buf + n) - buf);
        for (int f = 0; f < nf; ++f) {
            int32_t* dst = h.col_idx.data() + h.row_ptr[(size_t)f * N + a];
            for (int g = 0; g < nf; ++g)
                for (int i = 0; i < n; ++i) *dst++ = buf[i] + g * N;
        }
    });
    if (h.n_el > 0 && a_ptr[(size_t)N] > 0) {
        // an adof of a single element has at most nb neighbours; an interior node 27 (3D) / 9 (2D)
        for (int32_t a = 0; a < N; ++a)
            if (deg[(size_t)a] > (h.dim == 3 ? 27 : 9)) {
                *err = "adof with more neighbours than a lattice stencil";
                return MGPU_ERR_UNSUPPORTED;
            }
    }

    // cut elements
    h.elem_cut.assign((size_t)h.n_el, -1);
    h.cut_elems.clear();
    for (int32_t e = 0; e < h.n_el; ++e)
        if (d.ig_ptr[e + 1] > d.ig_ptr[e]) {
            h.elem_cut[(size_t)e] = (int32_t)h.cut_elems.size();
            h.cut_elems.push_back(e);
        }

    if (h.n_el == 0) return MGPU_OK;   // empty set: nothing to assemble

    // lattice maps
    h.lattice = d.lattice.n[0] > 0 && d.elem_cell != nullptr;
    if (!h.lattice) {
        *err = "sets without a lattice hint are not implemented";
        return MGPU_ERR_UNSUPPORTED;
    }
    h.lat = d.lattice;
    if (h.dim == 2) { h.lat.n[2] = 1; }
    for (int k = 0; k < h.dim; ++k)
        if (h.lat.n[k] < 1 || !(h.lat.h[k] > 0.0)) { *err = "bad lattice"; return MGPU_ERR_INVALID_ARG; }
    const int64_t nx = h.lat.n[0], ny = h.lat.n[1], nz = h.dim == 3 ? h.lat.n[2] : 1;
    h.n_cells = nx * ny * nz;
    h.n_nodes = h.dim == 3 ? (nx + 1) * (ny + 1) * (nz + 1) : (nx + 1) * (ny + 1);
    // the stencil kernel indexes lattice nodes in int32
    if (h.n_nodes >= INT32_MAX) { *err = "lattice has 2^31 or more nodes"; return MGPU_ERR_UNSUPPORTED; }
    h.cell_elem.assign((size_t)h.n_cells, -1);
    h.node_adof.assign((size_t)h.n_nodes, -1);
    for (int32_t e = 0; e < h.n_el; ++e) {
        const int64_t c = d.elem_cell[e];
        if (c < 0 || c >= h.n_cells || h.cell_elem[(size_t)c] >= 0) {
            *err = "elem_cell out of range or duplicated";
            return MGPU_ERR_INVALID_ARG;
        }
        h.cell_elem[(size_t)c] = e;
        const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        // the IP box must be the lattice cell
        const int64_t ijk[3] = {i, j, k};
        for (int q = 0; q < h.dim; ++q) {
            const double lo = h.lat.lo[q] + (double)ijk[q] * h.lat.h[q];
            if (d.ip_box[(size_t)e * 2 * h.dim + q] != lo || d.ip_box[(size_t)e * 2 * h.dim + h.dim + q] != h.lat.h[q]) {
                *err = "ip_box does not match the lattice cell";
                return MGPU_ERR_INVALID_ARG;
            }
        }
        for (int b = 0; b < nb; ++b) {
            const int* off = h.dim == 3 ? kHex8Off[b] : kQuad4Off[b];
            const int64_t nid = h.dim == 3 ? (i + off[0]) + (nx + 1) * ((j + off[1]) + (ny + 1) * (k + off[2]))
                                           : (i + off[0]) + (nx + 1) * (j + off[1]);
            const int32_t a = d.t_adof[d.t_ptr[(size_t)e * nb + b]];
            if (h.node_adof[(size_t)nid] >= 0 && h.node_adof[(size_t)nid] != a) {
                *err = "lattice node mapped to two adofs";
                return MGPU_ERR_UNSUPPORTED;
            }
            h.node_adof[(size_t)nid] = a;
        }
    }
    h.row_node.assign((size_t)N, -1);
    int32_t last = -1;
    for (int64_t nid = 0; nid < h.n_nodes; ++nid) {
        const int32_t a = h.node_adof[(size_t)nid];
        if (a < 0) continue;
        if (a <= last) {
            *err = "adof ids not ascending in lattice node order";
            return MGPU_ERR_UNSUPPORTED;
        }
        last = a;
        h.row_node[(size_t)a] = (int32_t)nid;
    }
    for (int32_t a = 0; a < N; ++a)
        if (h.row_node[(size_t)a] < 0) { *err = "adof not used by any element"; return MGPU_ERR_INVALID_ARG; }
    return MGPU_OK;
}

}  // namespace mgpu
