// enrich.cpp -- XTK subphases of a cut cell and basis-cluster enrichment numbering (enrich.hpp).
#include "enrich.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <numeric>
#include <utility>

#include "../common/fem_tables.hpp"

namespace mgpu {
namespace {

int find_root(std::vector<int>& parent, int a) {
    while (parent[(size_t)a] != a) {
        parent[(size_t)a] = parent[(size_t)parent[(size_t)a]];
        a = parent[(size_t)a];
    }
    return a;
}

void unite(std::vector<int>& parent, int a, int b) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a != b) parent[(size_t)std::max(a, b)] = std::min(a, b);
}

// corner offsets of vertex v (kHex8Off / kQuad4Off), 0 on unused axes
void corner_offset(int dim, int v, int* off) {
    for (int d = 0; d < 3; ++d) off[d] = 0;
    for (int d = 0; d < dim; ++d) off[d] = dim == 3 ? kHex8Off[v][d] : kQuad4Off[v][d];
}

int corner_of_offset(int dim, const int* off) {
    const int nc = 1 << dim;
    for (int v = 0; v < nc; ++v) {
        int o[3];
        corner_offset(dim, v, o);
        bool same = true;
        for (int d = 0; d < dim; ++d) same = same && o[d] == off[d];
        if (same) return v;
    }
    return -1;
}

}  // namespace

void split_subphases(int dim, int n_sub, double* xi, const int* corner_region, int phase,
                     std::vector<int32_t>& count, std::vector<uint8_t>& corner) {
    const int nv = dim + 1, stride = nv * dim;
    // faces of the sub-cells as sorted vertex lists; equal faces join their sub-cells
    using Key = std::array<double, 9>;
    std::vector<std::pair<Key, int>> faces;
    faces.reserve((size_t)n_sub * nv);
    for (int s = 0; s < n_sub; ++s) {
        const double* X = xi + (size_t)s * stride;
        for (int omit = 0; omit < nv; ++omit) {
            std::array<std::array<double, 3>, 3> vs{};
            int m = 0;
            for (int k = 0; k < nv; ++k)
                if (k != omit) {
                    for (int d = 0; d < dim; ++d) vs[(size_t)m][(size_t)d] = X[k * dim + d];
                    ++m;
                }
            std::sort(vs.begin(), vs.begin() + m);
            Key key{};
            for (int a = 0; a < m; ++a)
                for (int d = 0; d < 3; ++d) key[(size_t)(a * 3 + d)] = vs[(size_t)a][(size_t)d];
            faces.push_back({key, s});
        }
    }
    std::sort(faces.begin(), faces.end());
    std::vector<int> parent((size_t)n_sub);
    std::iota(parent.begin(), parent.end(), 0);
    for (size_t f = 1; f < faces.size(); ++f)
        if (faces[f].first == faces[f - 1].first) unite(parent, faces[f].second, faces[f - 1].second);
    // subphase of each sub-cell, numbered by first sub-cell
    std::vector<int> comp((size_t)n_sub, -1), root_comp((size_t)n_sub, -1);
    int n_sp = 0;
    for (int s = 0; s < n_sub; ++s) {
        const int r = find_root(parent, s);
        if (root_comp[(size_t)r] < 0) root_comp[(size_t)r] = n_sp++;
        comp[(size_t)s] = root_comp[(size_t)r];
    }
    count.assign((size_t)n_sp, 0);
    corner.assign((size_t)n_sp, 0);
    for (int s = 0; s < n_sub; ++s) ++count[(size_t)comp[(size_t)s]];
    // strictly-material corners: the subphase with a vertex at the corner
    const int nc = 1 << dim;
    for (int v = 0; v < nc; ++v) {
        if (corner_region[v] != phase) continue;
        const double* hat = dim == 3 ? kHex8XiHat[v] : kQuad4XiHat[v];
        for (int s = 0; s < n_sub; ++s) {
            const double* X = xi + (size_t)s * stride;
            bool hit = false;
            for (int k = 0; k < nv && !hit; ++k) {
                bool eq = true;
                for (int d = 0; d < dim; ++d) eq = eq && X[k * dim + d] == hat[d];
                hit = eq;
            }
            if (hit) {
                corner[(size_t)comp[(size_t)s]] |= (uint8_t)(1u << v);
                break;
            }
        }
    }
    if (n_sp < 2) return;
    // stable regroup by subphase
    std::vector<double> tmp((size_t)n_sub * stride);
    std::vector<int32_t> pos((size_t)n_sp, 0);
    for (int c = 1; c < n_sp; ++c) pos[(size_t)c] = pos[(size_t)c - 1] + count[(size_t)c - 1];
    for (int s = 0; s < n_sub; ++s) {
        const int32_t at = pos[(size_t)comp[(size_t)s]]++;
        std::memcpy(&tmp[(size_t)at * stride], xi + (size_t)s * stride, (size_t)stride * sizeof(double));
    }
    std::memcpy(xi, tmp.data(), tmp.size() * sizeof(double));
}

int enrich_local_count(int dim, int scheme) {
    const int per_axis = scheme == kEnrichLagrange1 ? 2 : 3;
    return dim == 3 ? per_axis * per_axis * per_axis : per_axis * per_axis;
}

int enrich_local_index(int dim, int scheme, const int* axis_offset) {
    const int per_axis = scheme == kEnrichLagrange1 ? 2 : 3;
    return axis_offset[0] + per_axis * (axis_offset[1] + (dim == 3 ? per_axis * axis_offset[2] : 0));
}

bool enrich_numbering(const EnrichInput& in, std::vector<int32_t>& adof, int32_t* n_adofs, std::string* err) {
    const int dim = in.dim, nloc = enrich_local_count(dim, in.scheme);
    const int nc = 1 << dim;
    int64_t nbas[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d)
        nbas[d] = in.scheme == kEnrichLagrange1 ? in.n[d] + 1 : in.scheme == kEnrichLagrange27 ? 2 * in.n[d] + 1 : in.n[d] + 2;
    // support cells of basis index m along one axis: [lo, hi]
    auto axis_support = [&](int d, int64_t m, int64_t& lo, int64_t& hi) {
        if (in.scheme == kEnrichLagrange1) {
            lo = m - 1;
            hi = m;
        } else if (in.scheme == kEnrichLagrange27) {
            lo = (m & 1) ? m / 2 : m / 2 - 1;
            hi = m / 2;
        } else {
            lo = m - 2;
            hi = m;
        }
        lo = std::max<int64_t>(lo, 0);
        hi = std::min<int64_t>(hi, in.n[d] - 1);
    };
    auto axis_local = [&](int64_t m, int64_t cell) {
        return (int)(in.scheme == kEnrichLagrange27 ? m - 2 * cell : m - cell);
    };
    adof.assign((size_t)in.n_el * nloc, -1);
    int32_t next = 0;
    std::vector<int64_t> cells;
    std::vector<int32_t> sp_elem;       // elements (subphases) in support, in order
    std::vector<int32_t> cell_sp0;      // first support-list position of each support cell
    std::vector<int> parent, level_of_root;
    // corner pairs across the face between a cell and its +d neighbour
    int face_pair[3][4][2];
    for (int d = 0; d < dim; ++d) {
        int m = 0;
        for (int v = 0; v < nc; ++v) {
            int o[3];
            corner_offset(dim, v, o);
            if (o[d] != 1) continue;
            o[d] = 0;
            face_pair[d][m][0] = v;
            face_pair[d][m][1] = corner_of_offset(dim, o);
            ++m;
        }
    }
    const int n_face_corners = nc / 2;
    const int64_t nx = in.n[0], ny = in.n[1];
    for (int64_t bz = 0; bz < nbas[2]; ++bz)
        for (int64_t by = 0; by < nbas[1]; ++by)
            for (int64_t bx = 0; bx < nbas[0]; ++bx) {
                const int64_t bm[3] = {bx, by, bz};
                int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                bool empty = false;
                for (int d = 0; d < dim; ++d) {
                    axis_support(d, bm[d], lo[d], hi[d]);
                    empty = empty || lo[d] > hi[d];
                }
                if (empty) continue;
                cells.clear();
                sp_elem.clear();
                cell_sp0.clear();
                for (int64_t k = lo[2]; k <= hi[2]; ++k)
                    for (int64_t j = lo[1]; j <= hi[1]; ++j)
                        for (int64_t i = lo[0]; i <= hi[0]; ++i) {
                            const int64_t c = i + nx * (j + ny * k);
                            if (in.cell_first[c] < 0) continue;
                            cells.push_back(c);
                            cell_sp0.push_back((int32_t)sp_elem.size());
                            for (int32_t q = 0; q < in.cell_count[c]; ++q) sp_elem.push_back(in.cell_first[c] + q);
                        }
                if (cells.empty()) continue;
                const int nsp = (int)sp_elem.size();
                parent.resize((size_t)nsp);
                std::iota(parent.begin(), parent.end(), 0);
                // subphase graph pruned to the support: face neighbours within the support
                for (size_t a = 0; a < cells.size(); ++a) {
                    const int64_t c = cells[a];
                    const int64_t cijk[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
                    for (int d = 0; d < dim; ++d) {
                        if (cijk[d] + 1 > hi[d]) continue;
                        const int64_t stride = d == 0 ? 1 : d == 1 ? nx : nx * ny;
                        const int64_t cn = c + stride;
                        const auto it = std::lower_bound(cells.begin(), cells.end(), cn);
                        if (it == cells.end() || *it != cn) continue;
                        const size_t b = (size_t)(it - cells.begin());
                        for (int f = 0; f < n_face_corners; ++f) {
                            const int va = face_pair[d][f][0], vb = face_pair[d][f][1];
                            int sa = -1, sb = -1;
                            for (int32_t q = 0; q < in.cell_count[c] && sa < 0; ++q)
                                if (in.elem_corner[in.cell_first[c] + q] >> va & 1u) sa = cell_sp0[a] + q;
                            for (int32_t q = 0; q < in.cell_count[cn] && sb < 0; ++q)
                                if (in.elem_corner[in.cell_first[cn] + q] >> vb & 1u) sb = cell_sp0[b] + q;
                            if (sa >= 0 && sb >= 0) unite(parent, sa, sb);
                        }
                    }
                }
                // enrichment levels in order of first appearance
                level_of_root.assign((size_t)nsp, -1);
                int n_lev = 0;
                for (int s = 0; s < nsp; ++s) {
                    const int r = find_root(parent, s);
                    if (level_of_root[(size_t)r] < 0) level_of_root[(size_t)r] = n_lev++;
                }
                if (n_lev > 255) {
                    *err = "a basis has more than 255 enrichment levels";
                    return false;
                }
                for (size_t a = 0; a < cells.size(); ++a) {
                    const int64_t c = cells[a];
                    const int64_t cijk[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
                    int off[3] = {0, 0, 0};
                    for (int d = 0; d < dim; ++d) off[d] = axis_local(bm[d], cijk[d]);
                    const int loc = enrich_local_index(dim, in.scheme, off);
                    for (int32_t q = 0; q < in.cell_count[c]; ++q) {
                        const int s = cell_sp0[a] + q;
                        adof[(size_t)(in.cell_first[c] + q) * nloc + loc] = next + level_of_root[(size_t)find_root(parent, s)];
                    }
                }
                next += n_lev;
            }
    *n_adofs = next;
    return true;
}

}  // namespace mgpu
