// extraction.cpp -- elemental extraction operators (T-matrices from the integration cells to the
// adofs) of a set descriptor, and their file output.
//
// The reference writes, per integration (IG) cell, the B-spline ids that interpolate into the cell
// and the weights that map them to the cell's vertices (Integration_Mesh::
// save_elemental_T_matrices_to_file and get_elemental_IG_to_BS_T_matrices,
// projects/MTK/src/mesh/cl_MTK_Integration_Mesh_write_T_matrices.cpp:38-160, 684-809; requested by
// Workflow_HMR_XTK::output_T_matrices, projects/WRK/src/cl_WRK_Workflow_HMR_XTK.cpp:542-575):
//   E_c[v][a] = sum_b N_b(xi_v) T[b][a],
// N the IP cell's Lagrange basis at the IG vertex's parametric coordinate xi_v and T the element's
// T rows.  The reference stores HDF5; HDF5 is not available to this library, so the same content is
// written in the binary layout documented at mgpu_write_extraction_operators (moris_gpu.h).  The
// IG cells of an element are its cut sub-cells, or for an uncut element the IP cell's box corners.
#include <algorithm>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../common/fem_tables.hpp"
#include "../setup/internal.hpp"

using namespace mgpu;

namespace {

// Lagrange basis of the IP cell at xi: QUAD4 / HEX8 (vertex order of kQuad4XiHat / kHex8XiHat) or
// HEX27 (tensor products of 1D quadratics in the order of kHex27Half).
void ip_basis(int dim, int nb, const double* xi, double* N) {
    if (nb == 27) {
        double l[3][3], dl[3];
        for (int d = 0; d < 3; ++d) lagrange_quadratic_1d(xi[d], l[d], dl);
        for (int a = 0; a < 27; ++a) N[a] = l[0][kHex27Half[a][0]] * l[1][kHex27Half[a][1]] * l[2][kHex27Half[a][2]];
        return;
    }
    if (dim == 3) {
        for (int k = 0; k < 8; ++k)
            N[k] = 0.125 * (1.0 + xi[0] * kHex8XiHat[k][0]) * (1.0 + xi[1] * kHex8XiHat[k][1]) *
                   (1.0 + xi[2] * kHex8XiHat[k][2]);
    } else {
        for (int k = 0; k < 4; ++k)
            N[k] = 0.25 * (1.0 + xi[0] * kQuad4XiHat[k][0]) * (1.0 + xi[1] * kQuad4XiHat[k][1]);
    }
}

struct Writer {
    FILE* f = nullptr;
    bool ok = true;
    template <class T>
    void put(const T* p, size_t n) {
        if (ok && n) ok = std::fwrite(p, sizeof(T), n, f) == n;
    }
    ~Writer() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" mgpu_status mgpu_write_extraction_operators(const mgpu_set_desc* d, const char* path, int64_t* n_cells) {
    MGPU_GUARD_BEGIN
    if (!d || !path) return set_error(MGPU_ERR_INVALID_ARG, "null argument");
    const int dim = d->spatial_dim;
    const int nb = d->n_bases;
    if ((dim != 2 && dim != 3) || !(nb == (dim == 3 ? 8 : 4) || (dim == 3 && nb == 27)))
        return set_error(MGPU_ERR_INVALID_ARG, "dimension and IP bases must be 2 / 4, 3 / 8 or 3 / 27");
    const int32_t ne = d->n_elements;
    if (ne < 0) return set_error(MGPU_ERR_INVALID_ARG, "negative element count");
    if (ne > 0 && (!d->t_ptr || !d->t_adof || !d->t_weight || !d->ig_ptr))
        return set_error(MGPU_ERR_INVALID_ARG, "descriptor array missing");
    Writer w;
    w.f = std::fopen(path, "wb");
    if (!w.f) return set_error(MGPU_ERR_INVALID_ARG, std::string("cannot open ") + path);
    int64_t count = 0;
    for (int32_t e = 0; e < ne; ++e) count += std::max<int32_t>(d->ig_ptr[e + 1] - d->ig_ptr[e], 1);
    const char magic[8] = {'M', 'G', 'P', 'U', 'E', 'X', 'T', '1'};
    const int32_t head[4] = {dim, nb, d->n_adofs_per_type, 0};
    w.put(magic, 8);
    w.put(head, 4);
    w.put(&count, 1);
    const int nv_cut = dim + 1, nv_box = dim == 3 ? 8 : 4;
    std::vector<int32_t> ids;
    std::vector<double> E, N((size_t)nb);
    for (int32_t e = 0; e < ne; ++e) {
        // the element's adofs, ascending, and its T as a dense [nb][na] block
        ids.clear();
        for (int32_t q = d->t_ptr[(int64_t)e * nb]; q < d->t_ptr[(int64_t)(e + 1) * nb]; ++q) ids.push_back(d->t_adof[q]);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        const int na = (int)ids.size();
        std::vector<double> T((size_t)nb * na, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int32_t q = d->t_ptr[(int64_t)e * nb + b]; q < d->t_ptr[(int64_t)e * nb + b + 1]; ++q) {
                const int a = (int)(std::lower_bound(ids.begin(), ids.end(), d->t_adof[q]) - ids.begin());
                T[(size_t)b * na + a] += d->t_weight[q];
            }
        const int32_t s0 = d->ig_ptr[e], s1 = d->ig_ptr[e + 1];
        const bool cut = s1 > s0;
        const int nv = cut ? nv_cut : nv_box;
        for (int32_t s = cut ? s0 : 0; s < (cut ? s1 : 1); ++s) {
            E.assign((size_t)nv * na, 0.0);
            for (int v = 0; v < nv; ++v) {
                double xi[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < dim; ++k)
                    xi[k] = cut ? d->ig_xi[((size_t)s * nv_cut + v) * dim + k] : (dim == 3 ? kHex8XiHat[v][k] : kQuad4XiHat[v][k]);
                ip_basis(dim, nb, xi, N.data());
                for (int b = 0; b < nb; ++b)
                    if (N[(size_t)b] != 0.0)
                        for (int a = 0; a < na; ++a) E[(size_t)v * na + a] += N[(size_t)b] * T[(size_t)b * na + a];
            }
            const int32_t rec[3] = {e, nv, na};
            w.put(rec, 3);
            w.put(ids.data(), ids.size());
            w.put(E.data(), E.size());
        }
    }
    if (!w.ok) return set_error(MGPU_ERR_INVALID_ARG, std::string("write failed: ") + path);
    if (n_cells) *n_cells = count;
    return MGPU_OK;
    MGPU_GUARD_END
}
