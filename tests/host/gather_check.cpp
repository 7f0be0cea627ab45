// Host check of the gather lists of generic sets (no GPU).  The lists are fixed by three
// properties, each checked here independently of how they are built: every element-matrix entry
// k = ke_off[e] + x na + y appears exactly once; it sits in the list of the CSR value of row
// topo[x] whose column is topo[y]; each value's list ascends in k (the order of summation).
// Usage: gather_check N REFINE   (N^3 lattice, gyroid cut, p = 2 B-splines, HEX27 cells).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "moris_gpu.h"
#include "set_internal.hpp"
int main(int argc, char** argv) {
    const int n = atoi(argv[1]), refine = atoi(argv[2]);
    mgpu_problem_spec s{};
    s.dim = 3; s.n[0] = s.n[1] = s.n[2] = n;
    s.hi[0] = s.hi[1] = s.hi[2] = 1.0;
    s.iwg = MGPU_IWG_STRUC_LINEAR_BULK; s.ls_kind = MGPU_LS_GYROID; s.gyroid_scale = 2 * M_PI * n / 4.0;
    s.material_phase = 1; s.E = 1.0; s.nu = 0.3; s.seed = 20261203;
    s.bspline_order = 2; s.lagrange_order = 2; s.refine_levels = refine;
    mgpu_problem* p = nullptr;
    if (mgpu_problem_build(&s, &p)) { printf("build failed\n"); return 1; }
    mgpu_set_desc d;
    mgpu_problem_desc(p, &d);
    setenv("MGPU_GENERIC_SCATTER", "0", 1);
    mgpu::SetHost h;
    std::string err;
    if (mgpu::build_set_host(d, h, &err) != MGPU_OK) { printf("set: %s\n", err.c_str()); return 1; }
    if (!h.generic || h.scatter) { printf("not a gather-mode generic set\n"); return 1; }
    const int64_t nnz = (int64_t)h.col_idx.size(), n_src = (int64_t)h.ke_off[h.n_el];
    if ((int64_t)h.gather_ptr.size() != nnz + 1 || (int64_t)h.gather_src.size() != n_src ||
        h.gather_ptr[0] != 0 || (int64_t)h.gather_ptr[nnz] != n_src) { printf("bad sizes\n"); return 1; }
    // the value each entry belongs to, from the topology and the pattern
    std::vector<int64_t> value(n_src, -1);
    for (int32_t e = 0; e < h.n_el; ++e) {
        const int32_t* t = h.topo.data() + h.topo_ptr[e];
        const int64_t na = h.topo_ptr[e + 1] - h.topo_ptr[e];
        for (int64_t x = 0; x < na; ++x) {
            const int32_t* c0 = h.col_idx.data() + h.row_ptr[t[x]];
            const int32_t* c1 = h.col_idx.data() + h.row_ptr[t[x] + 1];
            for (int64_t y = 0; y < na; ++y) {
                const int32_t* c = std::lower_bound(c0, c1, t[y]);
                if (c == c1 || *c != t[y]) { printf("column %d missing from row %d\n", t[y], t[x]); return 1; }
                value[h.ke_off[e] + x * na + y] = c - h.col_idx.data();
            }
        }
    }
    std::vector<char> seen(n_src, 0);
    for (int64_t q = 0; q < nnz; ++q) {
        if (h.gather_ptr[q + 1] < h.gather_ptr[q]) { printf("gather_ptr descends at %lld\n", (long long)q); return 1; }
        for (int64_t i = h.gather_ptr[q]; i < h.gather_ptr[q + 1]; ++i) {
            const int64_t k = h.gather_src[i];
            if (k < 0 || k >= n_src || seen[k]) { printf("entry %lld out of range or repeated\n", (long long)k); return 1; }
            seen[k] = 1;
            if (value[k] != q) { printf("entry %lld in value %lld, belongs to %lld\n", (long long)k, (long long)q, (long long)value[k]); return 1; }
            if (i > h.gather_ptr[q] && h.gather_src[i - 1] >= k) { printf("value %lld not ascending\n", (long long)q); return 1; }
        }
    }
    printf("OK: %d elements, %lld values, %lld entries\n", h.n_el, (long long)nnz, (long long)n_src);
    return 0;
}
