// Host check of the scatter mode of HEX27 sets (no GPU): every element matrix entry lands in its
// own CSR row at the column of its adof, and no two elements of one colour share an adof.
// Usage: scatter_check N REFINE   (N^3 lattice, gyroid cut, p = 2 B-splines, HEX27 cells).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <chrono>
#include "moris_gpu.h"
#include "set_internal.hpp"
int main(int argc, char** argv) {
    const int n = atoi(argv[1]), refine = atoi(argv[2]);
    mgpu_problem_spec s{};
    s.dim = 3; s.n[0] = s.n[1] = s.n[2] = n;
    s.hi[0] = s.hi[1] = s.hi[2] = 1.0;
    s.iwg = MGPU_IWG_STRUC_LINEAR_BULK; s.ls_kind = MGPU_LS_GYROID; s.gyroid_scale = 2 * M_PI * n / 4.0;
    s.material_phase = 1; s.E = 1.0; s.nu = 0.3; s.seed = 20261202;
    s.bspline_order = 2; s.lagrange_order = 2; s.refine_levels = refine;
    auto t0 = std::chrono::steady_clock::now();
    mgpu_problem* p = nullptr;
    if (mgpu_problem_build(&s, &p)) { printf("build failed\n"); return 1; }
    mgpu_set_desc d;
    mgpu_problem_desc(p, &d);
    auto t1 = std::chrono::steady_clock::now();
    setenv("MGPU_GENERIC_SCATTER", "1", 1);
    mgpu::SetHost h;
    std::string err;
    if (mgpu::build_set_host(d, h, &err) != MGPU_OK) { printf("set: %s\n", err.c_str()); return 1; }
    auto t2 = std::chrono::steady_clock::now();
    printf("problem %.1fs set %.1fs scatter %d elements %d colours %zu rows %d nnz %zu rank entries %zu\n",
           std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
           (int)h.scatter, h.n_el, h.colour_ptr.size() - 1, h.n_rows, h.col_idx.size(), h.rank.size());
    if (!h.scatter) return 1;
    const int nf = h.nf;
    // permutation
    std::vector<char> hit(h.n_el, 0);
    for (int32_t e : h.colour_el) { if (e < 0 || e >= h.n_el || hit[e]) { printf("bad perm\n"); return 1; } hit[e] = 1; }
    // colour classes share no adof; ascending within a class
    std::vector<int32_t> mark(h.n_adofs_type, -1);
    size_t cmin = 1 << 30, cmax = 0;
    for (size_t c = 0; c + 1 < h.colour_ptr.size(); ++c) {
        cmin = std::min(cmin, (size_t)(h.colour_ptr[c + 1] - h.colour_ptr[c]));
        cmax = std::max(cmax, (size_t)(h.colour_ptr[c + 1] - h.colour_ptr[c]));
        for (int32_t k = h.colour_ptr[c]; k < h.colour_ptr[c + 1]; ++k) {
            const int32_t e = h.colour_el[k];
            const int na = h.na_t[e];
            for (int x = 0; x < na; ++x) {
                const int32_t a = h.topo[h.topo_ptr[e] + x];
                if (mark[a] == (int32_t)c) { printf("colour %zu shares adof %d\n", c, a); return 1; }
                mark[a] = (int32_t)c;
            }
        }
    }
    // positions
    int64_t checked = 0;
    for (int32_t e = 0; e < h.n_el; ++e) {
        const int na = h.na_t[e];
        const int32_t* rows = h.topo.data() + h.topo_ptr[e];
        const uint16_t* rk = h.rank.data() + h.rank_off[e];
        if (h.topo_ptr[e + 1] - h.topo_ptr[e] != (int64_t)nf * na) { printf("topo size\n"); return 1; }
        if (h.rank_off[e + 1] - h.rank_off[e] != (int64_t)na * na) { printf("rank size\n"); return 1; }
        for (int f = 0; f < nf; ++f)
            for (int hb = 0; hb < nf; ++hb)
                for (int x = 0; x < na; ++x)
                    for (int y = 0; y < na; ++y) {
                        const int32_t row = rows[f * na + x];
                        if (row < 0 || row >= h.n_rows) { printf("row oob\n"); return 1; }
                        const int32_t r0 = h.row_ptr[row], r1 = h.row_ptr[row + 1], len = (r1 - r0) / nf;
                        if ((r1 - r0) % nf) { printf("row len\n"); return 1; }
                        const int64_t pos = (int64_t)r0 + hb * len + rk[x * na + y];
                        if (pos < r0 || pos >= r1 || rk[x * na + y] >= len) { printf("pos oob e %d\n", e); return 1; }
                        if (h.col_idx[pos] != rows[hb * na + y]) { printf("col mismatch e %d\n", e); return 1; }
                        ++checked;
                    }
    }
    printf("OK: %lld entries checked, colour sizes %zu..%zu, max na_t %d\n", (long long)checked, cmin, cmax, h.na_t_max);
    mgpu_problem_destroy(p);
    return 0;
}
