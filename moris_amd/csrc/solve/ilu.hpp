// ilu.hpp -- level-scheduled incomplete LU factors on the GPU (ilu.hip), shared with gmres.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

struct mgpu_ilu {
    int32_t n = 0;
    int64_t nnz = 0;          // entries of the factor pattern (the matrix pattern plus ILU(k) fill)
    int64_t nnz_in = 0;       // entries of the matrix pattern
    int32_t fill_level = 0;
    int32_t max_row = 0;      // longest row of the factor pattern
    int32_t* rp = nullptr;    // factor pattern [n + 1], [nnz]
    int32_t* ci = nullptr;
    int32_t* diag = nullptr;  // position of each row's diagonal
    int32_t* pos = nullptr;   // [nnz_in] position of each matrix entry in the factor pattern
    int32_t* rows_f = nullptr;  // rows by forward level (rows k < i of a row come first)
    int32_t* rows_b = nullptr;  // rows by backward level (rows k > i of a row come first)
    double* lu = nullptr;     // strict lower part: L (unit diagonal implied); diagonal and above: U
    int32_t* err = nullptr;   // zero-pivot flag of the last factorisation
    std::vector<int32_t> lev_f, lev_b;   // host: level offsets into rows_f / rows_b
    std::vector<void*> allocs;
    bool factored = false;
};

namespace mgpu {

// v <- (L U)^-1 v on the stream; every kernel returns at once while *skip != 0 (skip may be null).
void ilu_apply_launch(const mgpu_ilu* f, double* v, const int32_t* skip, hipStream_t s);

}  // namespace mgpu
