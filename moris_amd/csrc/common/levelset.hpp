// levelset.hpp -- analytic GEN level-set fields evaluated at vertex coordinates.
//   sphere/circle  projects/GEN/src/ADV/field/cl_GEN_Sphere.hpp, cl_GEN_Circle.hpp
//   gyroid         projects/GEN/src/ADV/field/cl_GEN_Gyroid.hpp
//   plane          projects/GEN/src/ADV/field/cl_GEN_Plane.hpp
// Derived vertices (face / cell-centre nodes of the regular subdivision) are evaluated at their own
// coordinates: use_multilinear_interpolation defaults to false
// (projects/PRM/src/fn_PRM_GEN_Parameters.hpp:207, cl_GEN_Design_Field.cpp:112-140).
// Host and device share these functions (the GPU set-up evaluates the same expressions).
#pragma once
#include <cmath>
#include <cstdint>
#include "../../../include/moris_gpu.h"
#include "fem_tables.hpp"

namespace mgpu {

struct LevelSet {
    int32_t kind = MGPU_LS_NONE;
    int32_t dim = 3;
    int32_t n_spheres = 0;
    const double* spheres = nullptr;   // [n][4]
    double gyroid_scale = 1.0;
    double plane[4] = {0, 0, 0, 0};

    MGPU_HD double operator()(const double* x) const {
        switch (kind) {
            case MGPU_LS_SPHERES: {
                double best = INFINITY;
                for (int i = 0; i < n_spheres; ++i) {
                    const double* s = spheres + 4 * i;
                    double r2 = 0.0;
                    for (int d = 0; d < dim; ++d) r2 += (x[d] - s[d]) * (x[d] - s[d]);
                    const double v = std::sqrt(r2) - s[3];
                    if (v < best) best = v;
                }
                return best;
            }
            case MGPU_LS_GYROID: {
                const double a = gyroid_scale * x[0], b = gyroid_scale * x[1];
                const double c = dim == 3 ? gyroid_scale * x[2] : 0.0;
                return std::sin(a) * std::cos(b) + std::sin(b) * std::cos(c) + std::sin(c) * std::cos(a);
            }
            case MGPU_LS_PLANE: {
                double v = plane[3];
                for (int d = 0; d < dim; ++d) v += plane[d] * x[d];
                return v;
            }
            default:
                return 1.0;
        }
    }
};

// splitmix64 (Steele, Lea, Flood 2014) as a counter-based generator: value idx of stream seed.
MGPU_HD inline uint64_t splitmix64_at(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
MGPU_HD inline double uniform01_at(uint64_t seed, uint64_t idx) {
    return (double)(splitmix64_at(seed, idx) >> 11) * (1.0 / 9007199254740992.0);
}
// Standard normal of lattice cell `cell` (Box-Muller on draws 2*cell, 2*cell+1).
MGPU_HD inline double normal_at(uint64_t seed, uint64_t cell) {
    const double u1 = 1.0 - uniform01_at(seed, 2 * cell);   // (0, 1]
    const double u2 = uniform01_at(seed, 2 * cell + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace mgpu
