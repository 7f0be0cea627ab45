// set_state.hpp -- host-side state of an mgpu_set and the launch sequence shared by the single-set
// entry points (assemble.hip) and the captured batches (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../setup/internal.hpp"
#include "../learned/lq_view.hpp"
#include "set_device.hpp"
#include "set_internal.hpp"

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return mgpu::set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct mgpu_set {
    mgpu::SetHost h;
    mgpu::DevView v{};
    hipStream_t stream = nullptr;
    int precision = MGPU_PREC_FP64;
    int quadrature = MGPU_QUAD_SUBDIVISION;
    int64_t stats[5] = {0, 0, 0, 0, 0};
    int64_t n_ip_cut = 0;
    std::vector<int32_t> cut_nip;           // integration points of each cut element's sub-cells
    std::vector<int32_t> cut_ig0;           // first sub-cell of each cut element
    std::vector<void*> allocs;
    double* d_u_tmp = nullptr;
    double* d_vals_tmp = nullptr;
    double* d_rhs_tmp = nullptr;
    // generic element path (T rows with several entries, or no lattice)
    mgpu::GenView g{};
    // launch shape (autotuned or set by mgpu_set_tuning)
    int stencil_variant = -1;               // index into the stencil row-block sizes, -1 default
    int cut_order = 0;                      // 1: cut elements longest-first
    const int32_t* d_cut_order = nullptr;   // device permutation (built on first use)
    const int32_t* d_gen_order = nullptr;   // generic sets, gather mode: elements most sub-cells first
    // learned cut-cell quadrature
    std::vector<double> vphi_cut;      // host copy of the cut elements' nodal level sets
    mgpu::LqView lv{};
    const mgpu_mlp* mlp = nullptr;
    bool lq_ready = false;
    // a set made from the GPU set-up holds its pattern, topology and cell map on the device only;
    // this fills the host copies (h.row_ptr, col_idx, topo, vert_pos, cell_elem) on first need
    std::function<mgpu_status(mgpu_set*)> fetch_host;

    mgpu_set() = default;
    mgpu_set(const mgpu_set&) = delete;
    mgpu_set& operator=(const mgpu_set&) = delete;
    // device buffers go with the handle, also when mgpu_set_create fails half-way
    ~mgpu_set() {
        for (void* p : allocs) (void)hipFree(p);
        if (d_u_tmp) (void)hipFree(d_u_tmp);
        if (d_vals_tmp) (void)hipFree(d_vals_tmp);
        if (d_rhs_tmp) (void)hipFree(d_rhs_tmp);
    }
};

namespace mgpu {
// The host copies of the pattern and topology, fetched from the device the first time a host
// entry point needs them (sets made from the GPU set-up).
inline mgpu_status ensure_host(mgpu_set* S) {
    if (!S->fetch_host) return MGPU_OK;
    auto f = std::move(S->fetch_host);
    S->fetch_host = nullptr;
    return f(S);
}
// Buffers of a lattice set once its inputs are on the device (assemble.hip).
mgpu_status finish_lattice_set(mgpu_set* S, const int32_t* ig_ptr_host);
// Enqueue one assemble call of the set on S->stream (no synchronisation, no allocation, no memset:
// safe to capture into a graph).  ev (optional, 6 events): [0] start, [1] after the learned
// inputs/network, [2] after the moment fit, [3] after the cut-element integrands, [4] after the
// cell records, [5] after the stencil scatter.
mgpu_status launch_assemble(mgpu_set* S, const double* u, double* vals, double* rhs, hipEvent_t* ev);
// Synchronise the set's stream and turn the kernels' row-length flag into a status.
mgpu_status check_err_flag(mgpu_set* S);
// Generic element path: device buffers of the set (at creation) and one assemble call (events as
// launch_assemble; [1], [2] coincide with [0]).
mgpu_status generic_upload(mgpu_set* S, const mgpu_set_desc& d);
mgpu_status launch_generic(mgpu_set* S, const double* u, double* vals, double* rhs, hipEvent_t* ev);
// Rows per workgroup of stencil variant `variant` for nf fields; *n_variants: number of variants.
int stencil_shape(int nf, int variant, int* n_variants);
}  // namespace mgpu
