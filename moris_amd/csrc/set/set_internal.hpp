// set_internal.hpp -- the state behind an mgpu_set handle: host copies of the descriptor arrays the
// set needs after creation, the CSR pattern, and device buffers.  Mirrors what fem::Set and
// MSI::Equation_Set own (cl_FEM_Set.hpp:64, cl_MSI_Equation_Set.hpp:61-62) plus the SOL graph
// (cl_Sparse_Matrix_EpetraFECrs.cpp:195-240).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/moris_gpu.h"

namespace mgpu {

// Largest element topology per dof type of the generic path (element-local T and the T^T K T
// products are staged in LDS at this size).
constexpr int kGenericMaxAdofsPerType = 64;
// The same for HEX27 IP cells (27 bases): level-0 quadratic B-splines or Lagrange nodes give 27,
// truncated hierarchical rows up to about 40.
constexpr int kHex27MaxAdofsPerType = 64;
// HEX27 sets whose element matrices hold more entries than this (a K_e pool of 2 GiB and gather
// lists of as much again) take the scatter mode; MGPU_GENERIC_SCATTER=1 / =0 forces it on / off.
constexpr int64_t kGatherMaxEntries = (int64_t)1 << 28;

// Host-side products of set creation (pure C++, no HIP).
struct SetHost {
    int dim = 3;
    int iwg = MGPU_IWG_SPATIALDIFF_BULK;
    int nf = 1;            // fields (dof types)
    int nb = 8;            // bases per field
    int na = 8;            // element adofs = nf * nb
    int nm = 1;            // material parameters per element
    int32_t n_el = 0;
    int32_t n_adofs_type = 0;
    int32_t n_rows = 0;    // nf * n_adofs_type
    // IG sub-cells whose detJ is below -1e-6 and the integration points skipped on them
    // (Element_Bulk::compute_jacobian_and_residual, cl_FEM_Element_Bulk.cpp:291)
    int64_t n_skip_subcells = 0;
    int64_t n_skip_points = 0;
    std::vector<int32_t> topo;       // lattice path [n_el][na]; generic path by topo_ptr; sorted
                                     // unique per type, global ids, types concatenated
    std::vector<int32_t> row_ptr;    // [n_rows + 1]
    std::vector<int32_t> col_idx;    // [nnz]
    int64_t nnz = 0;                 // CSR values (also while col_idx is not fetched)
    // element vertex b -> position of its adof in the sorted per-type topology
    std::vector<uint8_t> vert_pos;   // [n_el][nb]
    // lattice (stencil) path
    bool lattice = false;
    mgpu_lattice lat{};
    int64_t n_cells = 0, n_nodes = 0;
    std::vector<int32_t> cell_elem;  // [n_cells] element or -1
    std::vector<int32_t> node_adof;  // [n_nodes] type-local adof or -1
    std::vector<int32_t> row_node;   // [n_adofs_type] lattice node of each adof
    // cut elements
    std::vector<int32_t> cut_elems;  // element ids with sub-cells
    std::vector<int32_t> elem_cut;   // [n_el] cut index or -1
    // generic path (T rows with several entries, or no lattice): element-local T and the gather
    // lists of the CSR values
    bool generic = false;
    int32_t na_t_max = 0;                // largest per-type element topology
    std::vector<int32_t> na_t;           // [n_el] adofs per dof type of each element
    std::vector<int64_t> topo_ptr;       // [n_el + 1] into topo (nf * na_t per element; both paths)
    std::vector<int64_t> ke_off;         // [n_el + 1] K_e offsets, (nf * na_t)^2 per element
    std::vector<int32_t> t_col;          // [n T entries] element-local column of each T entry
    std::vector<int64_t> gather_ptr;     // [nnz + 1]
    std::vector<int64_t> gather_src;     // [sum of K_e sizes] K_e entry summed into each value
    // scatter mode of the HEX27 kernel (no K_e pool, no gather lists): elements coloured so that
    // no two of one colour share an adof; each element adds its K_e into the CSR values
    bool scatter = false;
    std::vector<int32_t> colour_ptr;     // [n_colours + 1] into colour_el
    std::vector<int32_t> colour_el;      // [n_el] elements by colour, most sub-cells first within one
    std::vector<int64_t> rank_off;       // [n_el + 1] na_t^2 per element
    std::vector<uint16_t> rank;          // [x][y]: position of adof y in the sorted neighbours of adof x
};

// Build topology, pattern and the lattice maps from a descriptor.  Returns MGPU_OK or an error
// code with *err set.
mgpu_status build_set_host(const mgpu_set_desc& d, SetHost& h, std::string* err);

}  // namespace mgpu
