// enrich.hpp -- XTK subphases of a cut cell and basis-cluster enrichment of a level-0 lattice.
//
//   subphases of a parent cell     the face-connected components of its material IG cells
//                                  (Cut_Integration_Mesh subphases: flood fill over the IG-cell
//                                  neighbourhood of one phase within a parent cell)
//   enrichment levels of a basis   Enrichment::perform_basis_cluster_enrichment,
//                                  cl_XTK_Enrichment.cpp:276-400: subphases in the basis' support
//                                  (get_subphase_clusters_in_support :622-650, support cells in
//                                  ascending index, subphases of a cell in order), the subphase
//                                  graph pruned to the support (:767-806), flood fill (:852-886,
//                                  fn_mesh_flood_fill.hpp: one level per component, in order of
//                                  first appearance)
//   enriched basis indices         assign_enriched_coefficients_identifiers, :1358-1510: basis
//                                  by basis in ascending index, levels consecutive
//
// Two material subphases of face-adjacent cells are neighbours when they share material facets
// on the common face.  On a node-hierarchy cut, every material piece of a face contains a corner
// node of the face strictly inside the material (the level set is linear on the face triangles
// and the face-centre value is the mean of the corner values), and within a cell exactly one
// subphase holds a given strictly-material corner (the sub-cells around a vertex form a
// face-connected star).  The neighbour relation is therefore: the subphase of cell A and the
// subphase of cell B that hold the same strictly-material corner of the common face.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mgpu {

// Split the material sub-cells of one cut cell into subphases, in place.
//   xi [n_sub][dim+1][dim]   sub-cell vertex parametric coordinates; reordered so that each
//                            subphase is contiguous (subphases in order of their first sub-cell,
//                            sub-cells in their original order within a subphase)
//   corner_region [2^dim]    region of the cell's corner nodes (0 / 1, 2 on the isocontour)
//   count [n_sp]             sub-cells per subphase
//   corner [n_sp]            bit v set: corner v (kHex8Off / kQuad4Off order) lies strictly in
//                            the material and is a vertex of the subphase
void split_subphases(int dim, int n_sub, double* xi, const int* corner_region, int phase,
                     std::vector<int32_t>& count, std::vector<uint8_t>& corner);

// Basis families of a level-0 lattice whose supports the enrichment walks.
enum EnrichScheme {
    kEnrichLagrange1 = 0,    // p = 1 Lagrange nodes: (n + 1) per axis, 2 cells per axis in support
    kEnrichLagrange27 = 1,   // HEX27 nodes on the half-step grid: (2n + 1) per axis
    kEnrichBspline2 = 2,     // quadratic B-splines: (n + 2) per axis, 3 cells per axis in support
};

struct EnrichInput {
    int dim;
    int scheme;
    int64_t n[3];                 // lattice cells per axis (1 for unused axes)
    const int32_t* cell_first;    // [n_cells] first element of the cell, -1 if none
    const int32_t* cell_count;    // [n_cells] elements (subphases) of the cell
    const uint8_t* elem_corner;   // [n_el] strictly-material corners of the element's subphase
    int64_t n_el;
};

// Bases of a cell (local index a + 3 b + 9 c over the per-axis offsets from the cell, 2 or 3 per
// axis) and their enriched adof ids.
int enrich_local_count(int dim, int scheme);
int enrich_local_index(int dim, int scheme, const int* axis_offset);

// Enriched adof of every (element, local basis): adof[e * local_count + local].  Adofs are
// numbered basis by basis in ascending basis index (x fastest), one per enrichment level.
// Returns false (with *err set) if a basis has more than 255 levels.
bool enrich_numbering(const EnrichInput& in, std::vector<int32_t>& adof, int32_t* n_adofs, std::string* err);

}  // namespace mgpu
