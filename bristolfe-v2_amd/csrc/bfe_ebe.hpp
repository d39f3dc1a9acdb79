// Host-side layout of the element-by-element stiffness operator (no device
// code here).
//
// K = sum_e s_e P_e' T[tau_e] P_e (the element matrices of
// code/internal/fn_build_global_matrices_v5.m:105-128 stored as templates,
// s_e the absorbing reduction of :116). Node v's rows of K u gather over the
// elements around v: slot k of node v holds the element in which v is the
// local vertex l_k of template tau_k, so one slot's template rows are the
// same for every node.
//
// On a structured mesh the vertices of every slot's element are also the
// same neighbours of v: they fall into a few neighbour classes (the node
// itself and its mesh neighbours), and node v's neighbour of class j is
// nbr[j][v]. The layout puts classes in order of the index offset w - v at
// an interior reference node, vertices within a slot in class order and
// slots in order of their class lists, so the (slot, vertex) -> class table
// does not depend on how the mesh generator ordered templates or vertices,
// and step kernels compiled for a known table (a "pattern") load each
// neighbour's u once and keep it in registers.
#pragma once

#include <cstdint>
#include <vector>

namespace bfe {

struct EbeLayout {
  int32_t npe = 0, dpn = 0;
  int32_t n_slot = 0;         // occupied (template, vertex) slots
  int32_t n_slot_padded = 0;  // table rows (multiple of the padding)
  int64_t n_nodes = 0;
  std::vector<int32_t> adj;      // n_slot_padded x n_nodes, element or -1
  std::vector<int32_t> nodes4;   // n_el x 4: element nodes (original order, padded)
  std::vector<double> srows;     // n_slot_padded x dpn x npe*dpn (original vertex order)
  // neighbour classes (empty when the mesh has no consistent class structure)
  int32_t n_cls = 0;
  int32_t self_cls = -1;         // class of the node itself
  std::vector<int64_t> cls_off;  // n_cls: index offset w - v of each class at the reference node
  // per node: bit k set if slot k is occupied; bit 31 set if some neighbour
  // is not at its class offset (the node must read nbr)
  std::vector<uint32_t> present;
  std::vector<int32_t> cls;      // n_slot x npe, class of canonical vertex b of slot k
  std::vector<int32_t> nbr;      // n_cls x n_nodes, node or -1
  std::vector<double> srows_c;   // n_slot_padded x dpn x npe*dpn (canonical vertex order)
};

// Builds the layout. Throws std::invalid_argument when the element operator
// cannot be laid out (two elements of one template at the same vertex of a
// node, more than max_slots slots).
EbeLayout build_ebe_layout(int64_t n_nodes, int32_t dpn, int64_t n_el, int32_t npe,
                           const int32_t* el_nodes, const int32_t* el_tmpl, int32_t n_tmpl,
                           const double* tmpl, int32_t pad, int32_t max_slots);

}  // namespace bfe
