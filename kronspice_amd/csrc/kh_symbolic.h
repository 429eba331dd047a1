// kh_symbolic.h — host-side symbolic phase shared by every circuit of a
// topology: value-slot layout of the stamps, minimum-degree ordering and the
// permuted column structure of each sparsity pattern.
#pragma once
#include <stdint.h>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/kronhip.h"
#include "kh_lu.h"

namespace kh {

// Exact minimum degree on the off-diagonal pattern of A + A^T, ties to the
// lowest index, eliminated vertices' neighbours joined into a clique
// (solver/sparse.go:412-451). Produces the same order as the reference's
// O(n^2) scan with an ordered (degree, index) set.
std::vector<int32_t> minimum_degree(int n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci);

// Structure of one pattern: ordering and permuted CSC (solver/sparse.go:154-187).
struct Pattern {
  std::vector<int32_t> q, qinv, ap, arow, aslot;
  int64_t fill_estimate = 0;  // structural |L| + |U| with diagonal pivots where possible
};

// Analyze the entries (ri[s], ci[s]) for the slots s listed in `slots`;
// aslot holds slot ids (indices into the lane's value array).
Pattern analyze_pattern(int n, const std::vector<int32_t>& slot_i, const std::vector<int32_t>& slot_j,
                        const std::vector<int32_t>& slots);

// Pool of analyzed patterns with their device descriptors.
struct PatternPool {
  std::vector<PatDesc> desc;
  std::vector<int32_t> pool;
  std::unordered_map<std::string, int32_t> by_key;
  int64_t max_fill = 0;
  int32_t add(const Pattern& p, const std::string& key);
};

// Topology: the element table of include/kronhip.h turned into the device
// element table (kinds, terminal unknowns, parameter and state offsets, slot
// tables), with the union of every value slot any Load can stamp.
struct Topology {
  int32_t num_nodes = 0, n = 0, nel = 0, nparams = 0, nstate = 0;
  std::vector<int32_t> et;         // [nel][ET_W]
  std::vector<int32_t> slot_i, slot_j, diag_slot, isV;
  int32_t nslots = 0;
  bool nonlin = false;
  uint32_t kinds = 0;  // bit k set when an element of kind k is present (kh_devices.h)
  // Element indices, in element order, of the nonlinear devices (D, M, Q), the
  // devices with stored charge or flux (C, L, D, M, Q) and the sources (V, I):
  // the loops that only concern them skip the rest.
  std::vector<int32_t> nl_idx, rx_idx, src_idx;
  // Parameter rows holding a MOSFET level, and the level circuit 0 has there:
  // the level is structural (it picks the compiled currents), so every
  // circuit of the topology must share it.
  std::vector<int32_t> level_rows;
  std::vector<double> level_vals;
  // Parameter rows whose sign decides an internal node (diode RS, MOSFET RD
  // and RS, BJT RC, RB and RE as conductances), and whether circuit 0 has that
  // node: the unknown layout follows circuit 0, so every circuit of the
  // topology must agree.
  std::vector<int32_t> layout_rows;
  std::vector<uint8_t> layout_vals;
  std::string error;
  bool build(int32_t numNodes, const kh_element* elems, int32_t n_elems, int32_t np, const double* params_c0);
};

// Stamping plan of the wavefront form (kh_wave.h, the fields of Circ that
// name it): consecutive elements in chunks of at most 64 whose records
// (record_capacity entries each) fit rec_limit entries; per chunk the targets
// its elements can touch — every slot of an element's slot table and every
// terminal's right-hand-side row (nslots + row) — each with the chunk's
// elements that can touch it, in element order. ro_*: the slots of each row
// in the order AbsProduct accumulates them, slot order for the sparse
// backend (solver/sparse.go:481-490) and row-major for the dense one
// (solver/dense.go:63-73).
struct WavePlan {
  std::vector<int32_t> ch_el, ch_t, tgt, tel_ptr, tel, rec_off;
  std::vector<int32_t> ro_ptr_sparse, ro_slot_sparse, ro_ptr_dense, ro_slot_dense;
  int32_t rec_cap = 0;  // the largest chunk's record entries
};
WavePlan build_wave_plan(const Topology& T, int rec_limit);

}  // namespace kh
