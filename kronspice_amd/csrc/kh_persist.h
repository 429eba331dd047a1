// kh_persist.h — the persistent batched-analysis kernel (k_persist).
//
// A launch covers every group of an uneven batch: workgroup w runs circuits
// of group bgroup[w] (kh_group.h); `cpb` circuits run in each one-wavefront
// workgroup, one per lane. Each circuit loops: factor + solve its stamped
// system, then advance its analysis to the next solve (run_circuit).
//
// Groups with lds_bytes > 0 run staged: the wavefront copies the staged
// arrays of each of its circuits (all per-circuit arrays when they fit, else
// the most used ones, plan_staging in kh_group.h) into LDS, the circuits run
// on the LDS copies and their slices of the rest, and the wavefront copies
// back what outlives the launch. The dependent accesses to staged arrays are
// then LDS accesses instead of trips to L2/HBM.
//
// The kernel is compiled once per register budget W (resident wavefronts per
// SIMD the allocation must allow): 1 leaves the allocator free; 2 and 4 trade
// spills for more circuits in flight, whose latencies then overlap. Each
// budget is its own translation unit (kh_persist.hip with KH_PERSIST_WAVES),
// so the device functions the kernel calls are allocated under that budget.
// It is also compiled once per set of element kinds (KH_KINDS, kh_devices.h):
// a batch of diode circuits runs a kernel without the transistor models.
#pragma once
#include <hip/hip_runtime.h>

#include "kh_group.h"

namespace kh {

constexpr int kBlock = 64;  // one wavefront per workgroup: lanes run long, divergent programs

// Kind sets k_persist is compiled for (R, C, L, V and I are in every one):
// diodes, level-1 MOSFETs, BJTs, and everything.
constexpr unsigned kKindsD = 126, kKindsM = 190, kKindsQ = 318, kKindsAll = KH_KINDS_ALL;

// The smallest compiled kind set holding `kinds`.
inline unsigned persist_kinds(unsigned kinds) {
  for (unsigned v : {kKindsD, kKindsM, kKindsQ})
    if ((kinds & ~v) == 0) return v;
  return kKindsAll;
}

// Launches k_persist compiled for `waves` (1, 2 or 4) resident wavefronts per
// SIMD and the kind set persist_kinds(kinds) on the null stream; returns the
// launch error.
hipError_t kh_launch_persist(int waves, unsigned kinds, unsigned grid, size_t lds_bytes, const Group* G,
                             const int32_t* bgroup, int32_t max_solves, int32_t* counters);

}  // namespace kh
