// kh_persist.hip — k_persist compiled for one register budget,
// KH_PERSIST_WAVES resident wavefronts per SIMD, and one set of element
// kinds, KH_KINDS (see kh_persist.h).
#include <hip/hip_runtime.h>

#define KH_CLOCK() ((uint64_t)clock64())
#include "kh_persist.h"

#ifndef KH_PERSIST_WAVES
#define KH_PERSIST_WAVES 1
#endif

namespace kh {
namespace {

__device__ inline void copy_wave(char* dst, const char* src, int64_t bytes) {
  const uintptr_t al = (uintptr_t)dst | (uintptr_t)src | (uintptr_t)bytes;
  if ((al & 7) == 0) {
    for (int64_t k = threadIdx.x; k < bytes / 8; k += kBlock) ((uint64_t*)dst)[k] = ((const uint64_t*)src)[k];
  } else if ((al & 3) == 0) {
    for (int64_t k = threadIdx.x; k < bytes / 4; k += kBlock) ((uint32_t*)dst)[k] = ((const uint32_t*)src)[k];
  } else {
    for (int64_t k = threadIdx.x; k < bytes; k += kBlock) dst[k] = src[k];
  }
}

__global__ void __launch_bounds__(kBlock, KH_PERSIST_WAVES) k_persist(const Group* G, const int32_t* bgroup,
                                                                       int32_t max_solves, int32_t* counters) {
  extern __shared__ __align__(16) char lds[];
  const Group& g = G[bgroup[blockIdx.x]];
  const int64_t c0 = (int64_t)(blockIdx.x - g.block0) * g.cpb;
  bool running = false, waiting = false;
  if (g.lds_bytes > 0) {
    const int nc = (int)min((int64_t)g.cpb, g.ctl.B - c0);  // uniform over the workgroup
    for (int t = 0; t < nc; t++) {
      char* base = lds + (int64_t)t * g.lds_bytes;
      for_each_staged(g, g.stage_mask, [&](const auto& p, int64_t bytes, bool in, bool, int64_t off) {
        if (in && p) copy_wave(base + off, (const char*)p + (c0 + t) * bytes, bytes);
      });
    }
    __syncthreads();
    if ((int)threadIdx.x < nc) {
      Group v = lds_view(g, lds + (int64_t)threadIdx.x * g.lds_bytes, c0 + threadIdx.x);
      // every array in LDS: the LU addresses it as LDS; otherwise generic
      // pointers reach the staged arrays and the rest alike
      if (g.stage_all)
        run_circuit<true>(v, 0, max_solves, &running, &waiting);
      else
        run_circuit<false>(v, 0, max_solves, &running, &waiting);
    }
    __syncthreads();
    for (int t = 0; t < nc; t++) {
      char* base = lds + (int64_t)t * g.lds_bytes;
      for_each_staged(g, g.stage_mask, [&](const auto& p, int64_t bytes, bool, bool out, int64_t off) {
        if (out && p) copy_wave((char*)p + (c0 + t) * bytes, base + off, bytes);
      });
    }
  } else {
    if ((int)threadIdx.x >= g.cpb) return;
    const int64_t cc = c0 + threadIdx.x;
    if (cc >= g.ctl.B) return;
    run_circuit(g, cc, max_solves, &running, &waiting);
  }
  if (running) atomicAdd(&counters[2 * g.gid], 1);
  if (waiting) atomicAdd(&counters[2 * g.gid + 1], 1);
}

}  // namespace

hipError_t KH_CAT(KH_CAT(kh_launch_persist_w, KH_PERSIST_WAVES), KH_CAT(_k, KH_KINDS))(
    unsigned grid, size_t lds_bytes, const Group* G, const int32_t* bgroup, int32_t max_solves, int32_t* counters) {
  static size_t opted = 64 * 1024;  // dynamic LDS above 64 KiB needs the kernel's opt-in
  if (lds_bytes > opted) {
    const hipError_t e = hipFuncSetAttribute((const void*)k_persist, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds_bytes);
    if (e != hipSuccess) return e;
    opted = lds_bytes;
  }
  hipLaunchKernelGGL(k_persist, dim3(grid), dim3(kBlock), lds_bytes, 0, G, bgroup, max_solves, counters);
  return hipGetLastError();
}

}  // namespace kh
