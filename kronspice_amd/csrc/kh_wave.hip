// kh_wave.hip — k_wave (kh_wave.h) compiled for one set of element kinds,
// KH_KINDS.
#include <hip/hip_runtime.h>

#define KH_CLOCK() ((uint64_t)clock64())
#include "kh_persist.h"
#include "kh_wave.h"

namespace kh {
namespace {

__global__ void __launch_bounds__(kWaveLanes) k_wave(const Group* G, const int32_t* gw, const int32_t* cw,
                                                     int32_t nwork, int32_t* next, int32_t max_solves,
                                                     int32_t* counters) {
  extern __shared__ __align__(16) char lds[];
  for (;;) {
    int w = 0;
    if (threadIdx.x == 0) w = atomicAdd(next, 1);
    w = __shfl(w, 0);
    if (w >= nwork) break;
    const Group& g = G[gw[w]];
    const int64_t c = cw[w];
    if (g.ctl.ci[c * NCI + CI_PHASE] == PH_DONE) continue;
    bool running = false, waiting = false;
    run_circuit_wave(g, c, max_solves, lds, &running, &waiting);
    if (threadIdx.x == 0) {
      if (running) atomicAdd(&counters[2 * g.gid], 1);
      if (waiting) atomicAdd(&counters[2 * g.gid + 1], 1);
    }
  }
}

}  // namespace

hipError_t KH_CAT(kh_launch_wave_k, KH_KINDS)(unsigned grid, size_t lds_bytes, const Group* G, const int32_t* gw,
                                             const int32_t* cw, int32_t nwork, int32_t* next, int32_t max_solves,
                                             int32_t* counters) {
  static size_t opted = 64 * 1024;  // dynamic LDS above 64 KiB needs the kernel's opt-in
  if (lds_bytes > opted) {
    const hipError_t e =
        hipFuncSetAttribute((const void*)k_wave, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    opted = lds_bytes;
  }
  hipLaunchKernelGGL(k_wave, dim3(grid), dim3(kWaveLanes), lds_bytes, 0, G, gw, cw, nwork, next, max_solves,
                     counters);
  return hipGetLastError();
}

}  // namespace kh
