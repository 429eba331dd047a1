// Host emulation of the HIP device built-ins the wavefront-per-circuit kernel
// (kronspice_amd/csrc/kh_wave.h) uses, so that its device code compiles with
// g++ and runs on the CPU. The 64 lanes of one wavefront are 64 fibers on one
// thread; every wavefront-wide operation (barrier, ballot, shuffle) is a
// barrier at which each lane hands over to the next, so between two barriers
// the lanes run one after another, and no lane passes a barrier before every
// lane has reached it. Test infrastructure only (tests/wave_host.cpp).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#ifndef KH_HD
#define KH_HD
#endif

typedef int hipError_t;

struct kh_emu_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
inline kh_emu_dim3 threadIdx;

// Saves the callee-saved registers on the current stack, stores the stack
// pointer in *save and resumes the context whose stack pointer is next.
extern "C" void kh_emu_switch(void** save, void* next);
asm(R"(
  .text
  .globl kh_emu_switch
  .type kh_emu_switch, @function
kh_emu_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
  .size kh_emu_switch, .-kh_emu_switch
)");

namespace kh_emu {

constexpr int kLanes = 64;
constexpr size_t kStack = (size_t)2 << 20;

struct Wave {
  void* sp[kLanes] = {};
  void* main_sp = nullptr;
  std::vector<char> stacks;
  void (*fn)(void*) = nullptr;
  void* arg = nullptr;
  int cur = 0;
  int done = 0;
};
inline Wave* g_wave = nullptr;
inline uint64_t slots[kLanes];

inline void switch_to(int from, int to) {
  g_wave->cur = to;
  threadIdx.x = (unsigned)to;
  kh_emu_switch(&g_wave->sp[from], g_wave->sp[to]);
}

// A lane's body: runs the wavefront function, then hands over to the next
// lane, the last one back to the caller of run().
extern "C" inline void kh_emu_lane_entry() {
  Wave* w = g_wave;
  const int me = w->cur;
  w->fn(w->arg);
  w->done++;
  if (me + 1 < kLanes) {
    g_wave->cur = me + 1;
    threadIdx.x = (unsigned)(me + 1);
    kh_emu_switch(&w->sp[me], w->sp[me + 1]);
  } else {
    kh_emu_switch(&w->sp[me], w->main_sp);
  }
  abort();  // a finished lane is never resumed
}

// Runs fn(arg) as the 64 lanes of one wavefront.
inline void run(void (*fn)(void*), void* arg) {
  Wave w;
  w.fn = fn;
  w.arg = arg;
  w.stacks.assign(kStack * kLanes, 0);
  for (int l = 0; l < kLanes; l++) {
    uintptr_t top = (uintptr_t)(w.stacks.data() + kStack * (l + 1));
    top &= ~(uintptr_t)15;
    void** s = (void**)top;
    *--s = nullptr;                          // alignment: entry sees rsp % 16 == 8 as after a call
    *--s = (void*)&kh_emu_lane_entry;        // the return address kh_emu_switch's ret takes
    for (int r = 0; r < 6; r++) *--s = nullptr;  // rbp rbx r12 r13 r14 r15
    w.sp[l] = s;
  }
  Wave* prev = g_wave;
  g_wave = &w;
  w.cur = 0;
  threadIdx.x = 0;
  kh_emu_switch(&w.main_sp, w.sp[0]);
  g_wave = prev;
}

// Barrier: hand over to the next lane; the last lane hands over to lane 0.
inline void barrier() {
  const int me = g_wave->cur;
  switch_to(me, (me + 1) % kLanes);
}

}  // namespace kh_emu

inline void __syncthreads() { kh_emu::barrier(); }

inline unsigned long long __ballot(int v) {
  kh_emu::slots[threadIdx.x] = v ? 1 : 0;
  __syncthreads();
  unsigned long long m = 0;
  for (int i = 0; i < 64; i++)
    if (kh_emu::slots[i]) m |= 1ull << i;
  __syncthreads();
  return m;
}

template <class T>
T __shfl(T v, int src, int width = 64) {
  static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
  (void)width;
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  kh_emu::slots[threadIdx.x] = u;
  __syncthreads();
  const uint64_t r = kh_emu::slots[src & 63];
  __syncthreads();
  T out;
  std::memcpy(&out, &r, sizeof(T));
  return out;
}
template <class T>
T __shfl_xor(T v, int mask, int width = 64) {
  return __shfl(v, (int)(threadIdx.x ^ (unsigned)mask), width);
}

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
template <class T>
inline T min(T a, T b) {
  return b < a ? b : a;
}
template <class T>
inline T max(T a, T b) {
  return a < b ? b : a;
}
inline int atomicAdd(int* p, int v) {
  const int old = *p;
  *p = old + v;
  return old;
}
