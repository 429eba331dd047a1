// CPU check of k_persist's LDS staging plan (kronspice_amd/csrc/kh_group.h:
// plan_staging, for_each_staged, lds_view) on groups of many shapes and LDS
// budgets, compiled for the host. For every circuit of the group:
//   * the staged arrays' views lie inside the circuit's LDS slice
//     [base, base + lds_bytes), do not overlap, and sit at the offsets the
//     kernel's copy loops use (for_each_staged);
//   * every array left in device memory is viewed at circuit cc's slice, the
//     same address circuit_view gives the in-place kernel;
//   * the plan respects the budget, stages everything when everything fits,
//     takes ranks in order, and keeps the count of 8-byte words odd.
// Prints "ok <cases>" and exits 0, or the first violation and exits 1.
// Test infrastructure only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kh_group.h"

using namespace kh;

namespace {

int fail(const char* what, int a, long long b, long long c) {
  printf("FAIL %s (%d, %lld, %lld)\n", what, a, b, c);
  return 1;
}

// Points every per-circuit array of g at a distinct fake device range for B
// circuits (never dereferenced); some LU arrays stay null when lu_null.
void fake_arrays(Group& g, int64_t B, bool lu_null) {
  uintptr_t next = 0x10000000;
  for_each_array(g, [&](auto& p, int64_t bytes, bool, bool, int) {
    using Ptr = std::remove_reference_t<decltype(p)>;
    p = (Ptr)next;
    next += (uintptr_t)((bytes * B + 4095) & ~(int64_t)4095) + 4096;
  });
  if (lu_null) {
    g.lu.lx = nullptr; g.lu.li = nullptr; g.lu.ux = nullptr; g.lu.ui = nullptr; g.lu.rch = nullptr;
  }
  g.lu.val = g.ctl.val;
  g.lu.rhs = g.ctl.rhs;
  g.lu.x = g.ctl.x;
}

int check(const Group& g0, int64_t budget, int64_t B, int* cases) {
  Group g = g0;
  plan_staging(g, budget);
  const int64_t full = lds_need(g);
  if (g.lds_bytes < 0 || g.lds_bytes > budget) return fail("lds_bytes outside the budget", g.lds_bytes, budget, full);
  if (g.lds_bytes > 0 && ((g.lds_bytes / 8) & 1) == 0) return fail("even word count", g.lds_bytes, budget, 0);
  if (full + 8 <= budget && !g.stage_all) return fail("everything fits but not all staged", 0, budget, full);
  if (g.stage_all && full + 8 > budget) return fail("all staged beyond the budget", 0, budget, full);
  // ranks in order: an unstaged allocated array of rank r that would have
  // fitted into the room left at the end means a rank was skipped
  int64_t staged = 0;
  for_each_staged(g, g.stage_mask, [&](const auto&, int64_t bytes, bool, bool, int64_t) { staged += round8(bytes); });
  if (g.lds_bytes > 0 && staged > g.lds_bytes) return fail("staged bytes above lds_bytes", 0, staged, g.lds_bytes);
  if (g.lds_bytes > 0 && !g.stage_all) {
    int i = 0;
    bool bad = false;
    for_each_array(g, [&](const auto& p, int64_t bytes, bool, bool, int) {
      const bool st = (g.stage_mask >> i) & 1;
      if (!st && p && staged + round8(bytes) + 8 <= budget) bad = true;
      if (st && !p) bad = true;
      i++;
    });
    if (bad) return fail("an array left out that fits, or a null array staged", 0, staged, budget);
  }
  if (g.lds_bytes == 0) {
    (*cases)++;
    return 0;
  }
  std::vector<char> lds((size_t)g.lds_bytes * 3);
  for (int64_t cc : {(int64_t)0, (int64_t)1, B - 1}) {
    char* base = lds.data() + (size_t)g.lds_bytes;
    Group v = lds_view(g, base, cc);
    CtlArgs a;
    LuArgs l;
    circuit_view(g, cc, a, l);
    Group ref = g;
    ref.ctl = a;
    ref.lu = l;
    // walk the view, the in-place reference and the staged offsets together
    std::vector<uintptr_t> vp, rp, gp;
    std::vector<int64_t> sz;
    for_each_array(v, [&](auto& p, int64_t bytes, bool, bool, int) { vp.push_back((uintptr_t)p); sz.push_back(bytes); });
    for_each_array(ref, [&](auto& p, int64_t, bool, bool, int) { rp.push_back((uintptr_t)p); });
    for_each_array(g, [&](const auto& p, int64_t, bool, bool, int) { gp.push_back((uintptr_t)p); });
    std::vector<int64_t> offs(vp.size(), -1);
    {
      int i = 0, k = 0;
      int64_t dummy = 0;
      (void)dummy;
      for_each_array(g, [&](const auto&, int64_t, bool, bool, int) {
        if ((g.stage_mask >> i) & 1) offs[i] = -2;
        i++;
      });
      for_each_staged(g, g.stage_mask, [&](const auto&, int64_t, bool, bool, int64_t off) {
        while (k < (int)offs.size() && offs[k] != -2) k++;
        offs[k++] = off;
      });
    }
    if (vp.size() != rp.size()) return fail("array count differs", (int)vp.size(), (long long)rp.size(), 0);
    std::vector<std::pair<int64_t, int64_t>> spans;
    for (size_t i = 0; i < vp.size(); i++) {
      const bool st = (g.stage_mask >> i) & 1;
      if (st) {
        const int64_t off = (int64_t)(vp[i] - (uintptr_t)base);
        if (off < 0 || off + sz[i] > g.lds_bytes) return fail("staged view outside the slice", (int)i, off, sz[i]);
        if (off != offs[i]) return fail("view offset differs from the copy offset", (int)i, off, offs[i]);
        spans.push_back({off, off + sz[i]});
      } else if (gp[i] == 0 && vp[i] != 0) {
        return fail("unallocated array viewed as allocated", (int)i, (long long)vp[i], 0);
      } else if (gp[i] != 0 && vp[i] != rp[i]) {
        return fail("unstaged view differs from circuit_view", (int)i, (long long)vp[i], (long long)rp[i]);
      }
    }
    std::sort(spans.begin(), spans.end());
    for (size_t i = 1; i < spans.size(); i++)
      if (spans[i].first < spans[i - 1].second) return fail("staged arrays overlap", (int)i, spans[i].first, 0);
    if (v.lu.val != v.ctl.val || v.lu.rhs != v.ctl.rhs || v.lu.x != v.ctl.x)
      return fail("LU aliases of val/rhs/x differ", 0, 0, 0);
    if (g.stage_all && !g.ctl.ckt.dense) {
      // the factorization then addresses every one of its arrays as LDS
      const void* lu[] = {v.lu.pid, v.lu.val, v.lu.rhs, v.lu.x, v.lu.xw, v.lu.w, v.lu.udiag, v.lu.lx, v.lu.ux,
                          v.lu.li, v.lu.ui, v.lu.lp, v.lu.up, v.lu.pinv, v.lu.prow, v.lu.xi, v.lu.pstack,
                          v.lu.mark, v.lu.err_step, v.lu.rch, v.lu.rp, v.lu.dk, v.lu.cpid};
      for (int j = 0; j < (int)(sizeof(lu) / sizeof(lu[0])); j++) {
        const int64_t off = (int64_t)((uintptr_t)lu[j] - (uintptr_t)base);
        if (off < 0 || off >= g.lds_bytes) return fail("LU array outside LDS under stage_all", j, off, g.lds_bytes);
      }
    }
    (*cases)++;
  }
  return 0;
}

}  // namespace

int main() {
  int cases = 0;
  for (int dense = 0; dense < 2; dense++)
    for (int extras = 0; extras < 4; extras++)
      for (int n : {3, 35, 200, 2000})
        for (bool lu_null : {false, true}) {
          Group g = {};
          g.ctl.ckt.n = n;
          g.ctl.ckt.nslots = 9 * n + 5;
          g.ctl.ckt.numNodes = n - 1 > 1 ? n - 1 : 2;
          g.ctl.ckt.dense = dense;
          g.npar = 4 * n + 3;
          g.nstate = 7 * n + 1;
          g.lu.n = n;
          // the sparse workspace's capacities; the dense one holds an n x n matrix in lx
          g.lu.capL = dense ? (int64_t)n * n : 5 * n + 11;
          g.lu.capU = dense ? 0 : g.lu.capL;
          g.lu.capR = g.lu.capL + g.lu.capU + n;
          g.cpb = 16;
          const int64_t B = 4099;
          g.ctl.B = B;
          // lane_spec / nodeset present or not (for_each_array visits them only when set)
          g.ctl.lane_spec = (extras & 1) ? (double*)0x1 : nullptr;
          g.ctl.nodeset = (extras & 2) ? (double*)0x1 : nullptr;
          fake_arrays(g, B, lu_null && !dense);
          for (int64_t budget : {0LL, 40LL, 64LL, 100LL, 1000LL, 4096LL, 10240LL, 10232LL, 65536LL, 163840LL, 1LL << 24}) {
            const int r = check(g, budget, B, &cases);
            if (r) {
              printf("  case dense=%d extras=%d n=%d lu_null=%d budget=%lld\n", dense, extras, n, (int)lu_null,
                     (long long)budget);
              return r;
            }
          }
        }
  printf("ok %d\n", cases);
  return 0;
}
