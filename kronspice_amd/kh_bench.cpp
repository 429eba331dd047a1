// kh_bench — runs a batch file through kh_batch_run from plain C++ (no
// Python, no PyTorch in the process), for timing and for profilers.
//
//   kh_bench <batch file> [steps] [warmup] [staging] [circuits per wave] [waves per SIMD] [precision]
//            [circuit per wavefront: 1 = k_wave] [backend]
//   kh_bench --copy <MiB>    one device-to-device copy kernel of that size:
//                            the reference point for hardware byte counters
//
// The batch file is written by kronspice_amd.workloads.write_batch_file:
//   "KHB1", int32 ngroups, int32 transient, double step, stop, max_step,
//   then per group: int32 num_nodes, nel, nparams, B, has_lane_spec;
//   kh_element[nel]; double params[nparams*B] (SoA); double lane[4*B] if present.
// Prints one JSON line: circuits per second over the timed steps, the
// library's counters of the last step and the status histogram.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../include/kronhip.h"

namespace {

struct HostGroup {
  int32_t num_nodes = 0, nel = 0, nparams = 0, B = 0, has_lane = 0;
  std::vector<kh_element> el;
  std::vector<double> params, lane;
};

bool read_exact(FILE* f, void* p, size_t bytes) { return fread(p, 1, bytes, f) == bytes; }

int die(const char* what) {
  fprintf(stderr, "kh_bench: %s\n", what);
  return 1;
}

#define CHECK_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "kh_bench: %s: %s\n", #call, hipGetErrorString(e_));           \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

#define CHECK_KH(call)                                                               \
  do {                                                                               \
    kh_status s_ = (call);                                                           \
    if (s_ != KH_OK) {                                                               \
      fprintf(stderr, "kh_bench: %s: %s (%s)\n", #call, kh_status_string(s_), kh_last_error()); \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

__global__ void __launch_bounds__(256) k_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int copy_mode(int mib) {
  const int64_t n = (int64_t)mib * (1 << 20) / 8;
  double *a = nullptr, *b = nullptr;
  CHECK_HIP(hipMalloc(&a, n * 8));
  CHECK_HIP(hipMalloc(&b, n * 8));
  CHECK_HIP(hipMemset(a, 0, n * 8));
  CHECK_HIP(hipDeviceSynchronize());
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  CHECK_HIP(hipGetLastError());
  CHECK_HIP(hipDeviceSynchronize());
  printf("{\"copy_bytes_read\": %lld, \"copy_bytes_written\": %lld}\n", (long long)(n * 8), (long long)(n * 8));
  CHECK_HIP(hipFree(a));
  CHECK_HIP(hipFree(b));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return die("usage: kh_bench <batch file> [steps] [warmup] | kh_bench --copy <MiB>");
  if (strcmp(argv[1], "--copy") == 0) return copy_mode(argc > 2 ? atoi(argv[2]) : 256);
  const int steps = argc > 2 ? atoi(argv[2]) : 1;
  const int warmup = argc > 3 ? atoi(argv[3]) : 0;
  const int staging = argc > 4 ? atoi(argv[4]) : 0;
  const int cpw = argc > 5 ? atoi(argv[5]) : 0;
  const int waves = argc > 6 ? atoi(argv[6]) : 0;
  const int precision = argc > 7 ? atoi(argv[7]) : 0;
  const int circuit_wave = argc > 8 ? atoi(argv[8]) : 0;
  const int backend = argc > 9 ? atoi(argv[9]) : 0;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return die("cannot open the batch file");
  char magic[4];
  int32_t ngroups = 0, transient = 0;
  double spec3[3];
  if (!read_exact(f, magic, 4) || memcmp(magic, "KHB1", 4) != 0) return die("not a KHB1 batch file");
  if (!read_exact(f, &ngroups, 4) || !read_exact(f, &transient, 4) || !read_exact(f, spec3, sizeof(spec3)))
    return die("truncated header");
  if (ngroups <= 0 || ngroups > (1 << 20)) return die("bad group count");
  std::vector<HostGroup> hg(ngroups);
  int64_t circuits = 0;
  for (auto& g : hg) {
    int32_t h[5];
    if (!read_exact(f, h, sizeof(h))) return die("truncated group header");
    g.num_nodes = h[0]; g.nel = h[1]; g.nparams = h[2]; g.B = h[3]; g.has_lane = h[4];
    if (g.nel <= 0 || g.nparams <= 0 || g.B <= 0) return die("bad group sizes");
    g.el.resize(g.nel);
    g.params.resize((size_t)g.nparams * g.B);
    if (!read_exact(f, g.el.data(), g.el.size() * sizeof(kh_element)) ||
        !read_exact(f, g.params.data(), g.params.size() * 8))
      return die("truncated group");
    if (g.has_lane) {
      g.lane.resize((size_t)4 * g.B);
      if (!read_exact(f, g.lane.data(), g.lane.size() * 8)) return die("truncated lane spec");
    }
    circuits += g.B;
  }
  fclose(f);

  CHECK_KH(kh_init());
  std::vector<kh_topology> topo(ngroups);
  std::vector<kh_group> groups(ngroups);
  std::vector<void*> owned;
  auto dalloc = [&](size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e == hipSuccess) owned.push_back(*p);
    return e;
  };
  for (int g = 0; g < ngroups; g++) {
    HostGroup& H = hg[g];
    std::vector<double> p0(H.nparams);
    for (int k = 0; k < H.nparams; k++) p0[k] = H.params[(size_t)k * H.B];
    CHECK_KH(kh_topology_create(H.num_nodes, H.el.data(), H.nel, H.nparams, p0.data(), &topo[g]));
    const int32_t n = kh_topology_size(topo[g]);
    kh_group& G = groups[g];
    memset(&G, 0, sizeof(G));
    G.topology = topo[g];
    G.B = H.B;
    void *dp, *dl = nullptr, *dx, *dst, *dit, *dacc, *drej, *dled;
    CHECK_HIP(dalloc(H.params.size() * 8, &dp));
    CHECK_HIP(hipMemcpy(dp, H.params.data(), H.params.size() * 8, hipMemcpyHostToDevice));
    if (H.has_lane) {
      CHECK_HIP(dalloc(H.lane.size() * 8, &dl));
      CHECK_HIP(hipMemcpy(dl, H.lane.data(), H.lane.size() * 8, hipMemcpyHostToDevice));
    }
    CHECK_HIP(dalloc((size_t)n * H.B * 8, &dx));
    CHECK_HIP(dalloc((size_t)H.B * 4, &dst));
    CHECK_HIP(dalloc((size_t)H.B * 4, &dit));
    CHECK_HIP(dalloc((size_t)H.B * 4, &dacc));
    CHECK_HIP(dalloc((size_t)H.B * 4, &drej));
    CHECK_HIP(dalloc((size_t)H.B * 8, &dled));
    G.d_params = (const double*)dp;
    G.d_lane_spec = (const double*)dl;
    G.d_x = (double*)dx;
    G.d_status = (int32_t*)dst;
    G.d_iters = (int32_t*)dit;
    G.d_accepted = (int32_t*)dacc;
    G.d_rejected = (int32_t*)drej;
    G.d_ledger = (double*)dled;
  }
  kh_options opt;
  memset(&opt, 0, sizeof(opt));
  opt.staging = staging;
  opt.circuits_per_wave = cpw;
  opt.waves_per_simd = waves;
  opt.precision = precision;
  opt.circuit_wave = circuit_wave;
  opt.backend = backend;
  kh_tran_spec spec;
  memset(&spec, 0, sizeof(spec));
  spec.step = spec3[0];
  spec.stop = spec3[1];
  spec.max_step = spec3[2];
  const kh_tran_spec* sp = transient ? &spec : nullptr;

  kh_stats st;
  for (int w = 0; w < warmup; w++) CHECK_KH(kh_batch_run(ngroups, groups.data(), &opt, sp, &st));
  CHECK_HIP(hipDeviceSynchronize());
  double lu_s = 0, bytes = 0, cyc_lu = 0, cyc_ctl = 0, cyc_part[5] = {0, 0, 0, 0, 0};
  int64_t solves = 0, launches = 0, passes = 0, redos = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; s++) {
    CHECK_KH(kh_batch_run(ngroups, groups.data(), &opt, sp, &st));
    lu_s += st.lu_seconds;
    bytes += (double)st.bytes_algorithmic;
    solves += st.solves;
    launches += st.lu_launches;
    cyc_lu += st.cycles_lu;
    cyc_ctl += st.cycles_control;
    passes += st.refine_passes;
    redos += st.fp64_redos;
    cyc_part[0] += st.cycles_stamp;
    cyc_part[1] += st.cycles_converge;
    cyc_part[2] += st.cycles_lte;
    cyc_part[3] += st.cycles_accept;
    cyc_part[4] += st.cycles_ledger;
  }
  CHECK_HIP(hipDeviceSynchronize());
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  std::map<int, int64_t> hist;
  for (auto& G : groups) {
    std::vector<int32_t> s(G.B);
    CHECK_HIP(hipMemcpy(s.data(), G.d_status, (size_t)G.B * 4, hipMemcpyDeviceToHost));
    for (int32_t v : s) hist[v]++;
  }
  std::string h = "{";
  for (auto& kv : hist) h += (h.size() > 1 ? ", \"" : "\"") + std::to_string(kv.first) + "\": " + std::to_string(kv.second);
  h += "}";
  printf("{\"circuits\": %lld, \"groups\": %d, \"staging\": %d, \"circuits_per_wave\": %d, \"waves_per_simd\": %d, "
         "\"precision\": %d, \"circuit_wave\": %d, \"backend\": %d, \"refine_passes_per_step\": %.1f, \"fp64_redos_per_step\": %.1f, \"steps\": %d, "
         "\"seconds_per_step\": %.6f, \"circuits_per_s\": %.3f, "
         "\"solves_per_step\": %.1f, \"launches_per_step\": %.1f, \"kernel_seconds_per_step\": %.6f, "
         "\"algorithmic_GBps\": %.4f, \"cycles_lu\": %.6g, \"cycles_control\": %.6g, \"cycles_stamp\": %.6g, "
         "\"cycles_converge\": %.6g, \"cycles_lte\": %.6g, \"cycles_accept\": %.6g, \"cycles_ledger\": %.6g, "
         "\"staged_bytes\": %lld, \"staged_partial\": %lld, \"status_histogram\": %s}\n",
         (long long)circuits, ngroups, staging, cpw, waves, precision, circuit_wave, backend, (double)passes / steps, (double)redos / steps, steps, dt / steps, circuits * steps / dt, (double)solves / steps,
         (double)launches / steps, lu_s / steps, lu_s > 0 ? bytes / lu_s / 1e9 : 0.0, cyc_lu / steps, cyc_ctl / steps,
         cyc_part[0] / steps, cyc_part[1] / steps, cyc_part[2] / steps, cyc_part[3] / steps, cyc_part[4] / steps,
         (long long)st.staged_bytes, (long long)st.staged_partial, h.c_str());
  for (void* p : owned) (void)hipFree(p);
  for (auto t : topo) kh_topology_destroy(t);
  return 0;
}
