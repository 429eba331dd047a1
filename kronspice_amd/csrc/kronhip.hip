// kronhip.hip — gfx950 kernels and the C ABI of include/kronhip.h.
//
// Kernels (one lane = one circuit, SoA with the circuit index fastest):
//   k_control  — advances every circuit's analysis state machine up to its
//                next linear solve and stamps that system (kh_control.h)
//   k_lu       — factors and solves every stamped system (kh_lu.h)
//   k_hash     — fingerprints the sparsity pattern of lanes whose stamps
//                created new entries, for the host's symbolic phase
// The host loop relaunches k_control / k_lu until every circuit is done and
// runs the symbolic phase (kh_symbolic.cpp) once per distinct pattern.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/kronhip.h"
#include "kh_control.h"
#include "kh_symbolic.h"

using namespace kh;

namespace {

thread_local std::string g_last_error;

kh_status fail(kh_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

#define KH_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(KH_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

constexpr int kBlock = 64;  // one wavefront per workgroup: lanes run long, divergent programs

// ---- kernels ----------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_control(CtlArgs a) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.B) return;
  Ctl ctl(a, c);
  ctl.step();
  ctl.store();
  if (ctl.ci[CI_PHASE] != PH_DONE) atomicAdd(&a.counters[0], 1);
  if (ctl.ci[CI_WAIT]) atomicAdd(&a.counters[1], 1);
}

// Factor + solve the lanes that requested a solve and have a structure.
// need == nullptr: every lane.
__global__ void __launch_bounds__(kBlock) k_lu(LuArgs a, const int32_t* need, const int32_t* wait, int32_t* status) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.B) return;
  if (need && (!need[c] || wait[c])) return;
  int st = lu_factor(a, c);
  if (st == LU_OK) st = lu_solve(a, c);
  status[c] = st;
}

__global__ void __launch_bounds__(kBlock) k_hash(const uint8_t* present, int32_t nslots, int64_t B,
                                                 const int32_t* wait, uint64_t* hash) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= B) return;
  if (!wait[c]) return;
  uint64_t h = 1469598103934665603ull;
  for (int s = 0; s < nslots; s++) {
    if (present[(int64_t)s * B + c]) {
      h ^= (uint64_t)(s + 1);
      h *= 1099511628211ull;
    }
  }
  hash[c] = h;
}

// ---- device buffers -----------------------------------------------------------------
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    release();
    if (b == 0) b = 8;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) { p = nullptr; return e; }
    bytes = b;
    return hipMemset(p, 0, b);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// LU workspace for B lanes of n unknowns with capL / capU entries.
struct LuWork {
  DBuf xw, w, udiag, lx, ux, li, ui, lp, up, pinv, prow, xi, pstack, mark, err;
  int64_t capL = 0, capU = 0, B = 0;
  int n = 0;
  hipError_t ensure(int n_, int64_t B_, int64_t cap) {
    if (n_ == n && B_ == B && cap <= capL && xw.p) return hipSuccess;
    n = n_;
    B = B_;
    capL = capU = cap;
    size_t nd = (size_t)n * B * 8, ni = (size_t)n * B * 4;
    hipError_t e;
#define KH_A(buf, bytes_) if ((e = buf.alloc(bytes_)) != hipSuccess) return e
    KH_A(xw, nd); KH_A(w, nd); KH_A(udiag, nd);
    KH_A(lx, (size_t)capL * B * 8); KH_A(ux, (size_t)capU * B * 8);
    KH_A(li, (size_t)capL * B * 4); KH_A(ui, (size_t)capU * B * 4);
    KH_A(lp, (size_t)(n + 1) * B * 4); KH_A(up, (size_t)(n + 1) * B * 4);
    KH_A(pinv, ni); KH_A(prow, ni); KH_A(xi, ni); KH_A(pstack, ni); KH_A(mark, ni);
    KH_A(err, (size_t)B * 4);
#undef KH_A
    return hipSuccess;
  }
  void fill(LuArgs& a) const {
    a.capL = capL; a.capU = capU;
    a.xw = xw.as<double>(); a.w = w.as<double>(); a.udiag = udiag.as<double>();
    a.lx = lx.as<double>(); a.ux = ux.as<double>(); a.li = li.as<int32_t>(); a.ui = ui.as<int32_t>();
    a.lp = lp.as<int32_t>(); a.up = up.as<int32_t>(); a.pinv = pinv.as<int32_t>(); a.prow = prow.as<int32_t>();
    a.xi = xi.as<int32_t>(); a.pstack = pstack.as<int32_t>(); a.mark = mark.as<int32_t>();
    a.err_step = err.as<int32_t>();
  }
};

int64_t lu_capacity(int64_t max_fill, int n) { return 2 * max_fill + 4 * (int64_t)n + 64; }

// Pattern pool mirrored on the device.
struct DevPatterns {
  PatternPool host;
  DBuf desc, pool;
  size_t up_desc = 0, up_pool = 0;
  hipError_t sync() {
    if (host.desc.size() != up_desc || host.pool.size() != up_pool) {
      hipError_t e;
      if ((e = desc.alloc(host.desc.size() * sizeof(PatDesc))) != hipSuccess) return e;
      if ((e = pool.alloc(host.pool.size() * sizeof(int32_t))) != hipSuccess) return e;
      if ((e = hipMemcpy(desc.p, host.desc.data(), host.desc.size() * sizeof(PatDesc), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
      if ((e = hipMemcpy(pool.p, host.pool.data(), host.pool.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
      up_desc = host.desc.size();
      up_pool = host.pool.size();
    }
    return hipSuccess;
  }
};

std::string slots_key(const std::vector<int32_t>& slots) {
  return std::string(reinterpret_cast<const char*>(slots.data()), slots.size() * sizeof(int32_t));
}

inline unsigned grid_for(int64_t B) { return (unsigned)((B + kBlock - 1) / kBlock); }

}  // namespace

// ---- handles -------------------------------------------------------------------------
struct kh_topology_s {
  Topology t;
  DBuf et, isV, slot_i, slot_j, diag;
  std::mutex mu;  // batch entry points on one topology run one at a time
};

struct kh_lu_batch_s {
  int n = 0;
  int64_t nnz = 0, B = 0;
  DevPatterns pats;
  Pattern pat;
  LuWork work;
  DBuf pid, status;
};

struct kh_sys_s {
  int n = 0;
  std::unordered_map<int64_t, int32_t> slot;
  std::vector<int32_t> ri, ci;
  std::vector<double> val;
  bool dirty = true;
  bool valid = false;
  Pattern pat;
  DevPatterns pats;
  LuWork work;
  DBuf dval, drhs, dx, pid, status;
  size_t dval_slots = 0;
};

extern "C" {

const char* kh_status_string(kh_status s) {
  switch (s) {
    case KH_OK: return "ok";
    case KH_SINGULAR: return "solver: singular matrix";
    case KH_NONFINITE: return "solver: non-finite value in system";
    case KH_BADSIZE: return "solver: invalid matrix size";
    case KH_NOCONV: return "analysis: no convergence";
    case KH_TSTEP_TOO_SMALL: return "analysis: timestep too small";
    case KH_TOO_MANY_STEPS: return "analysis: too many time steps";
    case KH_INTERRUPTED: return "analysis interrupted";
    case KH_LEDGER: return "em: energy ledger does not balance";
    case KH_BAD_TRAN_SPEC: return "analysis: invalid transient specification";
    case KH_UNSUPPORTED: return "device: unsupported element kind";
    case KH_BAD_ARGUMENT: return "invalid argument";
    case KH_HIP_ERROR: return "HIP runtime error";
    case KH_NO_DEVICE: return "no gfx950 device";
    case KH_CAPACITY: return "L/U capacity exceeded";
    case KH_RUNNING: return "running";
  }
  return "unknown status";
}

const char* kh_last_error(void) { return g_last_error.c_str(); }

kh_status kh_init(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(KH_NO_DEVICE, "no HIP device");
  int dev = 0;
  KH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  KH_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(KH_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", the library is built for gfx950");
  return KH_OK;
}

// ---- per-system seam --------------------------------------------------------------------
kh_status kh_sys_create(int32_t n, kh_sys* out) {
  if (!out) return fail(KH_BAD_ARGUMENT, "null handle");
  if (n <= 0 || n > (1 << 22)) return fail(KH_BADSIZE, "size " + std::to_string(n));
  auto* s = new kh_sys_s();
  s->n = n;
  *out = s;
  return KH_OK;
}

kh_status kh_sys_clear(kh_sys s) {
  if (!s) return fail(KH_BAD_ARGUMENT, "null handle");
  std::fill(s->val.begin(), s->val.end(), 0.0);
  s->valid = false;
  return KH_OK;
}

kh_status kh_sys_add_batch(kh_sys s, const int32_t* i, const int32_t* j, const double* v, int64_t count) {
  if (!s || (count > 0 && (!i || !j || !v))) return fail(KH_BAD_ARGUMENT, "null argument");
  for (int64_t k = 0; k < count; k++) {  // sparseLU.add, solver/sparse.go:114-143
    int ii = i[k], jj = j[k];
    if (ii < 0 || jj < 0 || ii >= s->n || jj >= s->n) continue;
    int64_t key = (int64_t)ii * s->n + jj;
    auto it = s->slot.find(key);
    int32_t sl;
    if (it == s->slot.end()) {
      sl = (int32_t)s->val.size();
      s->slot[key] = sl;
      s->ri.push_back(ii);
      s->ci.push_back(jj);
      s->val.push_back(0.0);
      s->dirty = true;
    } else {
      sl = it->second;
    }
    s->val[sl] += v[k];
    s->valid = false;
  }
  return KH_OK;
}

static kh_status sys_prepare(kh_sys s) {
  if (s->dirty) {
    std::vector<int32_t> slots(s->val.size());
    for (size_t k = 0; k < slots.size(); k++) slots[k] = (int32_t)k;
    s->pat = analyze_pattern(s->n, s->ri, s->ci, slots);
    s->pats.host = PatternPool();
    s->pats.up_desc = s->pats.up_pool = 0;
    s->pats.host.add(s->pat, slots_key(slots));
    KH_HIP(s->pats.sync());
    KH_HIP(s->work.ensure(s->n, 1, lu_capacity(s->pat.fill_estimate, s->n)));
    KH_HIP(s->dval.alloc(std::max<size_t>(1, s->val.size()) * 8));
    s->dval_slots = s->val.size();
    KH_HIP(s->drhs.alloc((size_t)s->n * 8));
    KH_HIP(s->dx.alloc((size_t)s->n * 8));
    KH_HIP(s->pid.alloc(4));
    KH_HIP(s->status.alloc(4));
    s->dirty = false;
  }
  return KH_OK;
}

kh_status kh_sys_solve(kh_sys s, const double* b, double* x) {
  if (!s || !b || !x) return fail(KH_BAD_ARGUMENT, "null argument");
  kh_status st = sys_prepare(s);
  if (st != KH_OK) return st;
  if (s->val.empty()) return fail(KH_SINGULAR, "empty matrix");
  KH_HIP(hipMemcpy(s->dval.p, s->val.data(), s->val.size() * 8, hipMemcpyHostToDevice));
  KH_HIP(hipMemcpy(s->drhs.p, b, (size_t)s->n * 8, hipMemcpyHostToDevice));
  LuArgs a = {};
  a.n = s->n;
  a.nslots = (int32_t)s->val.size();
  a.B = 1;
  a.pats = s->pats.desc.as<PatDesc>();
  a.pool = s->pats.pool.as<int32_t>();
  a.pid = s->pid.as<int32_t>();
  a.val = s->dval.as<double>();
  a.rhs = s->drhs.as<double>();
  a.x = s->dx.as<double>();
  s->work.fill(a);
  hipLaunchKernelGGL(k_lu, dim3(1), dim3(kBlock), 0, 0, a, nullptr, nullptr, s->status.as<int32_t>());
  KH_HIP(hipGetLastError());
  int32_t hs = 0;
  KH_HIP(hipMemcpy(&hs, s->status.p, 4, hipMemcpyDeviceToHost));
  if (hs == LU_OK || hs == LU_NONFINITE) KH_HIP(hipMemcpy(x, s->dx.p, (size_t)s->n * 8, hipMemcpyDeviceToHost));
  s->valid = hs == LU_OK;
  if (hs == LU_SINGULAR) return fail(KH_SINGULAR, "solver: singular matrix");
  if (hs == LU_NONFINITE) return fail(KH_NONFINITE, "solver: non-finite value in system");
  if (hs == LU_CAPACITY) return fail(KH_CAPACITY, "L/U capacity exceeded");
  return KH_OK;
}

kh_status kh_sys_absproduct(kh_sys s, const double* x, double* y) {
  if (!s || !x || !y) return fail(KH_BAD_ARGUMENT, "null argument");
  for (int i = 0; i < s->n; i++) y[i] = 0;
  for (size_t k = 0; k < s->val.size(); k++) y[s->ri[k]] += std::fabs(s->val[k]) * std::fabs(x[s->ci[k]]);
  return KH_OK;
}

kh_status kh_sys_nnz(kh_sys s, int64_t* a, int64_t* l, int64_t* u) {
  if (!s || !s->valid) return fail(KH_BAD_ARGUMENT, "no factorization held");
  int32_t lpn = 0, upn = 0;
  KH_HIP(hipMemcpy(&lpn, s->work.lp.as<int32_t>() + s->n, 4, hipMemcpyDeviceToHost));
  KH_HIP(hipMemcpy(&upn, s->work.up.as<int32_t>() + s->n, 4, hipMemcpyDeviceToHost));
  if (a) *a = (int64_t)s->val.size();
  if (l) *l = lpn;
  if (u) *u = (int64_t)upn + s->n;
  return KH_OK;
}

kh_status kh_sys_pivots(kh_sys s, int32_t* q, int32_t* prow) {
  if (!s || !s->valid) return fail(KH_BAD_ARGUMENT, "no factorization held");
  if (q) std::memcpy(q, s->pat.q.data(), (size_t)s->n * 4);
  if (prow) KH_HIP(hipMemcpy(prow, s->work.prow.p, (size_t)s->n * 4, hipMemcpyDeviceToHost));
  return KH_OK;
}

void kh_sys_destroy(kh_sys s) { delete s; }

// ---- batched numeric LU ------------------------------------------------------------------
kh_status kh_lu_batch_create(int32_t n, int64_t nnz, const int32_t* ri, const int32_t* cj, int32_t B,
                             kh_lu_batch* out) {
  if (!out || !ri || !cj || B <= 0 || nnz <= 0) return fail(KH_BAD_ARGUMENT, "bad argument");
  if (n <= 0 || n > (1 << 22)) return fail(KH_BADSIZE, "size " + std::to_string(n));
  std::vector<int32_t> r(ri, ri + nnz), c(cj, cj + nnz), slots(nnz);
  std::unordered_map<int64_t, int> seen;
  for (int64_t k = 0; k < nnz; k++) {
    if (r[k] < 0 || c[k] < 0 || r[k] >= n || c[k] >= n) return fail(KH_BAD_ARGUMENT, "entry out of range");
    if (!seen.emplace((int64_t)r[k] * n + c[k], 1).second) return fail(KH_BAD_ARGUMENT, "duplicate entry");
    slots[k] = (int32_t)k;
  }
  auto h = std::make_unique<kh_lu_batch_s>();
  h->n = n;
  h->nnz = nnz;
  h->B = B;
  h->pat = analyze_pattern(n, r, c, slots);
  h->pats.host.add(h->pat, slots_key(slots));
  KH_HIP(h->pats.sync());
  KH_HIP(h->work.ensure(n, B, lu_capacity(h->pat.fill_estimate, n)));
  KH_HIP(h->pid.alloc((size_t)B * 4));
  KH_HIP(h->status.alloc((size_t)B * 4));
  *out = h.release();
  return KH_OK;
}

kh_status kh_lu_batch_solve(kh_lu_batch h, const double* d_vals, const double* d_b, double* d_x, int32_t* d_status,
                            int32_t* d_prow, void* stream) {
  if (!h || !d_vals || !d_b || !d_x) return fail(KH_BAD_ARGUMENT, "null argument");
  LuArgs a = {};
  a.n = h->n;
  a.nslots = (int32_t)h->nnz;
  a.B = h->B;
  a.pats = h->pats.desc.as<PatDesc>();
  a.pool = h->pats.pool.as<int32_t>();
  a.pid = h->pid.as<int32_t>();
  a.val = d_vals;
  a.rhs = d_b;
  a.x = d_x;
  h->work.fill(a);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t* status = d_status ? d_status : h->status.as<int32_t>();
  hipLaunchKernelGGL(k_lu, dim3(grid_for(h->B)), dim3(kBlock), 0, st, a, nullptr, nullptr, status);
  KH_HIP(hipGetLastError());
  if (d_prow)
    KH_HIP(hipMemcpyAsync(d_prow, h->work.prow.p, (size_t)h->n * h->B * 4, hipMemcpyDeviceToDevice, st));
  return KH_OK;
}

kh_status kh_lu_batch_info(kh_lu_batch h, int32_t* q, int64_t* nnzL, int64_t* nnzU, int64_t* capL, int64_t* capU) {
  if (!h) return fail(KH_BAD_ARGUMENT, "null handle");
  if (q) std::memcpy(q, h->pat.q.data(), (size_t)h->n * 4);
  if (nnzL || nnzU) {
    int32_t lpn = 0, upn = 0;
    KH_HIP(hipMemcpy(&lpn, h->work.lp.as<int32_t>() + (size_t)h->n * h->B, 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(&upn, h->work.up.as<int32_t>() + (size_t)h->n * h->B, 4, hipMemcpyDeviceToHost));
    if (nnzL) *nnzL = lpn;
    if (nnzU) *nnzU = (int64_t)upn + h->n;
  }
  if (capL) *capL = h->work.capL;
  if (capU) *capU = h->work.capU;
  return KH_OK;
}

void kh_lu_batch_destroy(kh_lu_batch h) { delete h; }

// ---- topologies and batched analyses -------------------------------------------------------
kh_status kh_topology_create(int32_t num_nodes, const kh_element* elems, int32_t n_elems, int32_t nparams,
                             const double* params_c0, kh_topology* out) {
  if (!out || !elems || n_elems <= 0 || !params_c0 || num_nodes < 1) return fail(KH_BAD_ARGUMENT, "bad argument");
  auto t = std::make_unique<kh_topology_s>();
  if (!t->t.build(num_nodes, elems, n_elems, nparams, params_c0)) return fail(KH_UNSUPPORTED, t->t.error);
  const Topology& T = t->t;
  KH_HIP(t->et.alloc(T.et.size() * 4));
  KH_HIP(hipMemcpy(t->et.p, T.et.data(), T.et.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->isV.alloc(T.isV.size() * 4));
  KH_HIP(hipMemcpy(t->isV.p, T.isV.data(), T.isV.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->slot_i.alloc(T.slot_i.size() * 4));
  KH_HIP(hipMemcpy(t->slot_i.p, T.slot_i.data(), T.slot_i.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->slot_j.alloc(T.slot_j.size() * 4));
  KH_HIP(hipMemcpy(t->slot_j.p, T.slot_j.data(), T.slot_j.size() * 4, hipMemcpyHostToDevice));
  KH_HIP(t->diag.alloc(T.diag_slot.size() * 4));
  KH_HIP(hipMemcpy(t->diag.p, T.diag_slot.data(), T.diag_slot.size() * 4, hipMemcpyHostToDevice));
  *out = t.release();
  return KH_OK;
}

int32_t kh_topology_size(kh_topology t) { return t ? t->t.n : -1; }

void kh_topology_destroy(kh_topology t) { delete t; }

}  // extern "C"

// ---- the batch driver -----------------------------------------------------------------------
namespace {

struct BatchRun {
  kh_topology_s* topo;
  int64_t B;
  int n;
  DBuf st, val, present, rhs, x, xold, xsave, ptprev, xacc, absbuf, ci, cd, lustatus, counters, pid, hash;
  DevPatterns pats;
  LuWork work;
  int32_t* hcount = nullptr;
  const double* lane_spec = nullptr;
  kh_stats stats = {};

  ~BatchRun() {
    if (hcount) (void)hipHostFree(hcount);
  }

  kh_status setup() {
    const Topology& T = topo->t;
    n = T.n;
    size_t nd = (size_t)n * B * 8;
    KH_HIP(st.alloc((size_t)std::max(T.nstate, 1) * B * 8));
    KH_HIP(val.alloc((size_t)T.nslots * B * 8));
    KH_HIP(present.alloc((size_t)T.nslots * B));
    KH_HIP(rhs.alloc(nd)); KH_HIP(x.alloc(nd)); KH_HIP(xold.alloc(nd)); KH_HIP(xsave.alloc(nd));
    KH_HIP(ptprev.alloc(nd)); KH_HIP(xacc.alloc(nd)); KH_HIP(absbuf.alloc(nd));
    KH_HIP(ci.alloc((size_t)NCI * B * 4));
    KH_HIP(cd.alloc((size_t)NCD * B * 8));
    KH_HIP(lustatus.alloc((size_t)B * 4));
    KH_HIP(counters.alloc(16));
    KH_HIP(pid.alloc((size_t)B * 4));
    KH_HIP(hash.alloc((size_t)B * 8));
    KH_HIP(hipHostMalloc((void**)&hcount, 16, hipHostMallocDefault));
    return KH_OK;
  }

  CtlArgs ctl_args(const double* d_params, const CtlParams& p) const {
    const Topology& T = topo->t;
    CtlArgs a = {};
    a.ckt.n = T.n; a.ckt.nel = T.nel; a.ckt.nslots = T.nslots; a.ckt.numNodes = T.num_nodes;
    a.ckt.et = topo->et.as<int32_t>(); a.ckt.isV = topo->isV.as<int32_t>();
    a.ckt.slot_i = topo->slot_i.as<int32_t>(); a.ckt.slot_j = topo->slot_j.as<int32_t>();
    a.ckt.diag_slot = topo->diag.as<int32_t>();
    a.ckt.nonlin = T.nonlin ? 1 : 0;
    a.p = p;
    a.B = B;
    a.par = d_params;
    a.st = st.as<double>(); a.val = val.as<double>(); a.present = present.as<uint8_t>();
    a.rhs = rhs.as<double>(); a.x = x.as<double>(); a.xold = xold.as<double>();
    a.xsave = xsave.as<double>(); a.ptprev = ptprev.as<double>(); a.xacc = xacc.as<double>();
    a.absbuf = absbuf.as<double>();
    a.ci = ci.as<int32_t>(); a.cd = cd.as<double>();
    a.lu_status = lustatus.as<int32_t>();
    a.counters = counters.as<int32_t>();
    a.lane_spec = lane_spec;
    return a;
  }

  LuArgs lu_args() const {
    const Topology& T = topo->t;
    LuArgs a = {};
    a.n = T.n;
    a.nslots = T.nslots;
    a.B = B;
    a.pats = pats.desc.as<PatDesc>();
    a.pool = pats.pool.as<int32_t>();
    a.pid = pid.as<int32_t>();
    a.val = val.as<double>();
    a.rhs = rhs.as<double>();
    a.x = x.as<double>();
    work.fill(a);
    return a;
  }

  // Symbolic phase for lanes whose stamps grew their pattern.
  kh_status new_patterns() {
    const Topology& T = topo->t;
    int32_t* wait_row = ci.as<int32_t>() + (size_t)CI_WAIT * B;
    hipLaunchKernelGGL(k_hash, dim3(grid_for(B)), dim3(kBlock), 0, 0, present.as<uint8_t>(), T.nslots, B, wait_row,
                       hash.as<uint64_t>());
    KH_HIP(hipGetLastError());
    std::vector<int32_t> wait(B), hpid(B);
    std::vector<uint64_t> hh(B);
    KH_HIP(hipMemcpy(wait.data(), wait_row, (size_t)B * 4, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hh.data(), hash.p, (size_t)B * 8, hipMemcpyDeviceToHost));
    KH_HIP(hipMemcpy(hpid.data(), pid.p, (size_t)B * 4, hipMemcpyDeviceToHost));
    std::unordered_map<uint64_t, int32_t> by_hash;
    std::vector<uint8_t> col(T.nslots);
    for (int64_t c = 0; c < B; c++) {
      if (!wait[c]) continue;
      auto it = by_hash.find(hh[c]);
      if (it != by_hash.end()) {
        hpid[c] = it->second;
        continue;
      }
      KH_HIP(hipMemcpy2D(col.data(), 1, present.as<uint8_t>() + c, (size_t)B, 1, T.nslots, hipMemcpyDeviceToHost));
      std::vector<int32_t> slots;
      for (int s = 0; s < T.nslots; s++)
        if (col[s]) slots.push_back(s);
      std::string key = slots_key(slots);
      int32_t id;
      auto f = pats.host.by_key.find(key);
      if (f != pats.host.by_key.end()) {
        id = f->second;
      } else {
        Pattern P = analyze_pattern(T.n, T.slot_i, T.slot_j, slots);
        id = pats.host.add(P, key);
        stats.analyses++;
      }
      by_hash[hh[c]] = id;
      hpid[c] = id;
      wait[c] = 0;
    }
    for (int64_t c = 0; c < B; c++) wait[c] = 0;
    stats.patterns = (int64_t)pats.host.desc.size();
    KH_HIP(pats.sync());
    KH_HIP(work.ensure(T.n, B, std::max(work.capL, lu_capacity(pats.host.max_fill, T.n))));
    KH_HIP(hipMemcpy(pid.p, hpid.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    KH_HIP(hipMemcpy(wait_row, wait.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    return KH_OK;
  }

  kh_status run(const double* d_params, const CtlParams& p) {
    CtlArgs ca = ctl_args(d_params, p);
    // Kernel durations from HIP events on the launch stream (stream 0).
    hipEvent_t ev[4];
    for (auto& e : ev) KH_HIP(hipEventCreate(&e));
    struct EvGuard {
      hipEvent_t* e;
      ~EvGuard() { for (int k = 0; k < 4; k++) (void)hipEventDestroy(e[k]); }
    } guard{ev};
    bool lu_pending = false;
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      KH_HIP(hipMemsetAsync(counters.p, 0, 16, 0));
      KH_HIP(hipEventRecord(ev[0], 0));
      hipLaunchKernelGGL(k_control, dim3(grid_for(B)), dim3(kBlock), 0, 0, ca);
      KH_HIP(hipGetLastError());
      KH_HIP(hipEventRecord(ev[1], 0));
      KH_HIP(hipMemcpyAsync(hcount, counters.p, 8, hipMemcpyDeviceToHost, 0));
      KH_HIP(hipStreamSynchronize(0));
      float ms = 0;
      KH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      stats.control_seconds += ms * 1e-3;
      stats.control_launches++;
      if (lu_pending) {
        KH_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        stats.lu_seconds += ms * 1e-3;
        stats.lu_launches++;
        lu_pending = false;
      }
      int32_t running = hcount[0], waiting = hcount[1];
      if (running == 0) break;
      if (waiting > 0) {
        kh_status s = new_patterns();
        if (s != KH_OK) return s;
      }
      LuArgs la = lu_args();
      KH_HIP(hipEventRecord(ev[2], 0));
      hipLaunchKernelGGL(k_lu, dim3(grid_for(B)), dim3(kBlock), 0, 0, la,
                         ci.as<int32_t>() + (size_t)CI_NEED_SOLVE * B, ci.as<int32_t>() + (size_t)CI_WAIT * B,
                         lustatus.as<int32_t>());
      KH_HIP(hipGetLastError());
      KH_HIP(hipEventRecord(ev[3], 0));
      lu_pending = true;
      stats.ticks++;
    }
    KH_HIP(hipDeviceSynchronize());
    stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return KH_OK;
  }

  kh_status outputs(double* d_x, int32_t* d_status, int32_t* d_iters, int32_t* d_acc, int32_t* d_rej,
                    double* d_ledger, kh_stats* out) {
    if (d_x) KH_HIP(hipMemcpy(d_x, x.p, (size_t)n * B * 8, hipMemcpyDeviceToDevice));
    const int32_t* cip = ci.as<int32_t>();
    if (d_status) KH_HIP(hipMemcpy(d_status, cip + (size_t)CI_STATUS * B, (size_t)B * 4, hipMemcpyDeviceToDevice));
    if (d_iters) KH_HIP(hipMemcpy(d_iters, cip + (size_t)CI_ITERS * B, (size_t)B * 4, hipMemcpyDeviceToDevice));
    if (d_acc) KH_HIP(hipMemcpy(d_acc, cip + (size_t)CI_ACCEPTED * B, (size_t)B * 4, hipMemcpyDeviceToDevice));
    if (d_rej) KH_HIP(hipMemcpy(d_rej, cip + (size_t)CI_REJECTED * B, (size_t)B * 4, hipMemcpyDeviceToDevice));
    if (d_ledger) {
      const double* cdp = cd.as<double>();
      KH_HIP(hipMemcpy(d_ledger, cdp + (size_t)CD_MAXREL * B, (size_t)B * 8, hipMemcpyDeviceToDevice));
    }
    if (out) {
      std::vector<int32_t> it(B), acc(B), rej(B), sol(B);
      KH_HIP(hipMemcpy(it.data(), cip + (size_t)CI_ITERS * B, (size_t)B * 4, hipMemcpyDeviceToHost));
      KH_HIP(hipMemcpy(acc.data(), cip + (size_t)CI_ACCEPTED * B, (size_t)B * 4, hipMemcpyDeviceToHost));
      KH_HIP(hipMemcpy(rej.data(), cip + (size_t)CI_REJECTED * B, (size_t)B * 4, hipMemcpyDeviceToHost));
      KH_HIP(hipMemcpy(sol.data(), cip + (size_t)CI_SOLVES * B, (size_t)B * 4, hipMemcpyDeviceToHost));
      for (int64_t c = 0; c < B; c++) {
        stats.newton_iterations += it[c];
        stats.accepted_steps += acc[c];
        stats.rejected_steps += rej[c];
        stats.solves += sol[c];
      }
      // Algorithmic bytes per executed solve: 8*(nnzA + 2n), with nnzA the
      // largest pattern's entry count (patterns only grow).
      int64_t nnz = 0;
      for (const auto& d : pats.host.desc) nnz = std::max<int64_t>(nnz, d.nnz);
      stats.bytes_algorithmic = stats.solves * 8 * (nnz + 2 * (int64_t)n);
      *out = stats;
    }
    return KH_OK;
  }
};

CtlParams resolve(const kh_options* o, int analysis) {
  CtlParams p = {};
  p.relTol = (o && o->reltol > 0) ? o->reltol : 1e-3;
  p.absTol = (o && o->abstol > 0) ? o->abstol : 1e-12;
  p.vnTol = (o && o->vntol > 0) ? o->vntol : 1e-6;
  p.gmin = (o && o->gmin > 0) ? o->gmin : 1e-12;
  p.ledgerTol = (o && o->ledger_tol > 0) ? o->ledger_tol : 1e-9;
  p.itl1 = (o && o->itl1 > 0) ? o->itl1 : 100;
  p.itl4 = (o && o->itl4 > 0) ? o->itl4 : 10;
  p.noGmin = o ? o->no_gmin_stepping : 0;
  p.noSource = o ? o->no_source_stepping : 0;
  p.noPT = o ? o->no_pseudo_transient : 0;
  p.skipEnergy = o ? o->skip_energy_check : 0;
  p.analysis = analysis;
  return p;
}

}  // namespace

extern "C" {

kh_status kh_batch_op(kh_topology t, int32_t B, const double* d_params, const kh_options* opt, double* d_x,
                      int32_t* d_status, int32_t* d_iters, kh_stats* stats) {
  if (!t || B <= 0 || !d_params) return fail(KH_BAD_ARGUMENT, "bad argument");
  std::lock_guard<std::mutex> lk(t->mu);
  BatchRun r;
  r.topo = t;
  r.B = B;
  kh_status s = r.setup();
  if (s != KH_OK) return s;
  // polish (analysis/op.go:251-271) when the caller asks for round-off exact points
  CtlParams p = resolve(opt, (opt && opt->polish) ? AN_OP_POLISH : AN_OP);
  s = r.run(d_params, p);
  if (s != KH_OK) return s;
  return r.outputs(d_x, d_status, d_iters, nullptr, nullptr, nullptr, stats);
}

kh_status kh_batch_tran(kh_topology t, int32_t B, const double* d_params, const kh_options* opt,
                        const kh_tran_spec* spec, const double* d_lane_spec, double* d_x, int32_t* d_status,
                        int32_t* d_accepted, int32_t* d_rejected, double* d_ledger, kh_stats* stats) {
  if (!t || B <= 0 || !d_params || !spec) return fail(KH_BAD_ARGUMENT, "bad argument");
  // Tran's spec validation, analysis/tran.go:71-74
  if (!(spec->step > 0) || !(spec->stop > 0) || spec->start < 0 || spec->start >= spec->stop || spec->max_step < 0 ||
      std::isnan(spec->step + spec->stop + spec->start + spec->max_step) || std::isinf(spec->stop))
    return fail(KH_BAD_TRAN_SPEC, "invalid transient specification");
  std::lock_guard<std::mutex> lk(t->mu);
  BatchRun r;
  r.topo = t;
  r.B = B;
  kh_status s = r.setup();
  if (s != KH_OK) return s;
  CtlParams p = resolve(opt, AN_TRAN);
  p.step = spec->step;
  p.stop = spec->stop;
  p.start = spec->start;
  p.maxStep = spec->max_step > 0 ? spec->max_step : std::min(spec->step, (spec->stop - spec->start) / 50);
  p.chgTol = spec->chgtol > 0 ? spec->chgtol : 1e-14;
  p.trTol = spec->trtol > 0 ? spec->trtol : 7;
  p.minStep = spec->min_step > 0 ? spec->min_step : 1e-12 * spec->stop;
  p.maxSteps = spec->max_steps > 0 ? spec->max_steps : 1000000;
  p.uic = spec->uic;
  p.method = spec->method == 1 ? BE : TRAP;
  r.lane_spec = d_lane_spec;
  s = r.run(d_params, p);
  if (s != KH_OK) return s;
  return r.outputs(d_x, d_status, nullptr, d_accepted, d_rejected, d_ledger, stats);
}

}  // extern "C"
