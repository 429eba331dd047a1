// kh_mlp.hip — the learned Newton warm start: a small per-topology MLP that
// maps a circuit's value genes to its DC node voltages, whose predictions seed
// Newton through Options.NodeSet (kh_group.d_nodeset).
//
// Network: in -> H -> H -> out, ReLU after the two hidden layers. Weights and
// activations enter the matrix products as bf16, products accumulate in FP32 on
// the matrix cores (v_mfma_f32_16x16x32_bf16); master weights, biases, losses
// and the Adam moments are FP32. Every dimension is padded (inputs and outputs
// to multiples of 64, the batch to a multiple of 64) with zero rows/columns that
// never reach a loss.
//
// Products: one kernel, k_gemm_bf16, C[M][N] = sum_k A(m,k) B(k,n) with any
// strides for A and B (so A^T B and A B^T are the same kernel): a 256-thread
// workgroup stages a 64x32 tile of A and a 32x64 tile of B in LDS with the
// strides applied, then each of its four wavefronts runs 2x2 MFMA 16x16x32
// blocks of its 32x32 quarter of the 64x64 output tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kronhip.h"

typedef __bf16 kh_bf16x8 __attribute__((ext_vector_type(8)));
typedef float kh_f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kTile = 64, kTileK = 32;

// FP32 -> bf16 bits, round to nearest even (NaN stays NaN).
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct GemmArgs {
  const uint16_t* A;  // bf16 bits, element (m, k) at A[m * sam + k * sak]
  const uint16_t* B;  // bf16 bits, element (k, n) at B[k * sbk + n * sbn]
  float* C;           // C[m * ldc + n]
  int64_t sam, sak, sbk, sbn, ldc;
  int32_t M, N, K;    // multiples of 64, 64, 32
  float beta;         // C = A B + beta C
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ uint16_t As[kTile][kTileK + 8];  // As[m][k]
  __shared__ uint16_t Bt[kTile][kTileK + 8];  // Bt[n][k]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.y * kTile, n0 = (int64_t)blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  kh_f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = kh_f32x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < g.K; k0 += kTileK) {
    for (int e = t; e < kTile * kTileK; e += 256) {  // 64 x 32 of each operand
      const int r = e / kTileK, k = e % kTileK;
      As[r][k] = g.A[(m0 + r) * g.sam + (int64_t)(k0 + k) * g.sak];
      Bt[r][k] = g.B[(int64_t)(k0 + k) * g.sbk + (n0 + r) * g.sbn];
    }
    __syncthreads();
    // Operand layout of v_mfma_f32_16x16x32_bf16: lane l holds row (l % 16) of
    // A and column (l % 16) of B, reduction indices 8 * (l / 16) .. + 7.
    const int rr = lane & 15, kk = (lane >> 4) * 8;
    kh_bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) {
      uint16_t tmp[8];
      for (int q = 0; q < 8; q++) tmp[q] = As[wm + 16 * i + rr][kk + q];
      memcpy(&a[i], tmp, 16);
      for (int q = 0; q < 8; q++) tmp[q] = Bt[wn + 16 * i + rr][kk + q];
      memcpy(&b[i], tmp, 16);
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // Result layout: lane l holds column (l % 16), rows 4 * (l / 16) .. + 3.
  const int col = lane & 15, row4 = (lane >> 4) * 4;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        const int64_t m = m0 + wm + 16 * i + row4 + r, n = n0 + wn + 16 * j + col;
        float* c = g.C + m * g.ldc + n;
        *c = g.beta == 0.0f ? acc[i][j][r] : acc[i][j][r] + g.beta * *c;
      }
}

// Z += bias (per column); A = relu(Z) or Z; bf16 copies of A in [M][N] and
// transposed [N][M] (the next product's operand and the weight gradient's).
__global__ void k_bias_act(float* Z, const float* bias, int32_t M, int32_t N, int32_t relu, uint16_t* Ab,
                           uint16_t* AbT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int64_t m = i / N, n = i % N;
  float z = Z[i] + bias[n];
  Z[i] = z;
  const float a = (relu && z < 0.0f) ? 0.0f : z;
  if (Ab) Ab[i] = to_bf16(a);
  if (AbT) AbT[n * M + m] = to_bf16(a);
}

// Mean squared error over the first B rows and n_out columns; dY = 2 (Y - T) / (B n_out),
// as FP32 and as bf16 in [M][N] and [N][M]. loss[0] accumulates the sum of squares.
__global__ void __launch_bounds__(256) k_mse(const float* Y, const float* T, int32_t B, int32_t n_out, int32_t M,
                                             int32_t N, float* dY, uint16_t* dYb, uint16_t* dYbT, float* loss) {
  __shared__ float part[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (i < (int64_t)M * N) {
    const int64_t m = i / N, n = i % N;
    float d = 0.0f;
    if (m < B && n < n_out) {
      const float e = Y[i] - T[m * n_out + n];
      d = 2.0f * e / ((float)B * (float)n_out);
      sq = e * e;
    }
    dY[i] = d;
    dYb[i] = to_bf16(d);
    dYbT[n * M + m] = to_bf16(d);
  }
  part[threadIdx.x] = sq;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// dZ = dA * (Z > 0), as bf16 in [M][N] and [N][M].
__global__ void k_relu_back(const float* dA, const float* Z, int32_t M, int32_t N, uint16_t* dZb, uint16_t* dZbT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int64_t m = i / N, n = i % N;
  const float d = Z[i] > 0.0f ? dA[i] : 0.0f;
  dZb[i] = to_bf16(d);
  dZbT[n * M + m] = to_bf16(d);
}

// db[n] = sum over the M rows of dZ (rows beyond the batch are zero).
__global__ void k_colsum(const float* dZ, const float* Z, int32_t relu, int32_t M, int32_t N, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int64_t m = 0; m < M; m++) {
    const float d = dZ[m * N + n];
    if (!relu || Z[m * N + n] > 0.0f) s += d;
  }
  db[n] = s;
}

// Adam (Kingma & Ba 2015) on FP32 masters; refreshes the bf16 copies W [K][N]
// and W^T [N][K].
__global__ void k_adam(float* w, const float* g, float* m1, float* m2, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2, int32_t K, int32_t N, uint16_t* wb, uint16_t* wbT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = wi;
  if (wb) {
    const int64_t k = i / N, n = i % N;
    wb[i] = to_bf16(wi);
    wbT[n * K + k] = to_bf16(wi);
  }
}

__global__ void k_to_bf16_pad(const float* x, int32_t B, int32_t n, int32_t M, int32_t K, uint16_t* xb,
                              uint16_t* xbT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * K) return;
  const int64_t m = i / K, k = i % K;
  const float v = (m < B && k < n) ? x[m * n + k] : 0.0f;
  xb[i] = to_bf16(v);
  if (xbT) xbT[k * M + m] = to_bf16(v);
}

__global__ void k_unpad(const float* Y, int32_t B, int32_t n_out, int32_t N, float* y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * n_out) return;
  y[i] = Y[(i / n_out) * N + i % n_out];
}

int64_t pad(int64_t v, int64_t to) { return (v + to - 1) / to * to; }
unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  ~Buf() { if (p) (void)hipFree(p); }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b ? b : 8);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

thread_local std::string g_mlp_err;

}  // namespace

struct kh_mlp_s {
  int32_t nin = 0, nh = 0, nout = 0;
  int32_t Kin = 0, H = 0, Nout = 0;  // padded
  int64_t step = 0;
  // layer l: W [K_l][N_l] FP32 master, bias [N_l], Adam moments, bf16 W and W^T
  Buf W[3], b[3], mW[3], vW[3], mb[3], vb[3], Wb[3], WbT[3], gW[3], gb[3];
  // per-batch activations
  int64_t Mcap = 0;
  Buf xb, xbT, Z[3], Ab[2], AbT[2], dZb[3], dZbT[3], dA[3], loss;
  int32_t K(int l) const { return l == 0 ? Kin : H; }
  int32_t N(int l) const { return l == 2 ? Nout : H; }
};

namespace {

kh_status mlp_fail(const std::string& s) {
  g_mlp_err = s;
  return KH_BAD_ARGUMENT;
}
#define KH_MLP_HIP(x)                                                   \
  do {                                                                  \
    hipError_t e_ = (x);                                                \
    if (e_ != hipSuccess) {                                             \
      g_mlp_err = std::string(#x) + ": " + hipGetErrorString(e_);       \
      return KH_HIP_ERROR;                                              \
    }                                                                   \
  } while (0)

hipError_t gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn, float* C,
                int64_t ldc, int32_t M, int32_t N, int32_t K, float beta) {
  GemmArgs g{A, B, C, sam, sak, sbk, sbn, ldc, M, N, K, beta};
  hipLaunchKernelGGL(k_gemm_bf16, dim3(N / kTile, M / kTile), dim3(256), 0, 0, g);
  return hipGetLastError();
}

kh_status ensure_batch(kh_mlp m, int64_t M) {
  if (M <= m->Mcap) return KH_OK;
  KH_MLP_HIP(m->xb.ensure((size_t)M * m->Kin * 2));
  KH_MLP_HIP(m->xbT.ensure((size_t)M * m->Kin * 2));
  for (int l = 0; l < 3; l++) {
    KH_MLP_HIP(m->Z[l].ensure((size_t)M * m->N(l) * 4));
    KH_MLP_HIP(m->dA[l].ensure((size_t)M * m->N(l) * 4));
    KH_MLP_HIP(m->dZb[l].ensure((size_t)M * m->N(l) * 2));
    KH_MLP_HIP(m->dZbT[l].ensure((size_t)M * m->N(l) * 2));
  }
  for (int l = 0; l < 2; l++) {
    KH_MLP_HIP(m->Ab[l].ensure((size_t)M * m->H * 2));
    KH_MLP_HIP(m->AbT[l].ensure((size_t)M * m->H * 2));
  }
  KH_MLP_HIP(m->loss.ensure(8));
  m->Mcap = M;
  return KH_OK;
}

// Forward pass over M padded rows: Z[l] = act(prev) W_l + b_l.
kh_status forward(kh_mlp m, const float* d_x, int32_t B, int64_t M) {
  hipLaunchKernelGGL(k_to_bf16_pad, dim3(blocks(M * m->Kin)), dim3(256), 0, 0, d_x, B, m->nin, (int32_t)M, m->Kin,
                     m->xb.as<uint16_t>(), m->xbT.as<uint16_t>());
  KH_MLP_HIP(hipGetLastError());
  const uint16_t* in = m->xb.as<uint16_t>();
  for (int l = 0; l < 3; l++) {
    const int32_t K = m->K(l), N = m->N(l);
    // (m, k) of the input at in[m * K + k]; (k, n) of W via W^T [N][K]: WbT[n * K + k]
    KH_MLP_HIP(gemm(in, K, 1, m->WbT[l].as<uint16_t>(), 1, K, m->Z[l].as<float>(), N, (int32_t)M, N, K, 0.0f));
    uint16_t* ab = l < 2 ? m->Ab[l].as<uint16_t>() : nullptr;
    uint16_t* abt = l < 2 ? m->AbT[l].as<uint16_t>() : nullptr;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(M * N)), dim3(256), 0, 0, m->Z[l].as<float>(), m->b[l].as<float>(),
                       (int32_t)M, N, l < 2 ? 1 : 0, ab, abt);
    KH_MLP_HIP(hipGetLastError());
    in = ab;
  }
  return KH_OK;
}

// Loss and gradients of the mean squared error against T [B][n_out].
kh_status backward(kh_mlp m, const float* d_t, int32_t B, int64_t M, float* loss_out) {
  KH_MLP_HIP(hipMemset(m->loss.p, 0, 4));
  const int32_t Mi = (int32_t)M;
  hipLaunchKernelGGL(k_mse, dim3(blocks(M * m->Nout)), dim3(256), 0, 0, m->Z[2].as<float>(), d_t, B, m->nout, Mi,
                     m->Nout, m->dA[2].as<float>(), m->dZb[2].as<uint16_t>(), m->dZbT[2].as<uint16_t>(),
                     m->loss.as<float>());
  KH_MLP_HIP(hipGetLastError());
  for (int l = 2; l >= 0; l--) {
    const int32_t K = m->K(l), N = m->N(l);
    const uint16_t* inT = l == 0 ? m->xbT.as<uint16_t>() : m->AbT[l - 1].as<uint16_t>();  // [K][M]
    // gW[l] [K][N] = in^T dZ: (k, m) of in^T at inT[k * M + m]; (m, n) of dZ at dZbT[n * M + m]
    KH_MLP_HIP(gemm(inT, M, 1, m->dZbT[l].as<uint16_t>(), 1, M, m->gW[l].as<float>(), N, K, N, Mi, 0.0f));
    // gb[l] = column sums of dZ (for the output layer dZ = dY, no ReLU)
    hipLaunchKernelGGL(k_colsum, dim3(blocks(N)), dim3(256), 0, 0, m->dA[l].as<float>(), m->Z[l].as<float>(),
                       l < 2 ? 1 : 0, Mi, N, m->gb[l].as<float>());
    KH_MLP_HIP(hipGetLastError());
    if (l == 0) break;
    // dA[l-1] [M][K] = dZ W^T: (m, n) of dZ at dZb[m * N + n]; (n, k) of W^T at Wb[k * N + n]
    KH_MLP_HIP(gemm(m->dZb[l].as<uint16_t>(), N, 1, m->Wb[l].as<uint16_t>(), 1, N, m->dA[l - 1].as<float>(), K, Mi,
                    K, N, 0.0f));
    hipLaunchKernelGGL(k_relu_back, dim3(blocks(M * K)), dim3(256), 0, 0, m->dA[l - 1].as<float>(),
                       m->Z[l - 1].as<float>(), Mi, K, m->dZb[l - 1].as<uint16_t>(), m->dZbT[l - 1].as<uint16_t>());
    KH_MLP_HIP(hipGetLastError());
  }
  if (loss_out) {
    float s = 0;
    KH_MLP_HIP(hipMemcpy(&s, m->loss.p, 4, hipMemcpyDeviceToHost));
    *loss_out = s / ((float)B * (float)m->nout);
  }
  return KH_OK;
}

kh_status refresh_bf16(kh_mlp m, int l) {
  const int32_t K = m->K(l), N = m->N(l);
  std::vector<float> w((size_t)K * N);
  KH_MLP_HIP(hipMemcpy(w.data(), m->W[l].p, w.size() * 4, hipMemcpyDeviceToHost));
  std::vector<uint16_t> wb(w.size()), wbt(w.size());
  for (int k = 0; k < K; k++)
    for (int n = 0; n < N; n++) {
      wb[(size_t)k * N + n] = to_bf16(w[(size_t)k * N + n]);
      wbt[(size_t)n * K + k] = wb[(size_t)k * N + n];
    }
  KH_MLP_HIP(hipMemcpy(m->Wb[l].p, wb.data(), wb.size() * 2, hipMemcpyHostToDevice));
  KH_MLP_HIP(hipMemcpy(m->WbT[l].p, wbt.data(), wbt.size() * 2, hipMemcpyHostToDevice));
  return KH_OK;
}

}  // namespace

extern "C" {

kh_status kh_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, kh_mlp* out) {
  if (!out || n_in <= 0 || n_hidden <= 0 || n_out <= 0 || n_hidden % 64 != 0) return mlp_fail("bad MLP shape");
  auto m = new kh_mlp_s();
  m->nin = n_in;
  m->nh = n_hidden;
  m->nout = n_out;
  m->Kin = (int32_t)pad(n_in, 64);
  m->H = n_hidden;
  m->Nout = (int32_t)pad(n_out, 64);
  for (int l = 0; l < 3; l++) {
    const int32_t K = m->K(l), N = m->N(l);
    const size_t kn = (size_t)K * N;
    hipError_t e = hipSuccess;
    for (Buf* b : {&m->W[l], &m->mW[l], &m->vW[l], &m->gW[l]})
      if (e == hipSuccess) e = b->ensure(kn * 4);
    for (Buf* b : {&m->b[l], &m->mb[l], &m->vb[l], &m->gb[l]})
      if (e == hipSuccess) e = b->ensure((size_t)N * 4);
    for (Buf* b : {&m->Wb[l], &m->WbT[l]})
      if (e == hipSuccess) e = b->ensure(kn * 2);
    if (e != hipSuccess) {
      delete m;
      g_mlp_err = hipGetErrorString(e);
      return KH_HIP_ERROR;
    }
    // He initialisation from a SplitMix64 stream (Box–Muller normals), padding zero.
    const int32_t fan_in = l == 0 ? n_in : n_hidden, rows = l == 0 ? n_in : n_hidden, cols = l == 2 ? n_out : n_hidden;
    std::vector<float> w(kn, 0.0f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + (uint64_t)l + 1;
    auto next = [&]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    const double sd = std::sqrt(2.0 / fan_in);
    for (int k = 0; k < rows; k++)
      for (int n = 0; n < cols; n++) {
        const double u1 = next(), u2 = next();
        w[(size_t)k * N + n] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
      }
    std::vector<float> zero(kn, 0.0f);
    bool ok = hipMemcpy(m->W[l].p, w.data(), kn * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(m->mW[l].p, zero.data(), kn * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(m->vW[l].p, zero.data(), kn * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemset(m->b[l].p, 0, (size_t)N * 4) == hipSuccess &&
              hipMemset(m->mb[l].p, 0, (size_t)N * 4) == hipSuccess &&
              hipMemset(m->vb[l].p, 0, (size_t)N * 4) == hipSuccess;
    if (!ok || refresh_bf16(m, l) != KH_OK) {
      delete m;
      return KH_HIP_ERROR;
    }
  }
  *out = m;
  return KH_OK;
}

void kh_mlp_destroy(kh_mlp m) { delete m; }

const char* kh_mlp_last_error(void) { return g_mlp_err.c_str(); }

kh_status kh_mlp_forward(kh_mlp m, int32_t B, const float* d_x, float* d_y) {
  if (!m || B <= 0 || !d_x || !d_y) return mlp_fail("bad forward arguments");
  const int64_t M = pad(B, kTile);
  kh_status s = ensure_batch(m, M);
  if (s != KH_OK) return s;
  s = forward(m, d_x, B, M);
  if (s != KH_OK) return s;
  hipLaunchKernelGGL(k_unpad, dim3(blocks((int64_t)B * m->nout)), dim3(256), 0, 0, m->Z[2].as<float>(), B, m->nout,
                     m->Nout, d_y);
  KH_MLP_HIP(hipGetLastError());
  KH_MLP_HIP(hipDeviceSynchronize());
  return KH_OK;
}

kh_status kh_mlp_gradients(kh_mlp m, int32_t B, const float* d_x, const float* d_t, float* loss) {
  if (!m || B <= 0 || !d_x || !d_t) return mlp_fail("bad gradient arguments");
  const int64_t M = pad(B, kTile);
  kh_status s = ensure_batch(m, M);
  if (s != KH_OK) return s;
  s = forward(m, d_x, B, M);
  if (s != KH_OK) return s;
  return backward(m, d_t, B, M, loss);
}

kh_status kh_mlp_adam_step(kh_mlp m, float lr) {
  if (!m || !(lr > 0)) return mlp_fail("bad learning rate");
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  for (int l = 0; l < 3; l++) {
    const int32_t K = m->K(l), N = m->N(l);
    const int64_t kn = (int64_t)K * N;
    hipLaunchKernelGGL(k_adam, dim3(blocks(kn)), dim3(256), 0, 0, m->W[l].as<float>(), m->gW[l].as<float>(),
                       m->mW[l].as<float>(), m->vW[l].as<float>(), kn, lr, b1, b2, eps, c1, c2, K, N,
                       m->Wb[l].as<uint16_t>(), m->WbT[l].as<uint16_t>());
    KH_MLP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_adam, dim3(blocks(N)), dim3(256), 0, 0, m->b[l].as<float>(), m->gb[l].as<float>(),
                       m->mb[l].as<float>(), m->vb[l].as<float>(), (int64_t)N, lr, b1, b2, eps, c1, c2, 1, N,
                       (uint16_t*)nullptr, (uint16_t*)nullptr);
    KH_MLP_HIP(hipGetLastError());
  }
  return KH_OK;
}

// Host copies of parameters (which = 0) or gradients (which = 1), unpadded:
// W1 [n_in][H], b1 [H], W2 [H][H], b2 [H], W3 [H][n_out], b3 [n_out].
kh_status kh_mlp_copy(kh_mlp m, int32_t which, float* host, int64_t count, int32_t to_device) {
  if (!m || !host) return mlp_fail("bad copy arguments");
  int64_t need = 0;
  for (int l = 0; l < 3; l++) {
    const int32_t rows = l == 0 ? m->nin : m->nh, cols = l == 2 ? m->nout : m->nh;
    need += (int64_t)rows * cols + cols;
  }
  if (count != need) return mlp_fail("parameter count is " + std::to_string(need));
  int64_t off = 0;
  for (int l = 0; l < 3; l++) {
    const int32_t rows = l == 0 ? m->nin : m->nh, cols = l == 2 ? m->nout : m->nh, N = m->N(l), K = m->K(l);
    Buf& w = which ? m->gW[l] : m->W[l];
    Buf& b = which ? m->gb[l] : m->b[l];
    std::vector<float> dev((size_t)K * N);
    KH_MLP_HIP(hipMemcpy(dev.data(), w.p, dev.size() * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < rows; k++)
      for (int n = 0; n < cols; n++) {
        float& h = host[off + (int64_t)k * cols + n];
        if (to_device) dev[(size_t)k * N + n] = h;
        else h = dev[(size_t)k * N + n];
      }
    if (to_device) KH_MLP_HIP(hipMemcpy(w.p, dev.data(), dev.size() * 4, hipMemcpyHostToDevice));
    off += (int64_t)rows * cols;
    std::vector<float> bb(N);
    KH_MLP_HIP(hipMemcpy(bb.data(), b.p, (size_t)N * 4, hipMemcpyDeviceToHost));
    for (int n = 0; n < cols; n++) {
      if (to_device) bb[n] = host[off + n];
      else host[off + n] = bb[n];
    }
    if (to_device) KH_MLP_HIP(hipMemcpy(b.p, bb.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    off += cols;
    if (to_device && !which) {
      kh_status s = refresh_bf16(m, l);
      if (s != KH_OK) return s;
    }
  }
  return KH_OK;
}

}  // extern "C"
