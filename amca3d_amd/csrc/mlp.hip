// Temperature surrogate: a 5-128-128-128-128-1 MLP with SiLU, BF16 operands on
// the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16) and FP32 accumulation.
//
// Every layer, forward and backward, is one call of a tiled GEMM
//   C[M x N] = epilogue( opA(A)[M x K] * opB(B)[K x N] )
// with A and B read from FP32 or BF16 memory, optionally transposed, rounded
// to BF16 when staged into LDS, and an epilogue that adds the bias and applies
// SiLU (forward), multiplies by SiLU' (backward), or accumulates atomically
// (weight gradients, split over the batch).  A work-group of 4 waves owns a
// 128 x 128 block of C; each wave owns 32 rows and four 32 x 32 MFMA tiles.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/amca3d.h"
#include "surrogate.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace mlpk {

constexpr int kIn = 5, kHid = 128, kLayers = 5;
constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kWgradRows = 128;  // batch rows per work-group of a weight-gradient GEMM
constexpr int kPad = 8;  // bf16 elements of padding per LDS row
constexpr int kThreads = 256;

// parameter vector layout: W1[5][128] b1[128] W2[128][128] b2 W3 b3 W4 b4 W5[128][1] b5[1]
constexpr int kFanIn[kLayers] = {kIn, kHid, kHid, kHid, kHid};
constexpr int kFanOut[kLayers] = {kHid, kHid, kHid, kHid, 1};

struct Layout {
  int w[kLayers], b[kLayers];
  int total;
  __host__ __device__ Layout() {
    int o = 0;
    for (int l = 0; l < kLayers; l++) {
      w[l] = o;
      o += kFanIn[l] * kFanOut[l];
      b[l] = o;
      o += kFanOut[l];
    }
    total = o;
  }
};

enum Epi : int {
  kEpiStore = 0,     // C = acc (+ bias)
  kEpiBiasSilu = 1,  // Z = acc + bias (FP32), Aout = bf16(silu(Z))
  kEpiDSilu = 2,     // C = acc * silu'(Z)
  kEpiAtomic = 3,    // C += acc (atomic)
};

struct EpiArgs {
  float* C;          // output (FP32), row stride ldc
  int ldc;
  const float* bias; // per column, or null
  float* Z;          // kEpiBiasSilu: pre-activation out; kEpiDSilu: pre-activation in
  __bf16* Aout;      // kEpiBiasSilu: activation out (BF16), row stride ldc
  float* colsum;     // kEpiAtomic with opB = B: column sums of B over the block's K
                     // range (FP32, before rounding) added atomically, or null
};

// v_rcp_f32 (1 ulp) in place of the IEEE divide sequence: the result is
// rounded to bf16 or feeds an fp32 gradient, far coarser than the reciprocal.
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

template <typename T>
__device__ __forceinline__ float to_f(T v) { return (float)v; }

// C = epi(opA(A) * opB(B)).  opA(A)[m][k] = TrA ? A[k*lda + m] : A[m*lda + k];
// opB(B)[k][n] = TrB ? B[n*ldb + k] : B[k*ldb + n].  Out-of-range elements
// stage as zero.  Grid: (ceil(N/128), ceil(M/128), ksplit); block z covers
// K range [z*kchunk, (z+1)*kchunk).
template <typename TA, typename TB, bool TrA, bool TrB, int E>
__global__ void __launch_bounds__(kThreads)
k_gemm(const TA* __restrict__ A, int lda, const TB* __restrict__ B, int ldb, int M, int N, int K,
       int kchunk, EpiArgs e) {
  __shared__ __bf16 As[kBM][kBK + kPad];
  __shared__ __bf16 Bs[kBN][kBK + kPad];  // n-major: Bs[n][k]
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  floatx16 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int i = 0; i < 16; i++) acc[t][i] = 0.0f;
  // column sum of B (thread t stages column t % kBN of every B tile when !TrB)
  constexpr bool kColsum = E == kEpiAtomic && !TrB;
  float cs = 0.0f;

  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    // stage A (128 x 32) and B (32 x 128): 4096 elements each, 16 per thread.
    // Every load is issued before any is used: an element outside the matrix
    // (or the block's K range) loads element 0 and stages as zero.
    float va[(kBM * kBK) / kThreads], vb[(kBN * kBK) / kThreads];
#pragma unroll
    for (int q = 0; q < (kBM * kBK) / kThreads; q++) {
      const int idx = threadIdx.x + q * kThreads;
      int mm, kk;
      if (TrA) {  // consecutive threads walk m (contiguous in memory)
        mm = idx % kBM;
        kk = idx / kBM;
      } else {
        kk = idx % kBK;
        mm = idx / kBK;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      const bool in = gm < M && gk < kend;
      const size_t o = in ? (TrA ? (size_t)gk * lda + gm : (size_t)gm * lda + gk) : 0;
      va[q] = to_f(A[o]);
      if (!in) va[q] = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < (kBN * kBK) / kThreads; q++) {
      const int idx = threadIdx.x + q * kThreads;
      int nn, kk;
      if (TrB) {
        kk = idx % kBK;
        nn = idx / kBK;
      } else {
        nn = idx % kBN;
        kk = idx / kBN;
      }
      const int gn = n0 + nn, gk = k0 + kk;
      const bool in = gn < N && gk < kend;
      const size_t o = in ? (TrB ? (size_t)gn * ldb + gk : (size_t)gk * ldb + gn) : 0;
      vb[q] = to_f(B[o]);
      if (!in) vb[q] = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < (kBM * kBK) / kThreads; q++) {
      const int idx = threadIdx.x + q * kThreads;
      const int mm = TrA ? idx % kBM : idx / kBK, kk = TrA ? idx / kBM : idx % kBK;
      As[mm][kk] = (__bf16)va[q];
    }
#pragma unroll
    for (int q = 0; q < (kBN * kBK) / kThreads; q++) {
      const int idx = threadIdx.x + q * kThreads;
      const int nn = TrB ? idx / kBK : idx % kBN, kk = TrB ? idx % kBK : idx / kBN;
      if (kColsum) cs += vb[q];
      Bs[nn][kk] = (__bf16)vb[q];
    }
    __syncthreads();
    // two MFMA k-steps of 16: lane holds A[row = lane%32][k = (lane/32)*8 ..+8]
    // and B[k = (lane/32)*8 ..+8][col = lane%32]
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 16) {
      const int kq = ks + (lane >> 5) * 8;
      bf16x8 a;
#pragma unroll
      for (int j = 0; j < 8; j++) a[j] = As[wave * 32 + (lane & 31)][kq + j];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; j++) b[j] = Bs[t * 32 + (lane & 31)][kq + j];
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (kColsum && e.colsum) {
    static_assert(kThreads % kBN == 0, "a thread stages one column of B");
    const int gn = n0 + (int)(threadIdx.x % kBN);
    if (blockIdx.y == 0 && gn < N) atomicAdd(&e.colsum[gn], cs);
  }
  // D layout: element i of lane l is row (i/4)*8 + (l/32)*4 + i%4, column l%32.
  // kEpiDSilu reads the pre-activations of the whole tile before using any.
  float zin[4][16];
  if (E == kEpiDSilu) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int m = m0 + wave * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const int n = n0 + t * 32 + (lane & 31);
        const bool in = m < M && n < N;
        zin[t][i] = e.Z[in ? (size_t)m * e.ldc + n : 0];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; t++) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = m0 + wave * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
      const int n = n0 + t * 32 + (lane & 31);
      if (m >= M || n >= N) continue;
      float v = acc[t][i];
      const size_t o = (size_t)m * e.ldc + n;
      if (E == kEpiStore) {
        if (e.bias) v += e.bias[n];
        e.C[o] = v;
      } else if (E == kEpiBiasSilu) {
        const float z = v + (e.bias ? e.bias[n] : 0.0f);
        if (e.Z) e.Z[o] = z;
        e.Aout[o] = (__bf16)silu(z);
      } else if (E == kEpiDSilu) {
        e.C[o] = v * dsilu(zin[t][i]);
      } else {
        atomicAdd(&e.C[o], v);
      }
    }
  }
}

// dT = 2 (T - Y) / M (the gradient of the mean squared error), loss += sum (T - Y)^2
__global__ void k_mse(const float* __restrict__ T, const float* __restrict__ Y, int M,
                      float* __restrict__ dT, float* __restrict__ loss) {
  __shared__ float s[kThreads];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float r = 0.0f;
  if (i < M) {
    r = T[i] - Y[i];
    dT[i] = 2.0f * r / (float)M;
  }
  s[threadIdx.x] = r * r;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, s[0]);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace mlpk

namespace mlpk {
// fused inference (defined with the sampler below): mode 0 rows of X -> network
// output, mode 1 lattice points of (d, s, g) -> temperatures
hipError_t launch_fused(int mode, const float* p, long long n, const float* X, float* out,
                        const amca_surrogate_desc& d, const amca_scan& s, const SurLattice& g,
                        hipStream_t st);
hipError_t launch_fused_timed(amca_mlp* h, int mode, long long n, const float* X, float* out,
                              const amca_surrogate_desc& d, const amca_scan& s,
                              const SurLattice& g);
}  // namespace mlpk

using namespace mlpk;

// Activations of amca_mlp_forward_device_noalloc, which allocates nothing: two
// ping-pong [kNoAllocRows][128] BF16 buffers (16 MiB), one copy per GPU the
// library runs on, held until the process ends.  Host code never passes the
// name itself to a kernel (on the host it is only a stand-in for the symbol);
// each device's address comes from hipGetSymbolAddress (noalloc_scratch).
constexpr int kNoAllocRows = AMCA_MLP_NOALLOC_MAX_ROWS;
static __device__ __align__(256) __bf16 scratch[2 * (size_t)kNoAllocRows * kHid];

struct amca_mlp {
  hipStream_t stream = nullptr;
  int device = -1;                 // the device current at create
  __bf16* noalloc_act = nullptr;   // that device's copy of `scratch` (two [rows][128] buffers)
  Layout lay;
  float* params = nullptr;  // FP32 master copy
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  long long adam_t = 0;
  // per-batch work buffers
  int cap = 0;
  float* X = nullptr;            // [cap][5]
  float* Y = nullptr;            // [cap]
  float* Z[4] = {};              // pre-activations [cap][128]
  __bf16* Aact[4] = {};          // activations [cap][128]
  float* T = nullptr;            // outputs [cap]
  float* dT = nullptr;           // [cap]
  float* dZ[2] = {};             // ping-pong [cap][128]
  float* loss = nullptr;
  // device-resident training set (amca_surrogate_sample)
  float* setX = nullptr;
  float* setY = nullptr;
  long long set_n = 0;
  std::vector<void*> allocs;
  // HIP events around the fused kernel's launches (amca_mlp_enable_timing)
  bool timing = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double fused_ms = 0.0;
  long long fused_launches = 0, fused_rows = 0;
  ~amca_mlp() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocs) (void)hipFree(p);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

thread_local std::string g_mlp_error;

// per-device address of `scratch`, resolved once per device with that device
// current (amca_mlp_create)
std::mutex g_scratch_mu;
std::unordered_map<int, __bf16*> g_scratch_addr;

amca_status mlp_fail(amca_status s, const std::string& m) {
  g_mlp_error = m;
  return s;
}

#define MLP_HIP(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return mlp_fail(AMCA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define MLP_TRY(expr)             \
  do {                            \
    amca_status s_ = (expr);      \
    if (s_ != AMCA_OK) return s_; \
  } while (0)

amca_status noalloc_scratch(int device, __bf16** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mlp_fail(AMCA_ERR_HIP,
                      std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = g_scratch_addr.emplace(device, static_cast<__bf16*>(p)).first;
  }
  *out = it->second;
  return AMCA_OK;
}

template <class T>
amca_status mlp_alloc(amca_mlp* h, T** p, size_t n) {
  void* q = nullptr;
  MLP_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  h->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return AMCA_OK;
}

void mlp_release(amca_mlp* h, void* p) {
  if (!p) return;
  for (auto& a : h->allocs)
    if (a == p) {
      (void)hipFree(a);
      a = nullptr;
    }
}

amca_status ensure_cap(amca_mlp* h, int n) {
  if (n <= h->cap) return AMCA_OK;
  MLP_HIP(hipStreamSynchronize(h->stream));
  mlp_release(h, h->X);
  mlp_release(h, h->Y);
  mlp_release(h, h->T);
  mlp_release(h, h->dT);
  for (int l = 0; l < 4; l++) {
    mlp_release(h, h->Z[l]);
    mlp_release(h, h->Aact[l]);
  }
  for (int q = 0; q < 2; q++) mlp_release(h, h->dZ[q]);
  const size_t c = (size_t)n;
  MLP_TRY(mlp_alloc(h, &h->X, c * kIn));
  MLP_TRY(mlp_alloc(h, &h->Y, c));
  MLP_TRY(mlp_alloc(h, &h->T, c));
  MLP_TRY(mlp_alloc(h, &h->dT, c));
  for (int l = 0; l < 4; l++) {
    MLP_TRY(mlp_alloc(h, &h->Z[l], c * kHid));
    MLP_TRY(mlp_alloc(h, &h->Aact[l], c * kHid));
  }
  for (int q = 0; q < 2; q++) MLP_TRY(mlp_alloc(h, &h->dZ[q], c * kHid));
  h->cap = n;
  return AMCA_OK;
}

template <typename TA, typename TB, bool TrA, bool TrB, int E>
hipError_t gemm(hipStream_t s, const TA* A, int lda, const TB* B, int ldb, int M, int N, int K,
                const EpiArgs& e, int ksplit = 1) {
  const int kchunk = ((K + ksplit - 1) / ksplit + kBK - 1) / kBK * kBK;
  const int kz = (K + kchunk - 1) / kchunk;
  dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, kz);
  hipLaunchKernelGGL((k_gemm<TA, TB, TrA, TrB, E>), grid, dim3(kThreads), 0, s, A, lda, B, ldb, M,
                     N, K, kchunk, e);
  return hipGetLastError();
}

// forward of n rows already in h->X; keeps Z (for a backward pass) when keep_z
// forward of n rows X[n][5] -> T[n] through the activation buffers act[4]
// ([n][128] each; act[l] is written by layer l and read by layer l + 1), keeping
// the pre-activations in h->Z when keep_z
amca_status forward_rows(amca_mlp* h, int n, const float* X, __bf16* const act[4], float* T,
                         bool keep_z) {
  const Layout& L = h->lay;
  hipStream_t s = h->stream;
  for (int l = 0; l < 4; l++) {
    EpiArgs e{nullptr, kHid, h->params + L.b[l], keep_z ? h->Z[l] : nullptr, act[l]};
    if (l == 0)
      MLP_HIP((gemm<float, float, false, false, kEpiBiasSilu>(s, X, kIn, h->params + L.w[0], kHid,
                                                               n, kHid, kIn, e)));
    else
      MLP_HIP((gemm<__bf16, float, false, false, kEpiBiasSilu>(s, act[l - 1], kHid,
                                                                h->params + L.w[l], kHid, n, kHid,
                                                                kHid, e)));
  }
  EpiArgs e{T, 1, h->params + L.b[4], nullptr, nullptr};
  MLP_HIP((gemm<__bf16, float, false, false, kEpiStore>(s, act[3], kHid, h->params + L.w[4], 1, n,
                                                         1, kHid, e)));
  return AMCA_OK;
}

amca_status forward(amca_mlp* h, int n, bool keep_z = true) {
  return forward_rows(h, n, h->X, h->Aact, h->T, keep_z);
}

// gradients of the mean squared error over n rows (X, Y on the device) into h->grads
amca_status backward(amca_mlp* h, int n) {
  const Layout& L = h->lay;
  hipStream_t s = h->stream;
  MLP_HIP(hipMemsetAsync(h->grads, 0, sizeof(float) * L.total, s));
  MLP_HIP(hipMemsetAsync(h->loss, 0, sizeof(float), s));
  hipLaunchKernelGGL(k_mse, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, h->T, h->Y,
                     n, h->dT, h->loss);
  MLP_HIP(hipGetLastError());
  // weight gradients: one 128 x 128 tile per layer, split over the batch in
  // chunks of kWgradRows rows so that the split fills the chip; the bias
  // gradient (column sums of dZ) is taken while dZ is staged
  const int ksplit = std::max(1, (n + kWgradRows - 1) / kWgradRows);
  // layer 5: dW5 = A4^T dT, db5 = sum dT, dZ4 = (dT W5^T) * silu'(Z4)
  {
    EpiArgs e{h->grads + L.w[4], 1, nullptr, nullptr, nullptr, h->grads + L.b[4]};
    MLP_HIP((gemm<__bf16, float, true, false, kEpiAtomic>(s, h->Aact[3], kHid, h->dT, 1, kHid, 1, n,
                                                           e, ksplit)));
    EpiArgs d{h->dZ[0], kHid, nullptr, h->Z[3], nullptr};
    MLP_HIP((gemm<float, float, false, true, kEpiDSilu>(s, h->dT, 1, h->params + L.w[4], 1, n, kHid,
                                                         1, d)));
  }
  int cur = 0;
  for (int l = 3; l >= 0; l--) {
    float* dz = h->dZ[cur];
    // dW_l = A_{l-1}^T dZ_l (A_{-1} = X), db_l = sum dZ_l
    EpiArgs e{h->grads + L.w[l], kHid, nullptr, nullptr, nullptr, h->grads + L.b[l]};
    if (l == 0)
      MLP_HIP((gemm<float, float, true, false, kEpiAtomic>(s, h->X, kIn, dz, kHid, kIn, kHid, n, e,
                                                            ksplit)));
    else
      MLP_HIP((gemm<__bf16, float, true, false, kEpiAtomic>(s, h->Aact[l - 1], kHid, dz, kHid, kHid,
                                                             kHid, n, e, ksplit)));
    if (l > 0) {
      // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1})
      EpiArgs d{h->dZ[cur ^ 1], kHid, nullptr, h->Z[l - 1], nullptr};
      MLP_HIP((gemm<float, float, false, true, kEpiDSilu>(s, dz, kHid, h->params + L.w[l], kHid, n,
                                                           kHid, kHid, d)));
      cur ^= 1;
    }
  }
  return AMCA_OK;
}

}  // namespace

extern "C" {

const char* amca_mlp_last_error(void) { return g_mlp_error.c_str(); }

int32_t amca_mlp_param_count(void) { return Layout().total; }

amca_status amca_mlp_create(uint64_t seed, amca_mlp** out) {
  if (!out) return mlp_fail(AMCA_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return mlp_fail(AMCA_ERR_NO_DEVICE, "no HIP device visible");
  auto* h = new amca_mlp();
  amca_status st = AMCA_OK;
  do {
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
      st = mlp_fail(AMCA_ERR_HIP, "stream");
      break;
    }
    if (hipGetDevice(&h->device) != hipSuccess) {
      st = mlp_fail(AMCA_ERR_HIP, "hipGetDevice");
      break;
    }
    if ((st = noalloc_scratch(h->device, &h->noalloc_act)) != AMCA_OK) break;
    const int P = h->lay.total;
    if ((st = mlp_alloc(h, &h->params, P)) != AMCA_OK) break;
    if ((st = mlp_alloc(h, &h->grads, P)) != AMCA_OK) break;
    if ((st = mlp_alloc(h, &h->adam_m, P)) != AMCA_OK) break;
    if ((st = mlp_alloc(h, &h->adam_v, P)) != AMCA_OK) break;
    if ((st = mlp_alloc(h, &h->loss, 1)) != AMCA_OK) break;
    // He-uniform weights for SiLU layers, zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> p(P, 0.0f);
    for (int l = 0; l < kLayers; l++) {
      const float a = std::sqrt(6.0f / (float)kFanIn[l]);
      std::uniform_real_distribution<float> U(-a, a);
      for (int i = 0; i < kFanIn[l] * kFanOut[l]; i++) p[h->lay.w[l] + i] = U(rng);
    }
    if (hipMemcpy(h->params, p.data(), sizeof(float) * P, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->adam_m, 0, sizeof(float) * P) != hipSuccess ||
        hipMemset(h->adam_v, 0, sizeof(float) * P) != hipSuccess) {
      st = mlp_fail(AMCA_ERR_HIP, "upload");
      break;
    }
  } while (0);
  if (st != AMCA_OK) {
    delete h;
    return st;
  }
  *out = h;
  return AMCA_OK;
}

amca_status amca_mlp_destroy(amca_mlp* h) {
  delete h;
  return AMCA_OK;
}

amca_status amca_mlp_set_params(amca_mlp* h, const float* p) {
  if (!h || !p) return mlp_fail(AMCA_ERR_INVALID, "null argument");
  MLP_HIP(hipMemcpy(h->params, p, sizeof(float) * h->lay.total, hipMemcpyHostToDevice));
  MLP_HIP(hipMemset(h->adam_m, 0, sizeof(float) * h->lay.total));
  MLP_HIP(hipMemset(h->adam_v, 0, sizeof(float) * h->lay.total));
  h->adam_t = 0;
  return AMCA_OK;
}

amca_status amca_mlp_get_params(amca_mlp* h, float* p) {
  if (!h || !p) return mlp_fail(AMCA_ERR_INVALID, "null argument");
  MLP_HIP(hipStreamSynchronize(h->stream));
  MLP_HIP(hipMemcpy(p, h->params, sizeof(float) * h->lay.total, hipMemcpyDeviceToHost));
  return AMCA_OK;
}

amca_status amca_mlp_forward(amca_mlp* h, int32_t n, const float* x5, float* t) {
  if (!h || n < 1 || !x5 || !t) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_cap(h, n));
  MLP_HIP(hipMemcpyAsync(h->X, x5, sizeof(float) * kIn * n, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(forward(h, n));
  MLP_HIP(hipMemcpyAsync(t, h->T, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

amca_status amca_mlp_gradients(amca_mlp* h, int32_t n, const float* x5, const float* y,
                               float* grads, float* loss) {
  if (!h || n < 1 || !x5 || !y) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_cap(h, n));
  MLP_HIP(hipMemcpyAsync(h->X, x5, sizeof(float) * kIn * n, hipMemcpyHostToDevice, h->stream));
  MLP_HIP(hipMemcpyAsync(h->Y, y, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  MLP_TRY(forward(h, n));
  MLP_TRY(backward(h, n));
  if (grads)
    MLP_HIP(hipMemcpyAsync(grads, h->grads, sizeof(float) * h->lay.total, hipMemcpyDeviceToHost,
                           h->stream));
  float l = 0.0f;
  MLP_HIP(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MLP_HIP(hipStreamSynchronize(h->stream));
  if (loss) *loss = l / (float)n;
  return AMCA_OK;
}

// One Adam step per batch over device-resident samples X[n][5], Y[n].
amca_status amca_mlp_train_device(amca_mlp* h, int64_t n, const float* dX, const float* dY,
                                  int32_t batch, int32_t steps, float lr, uint64_t seed,
                                  float* loss_trace) {
  if (!h || n < batch || batch < 1 || steps < 0 || !dX || !dY)
    return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_cap(h, batch));
  std::mt19937_64 rng(seed);
  const long long nb = n / batch;
  for (int it = 0; it < steps; it++) {
    const long long b = (long long)(rng() % (unsigned long long)nb);
    MLP_HIP(hipMemcpyAsync(h->X, dX + (size_t)b * batch * kIn, sizeof(float) * kIn * batch,
                           hipMemcpyDeviceToDevice, h->stream));
    MLP_HIP(hipMemcpyAsync(h->Y, dY + (size_t)b * batch, sizeof(float) * batch,
                           hipMemcpyDeviceToDevice, h->stream));
    MLP_TRY(forward(h, batch));
    MLP_TRY(backward(h, batch));
    h->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)h->adam_t);
    const float c2 = 1.0f - std::pow(b2, (float)h->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((h->lay.total + 255) / 256), dim3(256), 0, h->stream, h->params,
                       h->grads, h->adam_m, h->adam_v, h->lay.total, lr, b1, b2, 1e-8f, c1, c2);
    MLP_HIP(hipGetLastError());
    if (loss_trace) {
      float l = 0.0f;
      MLP_HIP(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
      MLP_HIP(hipStreamSynchronize(h->stream));
      loss_trace[it] = l / (float)batch;
    }
  }
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

// Forward over device-resident rows (no copies): T[n] = mlp(X[n][5]).
amca_status amca_mlp_forward_fused(amca_mlp* h, int32_t n, const float* x5, float* t) {
  if (!h || n < 1 || !x5 || !t) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_cap(h, n));
  MLP_HIP(hipMemcpyAsync(h->X, x5, sizeof(float) * kIn * n, hipMemcpyHostToDevice, h->stream));
  MLP_HIP(launch_fused_timed(h, 0, n, h->X, h->T, amca_surrogate_desc{}, amca_scan{},
                             SurLattice{}));
  MLP_HIP(hipMemcpyAsync(t, h->T, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

amca_status amca_mlp_enable_timing(amca_mlp* h, int32_t on) {
  if (!h) return mlp_fail(AMCA_ERR_INVALID, "null handle");
  for (hipEvent_t& e : h->ev)
    if (!e) MLP_HIP(hipEventCreate(&e));
  h->timing = on != 0;
  h->fused_ms = 0.0;
  h->fused_launches = h->fused_rows = 0;
  return AMCA_OK;
}

amca_status amca_mlp_fused_times(amca_mlp* h, double* ms, int64_t* launches, int64_t* rows) {
  if (!h || !ms || !launches || !rows) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  *ms = h->fused_ms;
  *launches = h->fused_launches;
  *rows = h->fused_rows;
  return AMCA_OK;
}

amca_status amca_mlp_forward_device(amca_mlp* h, int64_t n, const float* dX, float* dT) {
  if (!h || n < 1 || !dX || !dT) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  const int chunk = 1 << 20;
  MLP_TRY(ensure_cap(h, (int)std::min<int64_t>(n, chunk)));
  for (int64_t r = 0; r < n; r += chunk) {
    const int m = (int)std::min<int64_t>(chunk, n - r);
    MLP_HIP(hipMemcpyAsync(h->X, dX + r * kIn, sizeof(float) * kIn * m, hipMemcpyDeviceToDevice,
                           h->stream));
    MLP_TRY(forward(h, m, false));
    MLP_HIP(hipMemcpyAsync(dT + r, h->T, sizeof(float) * m, hipMemcpyDeviceToDevice, h->stream));
  }
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

// amca_mlp_forward_device's kernels with the activations in the device's copy
// of `scratch` and the caller's rows read and written in place: no hipMalloc,
// no hipFree, no staging copies.  Every buffer element a layer reads was written
// by the layer before it in the same call, so nothing needs clearing.
amca_status amca_mlp_forward_device_noalloc(amca_mlp* h, int64_t n, const float* dX, float* dT) {
  if (!h || n < 1 || !dX || !dT) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  if (n > kNoAllocRows)
    return mlp_fail(AMCA_ERR_INVALID, "n = " + std::to_string(n) +
                                          " rows exceeds the no-allocation maximum of " +
                                          std::to_string(kNoAllocRows) + " rows per call");
  int dev = -1;
  MLP_HIP(hipGetDevice(&dev));
  if (dev != h->device || !h->noalloc_act)
    return mlp_fail(AMCA_ERR_INVALID, "make the network's device (" + std::to_string(h->device) +
                                          ") current before the call");
  __bf16* a0 = h->noalloc_act;
  __bf16* a1 = h->noalloc_act + (size_t)kNoAllocRows * kHid;
  __bf16* const act[4] = {a0, a1, a0, a1};
  MLP_TRY(forward_rows(h, (int)n, dX, act, dT, false));
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

}  // extern "C"

extern "C" amca_status amca_mlp_gemm_selftest(int32_t M, int32_t N, int32_t K, int32_t trans_a,
                                              int32_t trans_b, const float* A, const float* B,
                                              float* C) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  amca_status st = AMCA_OK;
  hipStream_t s = nullptr;
  do {
    if (hipMalloc(&dA, sizeof(float) * M * K) != hipSuccess ||
        hipMalloc(&dB, sizeof(float) * K * N) != hipSuccess ||
        hipMalloc(&dC, sizeof(float) * M * N) != hipSuccess || hipStreamCreate(&s) != hipSuccess) {
      st = mlp_fail(AMCA_ERR_HIP, "alloc");
      break;
    }
    (void)hipMemcpy(dA, A, sizeof(float) * M * K, hipMemcpyHostToDevice);
    (void)hipMemcpy(dB, B, sizeof(float) * K * N, hipMemcpyHostToDevice);
    EpiArgs e{dC, N, nullptr, nullptr, nullptr};
    const int lda = trans_a ? M : K, ldb = trans_b ? K : N;
    hipError_t r;
    if (!trans_a && !trans_b) r = gemm<float, float, false, false, kEpiStore>(s, dA, lda, dB, ldb, M, N, K, e);
    else if (trans_a && !trans_b) r = gemm<float, float, true, false, kEpiStore>(s, dA, lda, dB, ldb, M, N, K, e);
    else if (!trans_a && trans_b) r = gemm<float, float, false, true, kEpiStore>(s, dA, lda, dB, ldb, M, N, K, e);
    else r = gemm<float, float, true, true, kEpiStore>(s, dA, lda, dB, ldb, M, N, K, e);
    if (r != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      st = mlp_fail(AMCA_ERR_HIP, "gemm");
      break;
    }
    (void)hipMemcpy(C, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost);
  } while (0);
  if (dA) (void)hipFree(dA);
  if (dB) (void)hipFree(dB);
  if (dC) (void)hipFree(dC);
  if (s) (void)hipStreamDestroy(s);
  return st;
}

// ===========================================================================
// Training data and lattice evaluation of the temperature surrogate.


namespace mlpk {

struct Sampler {
  amca_surrogate_desc d;
  amca_scan s[16];
  int n_scans;
};

__device__ __forceinline__ unsigned long long splitmix(unsigned long long& st) {
  unsigned long long z = (st += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(unsigned long long& st) {
  return (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double rosenthal(const amca_surrogate_desc& d, const amca_scan& s,
                                            double x, double y, double z, double t) {
  const double xi = x - (s.x_start + s.v * t);
  const double dy = y - s.y_c, dz = z - d.z_top;
  double R = sqrt(xi * xi + dy * dy + dz * dz);
  if (R < d.r_min) R = d.r_min;
  return d.T0 + d.eta * s.P / (2.0 * 3.14159265358979323846 * d.k) / R *
                    exp(-s.v * (R + xi) / (2.0 * d.alpha));
}

// the coupled domain's FE path at (x, y, z, t): nodal frames of the source at
// t_f and t_f+1, linear in time (void rule), hex8 in space
__device__ double fe_temperature(const amca_surrogate_desc& d, const amca_scan& s, double x,
                                 double y, double z, double t) {
  int e[3];
  double p[3];
  const double o[3] = {d.fx0, d.fy0, d.fz0};
  const int n[3] = {d.fnx, d.fny, d.fnz};
  const double q[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    int i = (int)floor((q[a] - o[a]) / d.fdx);
    i = i < 0 ? 0 : (i > n[a] - 2 ? n[a] - 2 : i);
    e[a] = i;
    p[a] = 2.0 * ((q[a] - (o[a] + i * d.fdx)) / d.fdx) - 1.0;
  }
  int f = (int)floor(t / d.frame_dt);
  f = f < 0 ? 0 : (f > d.n_frames - 2 ? d.n_frames - 2 : f);
  const double r = (t - f * d.frame_dt) / d.frame_dt;
  const double xi = p[0], eta = p[1], zeta = p[2];
  const double sf[8] = {
      0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 - zeta), 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 - zeta),
      0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 - zeta), 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 - zeta),
      0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 + zeta), 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 + zeta),
      0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 + zeta), 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 + zeta)};
  const int di[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1},
            dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  double T = 0.0;
  for (int m = 0; m < 8; m++) {
    const double nx = d.fx0 + (e[0] + di[m]) * d.fdx, ny = d.fy0 + (e[1] + dj[m]) * d.fdx,
                 nz = d.fz0 + (e[2] + dk[m]) * d.fdx;
    const double th1 = rosenthal(d, s, nx, ny, nz, f * d.frame_dt);
    const double th0 = rosenthal(d, s, nx, ny, nz, (f + 1) * d.frame_dt);
    const double v = d.void_temperature;
    const double th = (th0 <= v) ? v : ((th1 <= v) ? th0 : r * (th0 - th1) + th1);
    T += sf[m] * th;
  }
  return T;
}

__device__ __forceinline__ void inputs(const amca_surrogate_desc& d, const amca_scan& s, double x,
                                       double y, double z, double t, float* o) {
  o[0] = (float)((x - s.x_start - s.v * t) / d.xi_scale);
  o[1] = (float)((y - s.y_c) / d.y_scale);
  o[2] = (float)((z - d.z_top) / d.z_scale);
  o[3] = (float)(s.P / d.P_scale);
  o[4] = (float)(s.v / d.v_scale);
}

__device__ __host__ inline double to_target(const amca_surrogate_desc& d, double T) {
  const double u = (fmin(T, d.T_clip) - d.T_ref) / d.T_scale;
  return d.log_target ? log1p(fmax(u, 0.0)) : u;
}
__device__ __host__ inline double from_target(const amca_surrogate_desc& d, double y) {
  return d.T_ref + d.T_scale * (d.log_target ? expm1(y) : y);
}

__global__ void k_sample(Sampler S, long long n, double near_fraction, unsigned long long seed,
                         float* __restrict__ X, float* __restrict__ Y) {
  const amca_surrogate_desc& d = S.d;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    unsigned long long st = seed ^ ((unsigned long long)i * 0xd1b54a32d192ed03ull);
    const amca_scan& s = S.s[i % S.n_scans];
    const double t = unit(st) * (d.n_frames - 1) * d.frame_dt;
    const double lx = (d.fnx - 1) * d.fdx, ly = (d.fny - 1) * d.fdx, lz = (d.fnz - 1) * d.fdx;
    double x, y, z;
    if (unit(st) < near_fraction) {
      // close behind the source, where the field varies: 40 xi_scale behind,
      // 4 ahead, 8 y_scale across, 8 z_scale deep
      x = s.x_start + s.v * t + (unit(st) * 44.0 - 40.0) * d.xi_scale;
      y = s.y_c + (unit(st) * 2.0 - 1.0) * 8.0 * d.y_scale;
      z = d.z_top - unit(st) * 8.0 * d.z_scale;
    } else {
      x = d.fx0 + unit(st) * lx;
      y = d.fy0 + unit(st) * ly;
      z = d.fz0 + unit(st) * lz;
    }
    x = fmin(fmax(x, d.fx0), d.fx0 + lx);
    y = fmin(fmax(y, d.fy0), d.fy0 + ly);
    z = fmin(fmax(z, d.fz0), d.fz0 + lz);
    const double T = fe_temperature(d, s, x, y, z, t);
    inputs(d, s, x, y, z, t, X + 5 * i);
    Y[i] = (float)to_target(d, T);
  }
}

__global__ void k_lattice_inputs(amca_surrogate_desc d, amca_scan s, SurLattice g, long long r0,
                                 int m, float* __restrict__ X) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const long long q = r0 + r;
  const int i = (int)(q % g.nxi);
  const int j = (int)((q / g.nxi) % g.ny);
  const int k = (int)(q / ((long long)g.nxi * g.ny));
  // a lattice point is (xi, y, z) in the moving frame: feed x = xi + x_start at t = 0
  inputs(d, s, g.xi0 + i * g.dxi + s.x_start, g.y0 + j * g.dy, g.z0 + k * g.dz, 0.0, X + 5 * r);
}

__global__ void k_denorm(const float* __restrict__ t, int m, amca_surrogate_desc d,
                         float* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  out[r] = (float)from_target(d, (double)t[r]);
}

// ---------------------------------------------------------------------------
// Fused inference: the whole network per 64 rows in one wave (two halves of 32
// whose MFMAs and SiLU epilogues interleave), weights in LDS, activations in
// registers.  Layer l computes D'[out][row] = W_l^T act^T with
// v_mfma_f32_32x32x16_bf16 (A = weights, rows = output features; B =
// activations, columns = batch rows).  The accumulator of a 32 x 32 tile holds,
// in lane L and register i, output feature (i/4)*8 + (L/32)*4 + i%4 of batch
// row L%32: after bias, SiLU and BF16 rounding, registers 8s..8s+7 of tile t
// are exactly the B operand of the next layer's k-step 2t+s, provided that
// k-step's k order is the one those registers carry.  The weights are laid out
// in LDS in that order, so no activation moves between lanes.
//   k-step ks, lane group g (= L/32), element ii  <->  input feature
//   16*ks + 8*(ii/4) + 4*g + ii%4
// The hidden layers carry y = -log2(e) z instead of the pre-activation z, so
// that sigmoid(z) = 1 / (1 + 2^y) is one v_exp_f32 and one v_rcp_f32 with no
// scaling multiply: the stored activation is a' = y sigmoid(z) = -log2(e)
// silu(z), whose factor cancels against the -log2(e) the next layer's bias
// carries (y_next = W a' - log2(e) b = -log2(e) z_next).  Layer 1 scales its
// accumulator once (one FMA with the scaled bias); layer 5 is an MFMA with W5
// in every row of A and scales its output by -1/log2(e) before adding b5.  The
// biases of layers 2..4 are loaded into the accumulators before the first
// MFMA, so those layers add no bias either.
// Inputs and weights round to BF16 as in the layered path; the activations
// round to BF16 as scaled values, which differ from the layered path's in the
// last BF16 place.

constexpr int kFusedWaves = 8;
constexpr int kWtStride = kHid + 8;  // bf16 per LDS row (padding spreads the banks)

struct FusedLds {
  __bf16 wt[3][kHid][kWtStride];  // layers 2..4: wt[l][out][16*ks + 8*g + ii]
  __bf16 w1t[kHid][8];            // layer 1: W1[k][out] for k < 5, zero above
  float b[4][kHid];               // biases of layers 1..4, times -log2(e)
  __bf16 w5[8][2][8];             // W5 in BF16, as the A fragment of k-step ks, lane group g
  float b5;
};

constexpr float kNegLog2e = -1.4426950408889634f;
constexpr float kNegLn2 = -0.6931471805599453f;  // -1/log2(e)

__device__ __forceinline__ int fused_feature(int ks, int g, int ii) {
  return 16 * ks + 8 * (ii >> 2) + 4 * g + (ii & 3);
}

// parameter offsets of the layered layout: W1 at 0, b1 at 640; layer l >= 2
// (0-based 1..3) at 768 + (l-1) * 16512, its bias 16384 further; W5, b5 last
__device__ __forceinline__ int fused_w(int l) { return l == 0 ? 0 : 768 + (l - 1) * 16512; }
__device__ __forceinline__ int fused_b(int l) { return l == 0 ? 640 : fused_w(l) + kHid * kHid; }

__device__ __forceinline__ void fused_load(const float* __restrict__ p, FusedLds& L) {
  for (int q = threadIdx.x; q < 3 * kHid * kHid; q += blockDim.x) {
    const int l = q / (kHid * kHid), out = (q / kHid) % kHid, kk = q % kHid;
    const int in = fused_feature(kk >> 4, (kk >> 3) & 1, kk & 7);
    L.wt[l][out][kk] = (__bf16)p[fused_w(l + 1) + in * kHid + out];
  }
  for (int q = threadIdx.x; q < kHid * 8; q += blockDim.x) {
    const int out = q / 8, k = q % 8;
    L.w1t[out][k] = (__bf16)(k < kIn ? p[k * kHid + out] : 0.0f);
  }
  for (int q = threadIdx.x; q < 4 * kHid; q += blockDim.x)
    L.b[q / kHid][q % kHid] = kNegLog2e * p[fused_b(q / kHid) + q % kHid];
  const int w5 = fused_b(3) + kHid;
  for (int q = threadIdx.x; q < kHid; q += blockDim.x) {
    const int ks = q >> 4, g = (q >> 3) & 1, ii = q & 7;
    L.w5[ks][g][ii] = (__bf16)p[w5 + fused_feature(ks, g, ii)];
  }
  if (threadIdx.x == 0) L.b5 = p[w5 + kHid];
  __syncthreads();
}

// Chunk ks of the epilogue: a' = y / (1 + 2^y) of the eight accumulator
// registers 8(ks%2).. of tile ks/2, rounded to BF16, into the next layer's B
// operand bop[ks]; kFirst: y = -log2(e) acc + bias (layer 1), else y = acc.
template <bool kFirst>
__device__ __forceinline__ void fused_epi_chunk(const floatx16* acc, const float* bias, int g,
                                                int ks, bf16x8* bop) {
  const int t = ks >> 1, s = ks & 1;
#pragma unroll
  for (int ii = 0; ii < 8; ii++) {
    float y = acc[t][8 * s + ii];
    if (kFirst) y = fmaf(y, kNegLog2e, bias[fused_feature(ks, g, ii)]);
    bop[ks][ii] = (__bf16)(y * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y)));
  }
}

// One phase of the two-half schedule: the 32 MFMAs of hidden layer l + 2
// (l = 0..2) for half P, interleaved with the epilogue of the previous layer
// of half Q, which needs no MFMA result of this phase.  The accumulators of P
// start at the layer's bias (register 4q+j of tile ot holds output feature
// 32 ot + 8 q + 4 g + j); the weight fragments of k-step ks+1 are read from
// LDS while those of k-step ks feed the MFMAs.
template <bool kFirstEpi>
__device__ __forceinline__ void fused_phase(const FusedLds& L, int l, int r, int g,
                                            floatx16* accP, const bf16x8* bopP,
                                            const floatx16* accQ, bf16x8* bopQ) {
#pragma unroll
  for (int ot = 0; ot < 4; ot++) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float4 b = *reinterpret_cast<const float4*>(&L.b[l + 1][32 * ot + 8 * q + 4 * g]);
      accP[ot][4 * q] = b.x;
      accP[ot][4 * q + 1] = b.y;
      accP[ot][4 * q + 2] = b.z;
      accP[ot][4 * q + 3] = b.w;
    }
  }
  const __bf16* wl = &L.wt[l][r][8 * g];
  bf16x8 a[2][4];
#pragma unroll
  for (int ot = 0; ot < 4; ot++)
    a[0][ot] = *reinterpret_cast<const bf16x8*>(wl + 32 * ot * kWtStride);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int ks = 0; ks < 8; ks++) {
    if (ks + 1 < 8) {
#pragma unroll
      for (int ot = 0; ot < 4; ot++)
        a[(ks + 1) & 1][ot] =
            *reinterpret_cast<const bf16x8*>(wl + 32 * ot * kWtStride + 16 * (ks + 1));
    }
#pragma unroll
    for (int ot = 0; ot < 4; ot++)
      accP[ot] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks & 1][ot], bopP[ks], accP[ot], 0, 0, 0);
    fused_epi_chunk<kFirstEpi>(accQ, L.b[0], g, ks, bopQ);
    // LDS reads first, then each MFMA followed by a share of the epilogue
    __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 10, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Layer 5 of half P (8 MFMAs, A = W5 in every row, so every accumulator
// register of lane L holds the output of row L % 32), optionally interleaved
// with the last epilogue of half Q.
template <bool kEpi>
__device__ __forceinline__ float fused_out(const FusedLds& L, int g, const bf16x8* bopP,
                                           const floatx16* accQ, bf16x8* bopQ) {
  floatx16 acc;
#pragma unroll
  for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
  for (int ks = 0; ks < 8; ks++) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&L.w5[ks][g][0]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bopP[ks], acc, 0, 0, 0);
    if (kEpi) {
      fused_epi_chunk<false>(accQ, nullptr, g, ks, bopQ);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 40, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  return fmaf(acc[0], kNegLn2, L.b5);
}

// Network output (the training target scale) of the 64 rows whose inputs the
// lanes hold in x[0..4] (lane L: row L of the group), returned in lane L.
// Half X is rows 0..31, half Y rows 32..63; each runs the whole network, and
// the MFMAs of one half overlap the SiLU epilogue of the other.
__device__ __forceinline__ float fused_rows(const FusedLds& L, const float* x) {
  const int lane = threadIdx.x & 63, g = lane >> 5, r = lane & 31;
  floatx16 accX[4], accY[4];
  bf16x8 bopX[8], bopY[8];
  // layer 1: K = 5, padded to one k-step of 16.  Half X takes k 0..7 (lane
  // group 0 holds its rows), half Y k 8..15 (lane group 1); the weight rows
  // repeat in both k halves and the other half's inputs are zero.
  bf16x8 bx, by;
#pragma unroll
  for (int ii = 0; ii < 8; ii++) {
    const __bf16 v = (__bf16)(ii < kIn ? x[ii] : 0.0f);
    bx[ii] = g == 0 ? v : (__bf16)0.0f;
    by[ii] = g == 1 ? v : (__bf16)0.0f;
  }
#pragma unroll
  for (int ot = 0; ot < 4; ot++) {
    const bf16x8 w = *reinterpret_cast<const bf16x8*>(&L.w1t[32 * ot + r][0]);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      accX[ot][i] = 0.0f;
      accY[ot][i] = 0.0f;
    }
    accX[ot] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, bx, accX[ot], 0, 0, 0);
    accY[ot] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, by, accY[ot], 0, 0, 0);
  }
#pragma unroll
  for (int ks = 0; ks < 8; ks++) fused_epi_chunk<true>(accX, L.b[0], g, ks, bopX);
  // hidden layers 2..4
  fused_phase<true>(L, 0, r, g, accX, bopX, accY, bopY);
  fused_phase<false>(L, 0, r, g, accY, bopY, accX, bopX);
  fused_phase<false>(L, 1, r, g, accX, bopX, accY, bopY);
  fused_phase<false>(L, 1, r, g, accY, bopY, accX, bopX);
  fused_phase<false>(L, 2, r, g, accX, bopX, accY, bopY);
  fused_phase<false>(L, 2, r, g, accY, bopY, accX, bopX);
  const float tx = fused_out<true>(L, g, bopX, accY, bopY);
  const float ty = fused_out<false>(L, g, bopY, nullptr, nullptr);
  return g == 0 ? tx : ty;
}

// mode 0: rows of X (n x 5) -> T (network output); mode 1: lattice points ->
// temperatures (from_target) into out
template <int kMode>
__global__ void __launch_bounds__(64 * kFusedWaves)
k_mlp_fused(const float* __restrict__ p, long long n, const float* __restrict__ X,
            float* __restrict__ out, amca_surrogate_desc d, amca_scan s, SurLattice g) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  FusedLds& L = *reinterpret_cast<FusedLds*>(lds_raw);
  fused_load(p, L);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long groups = (n + 63) / 64;
  for (long long gr = (long long)blockIdx.x * kFusedWaves + wave; gr < groups;
       gr += (long long)gridDim.x * kFusedWaves) {
    const long long row = gr * 64 + lane;
    float x[kIn] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (row < n) {
      if (kMode == 0) {
#pragma unroll
        for (int q = 0; q < kIn; q++) x[q] = X[row * kIn + q];
      } else {
        const int i = (int)(row % g.nxi);
        const int j = (int)((row / g.nxi) % g.ny);
        const int k = (int)(row / ((long long)g.nxi * g.ny));
        inputs(d, s, g.xi0 + i * g.dxi + s.x_start, g.y0 + j * g.dy, g.z0 + k * g.dz, 0.0, x);
      }
    }
    const float t = fused_rows(L, x);
    if (row < n) out[row] = kMode == 0 ? t : (float)from_target(d, (double)t);
  }
}

hipError_t launch_fused(int mode, const float* p, long long n, const float* X, float* out,
                        const amca_surrogate_desc& d, const amca_scan& s, const SurLattice& g,
                        hipStream_t st) {
  const size_t lds = sizeof(FusedLds);
  long long blocks = (n + 64 * kFusedWaves - 1) / (64 * kFusedWaves);
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  if (mode == 0) {
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_fused<0>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_fused<0>, dim3((unsigned)blocks), dim3(64 * kFusedWaves), lds, st, p,
                       n, X, out, d, s, g);
  } else {
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_fused<1>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_fused<1>, dim3((unsigned)blocks), dim3(64 * kFusedWaves), lds, st, p,
                       n, X, out, d, s, g);
  }
  return hipGetLastError();
}

// launch_fused, bracketed by HIP events on the handle's stream when timing is
// on (the launch is waited for and its duration added to the totals)
hipError_t launch_fused_timed(amca_mlp* h, int mode, long long n, const float* X, float* out,
                              const amca_surrogate_desc& d, const amca_scan& s,
                              const SurLattice& g) {
  if (!h->timing) return launch_fused(mode, h->params, n, X, out, d, s, g, h->stream);
  hipError_t e = hipEventRecord(h->ev[0], h->stream);
  if (e == hipSuccess) e = launch_fused(mode, h->params, n, X, out, d, s, g, h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev[1], h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev[1]);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  if (e == hipSuccess) {
    h->fused_ms += ms;
    h->fused_launches++;
    h->fused_rows += n;
  }
  return e;
}

__global__ void k_eval_points(amca_surrogate_desc d, amca_scan s, int n, const double* __restrict__ p,
                              float* __restrict__ X, double* __restrict__ Tfe) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double x = p[4 * r], y = p[4 * r + 1], z = p[4 * r + 2], t = p[4 * r + 3];
  inputs(d, s, x, y, z, t, X + 5 * r);
  Tfe[r] = fe_temperature(d, s, x, y, z, t);
}

}  // namespace mlpk

const char* surrogate_error() { return g_mlp_error.c_str(); }

amca_status surrogate_eval_lattice(amca_mlp* h, const amca_surrogate_desc& d, const amca_scan& s,
                                   const SurLattice& g, float* out) {
  const long long total = (long long)g.nxi * g.ny * g.nz;
  MLP_HIP(launch_fused_timed(h, 1, total, nullptr, out, d, s, g));
  MLP_HIP(hipStreamSynchronize(h->stream));
  return AMCA_OK;
}

extern "C" {

amca_status amca_surrogate_sample(amca_mlp* h, const amca_surrogate_desc* d, const amca_scan* scans,
                                  int32_t n_scans, int64_t n, double near_fraction, uint64_t seed) {
  if (!h || !d || !scans || n_scans < 1 || n_scans > 16 || n < 1)
    return mlp_fail(AMCA_ERR_INVALID, "bad arguments (1..16 scans)");
  if (d->fnx < 2 || d->fny < 2 || d->fnz < 2 || d->n_frames < 2 || !(d->fdx > 0) ||
      !(d->frame_dt > 0) || !(d->xi_scale > 0) || !(d->T_scale > 0))
    return mlp_fail(AMCA_ERR_INVALID, "bad surrogate description");
  MLP_HIP(hipStreamSynchronize(h->stream));
  mlp_release(h, h->setX);
  mlp_release(h, h->setY);
  h->setX = h->setY = nullptr;
  h->set_n = 0;
  MLP_TRY(mlp_alloc(h, &h->setX, (size_t)n * kIn));
  MLP_TRY(mlp_alloc(h, &h->setY, (size_t)n));
  Sampler S;
  S.d = *d;
  S.n_scans = n_scans;
  for (int i = 0; i < n_scans; i++) S.s[i] = scans[i];
  const long long blocks = std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_sample, dim3((unsigned)blocks), dim3(256), 0, h->stream, S, (long long)n,
                     near_fraction, (unsigned long long)seed, h->setX, h->setY);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(h->stream));
  h->set_n = n;
  return AMCA_OK;
}

amca_status amca_surrogate_train(amca_mlp* h, int32_t batch, int32_t steps, float lr, uint64_t seed,
                                 float* loss_trace) {
  if (!h || !h->set_n) return mlp_fail(AMCA_ERR_INVALID, "no training set: call amca_surrogate_sample");
  return amca_mlp_train_device(h, h->set_n, h->setX, h->setY, batch, steps, lr, seed, loss_trace);
}

amca_status amca_surrogate_eval_lattice(amca_mlp* h, const amca_surrogate_desc* d,
                                        const amca_scan* s, int32_t nxi, double xi0, double dxi,
                                        int32_t ny, double y0, double dy, int32_t nz, double z0,
                                        double dz, float* out) {
  if (!h || !d || !s || !out || nxi < 1 || ny < 1 || nz < 1)
    return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  const long long n = (long long)nxi * ny * nz;
  float* dev = nullptr;
  MLP_HIP(hipMalloc(&dev, sizeof(float) * n));
  const SurLattice g{xi0, dxi, nxi, y0, dy, ny, z0, dz, nz};
  amca_status st = surrogate_eval_lattice(h, *d, *s, g, dev);
  if (st == AMCA_OK && hipMemcpy(out, dev, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
    st = mlp_fail(AMCA_ERR_HIP, "readback");
  (void)hipFree(dev);
  return st;
}

amca_status amca_surrogate_eval(amca_mlp* h, const amca_surrogate_desc* d, const amca_scan* s,
                                int32_t n, const double* xyzt, double* Ts, double* Tfe) {
  if (!h || !d || !s || n < 1 || !xyzt || !Ts || !Tfe) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  MLP_TRY(ensure_cap(h, n));
  double *dp = nullptr, *dfe = nullptr;
  MLP_HIP(hipMalloc(&dp, sizeof(double) * 4 * n));
  MLP_HIP(hipMalloc(&dfe, sizeof(double) * n));
  amca_status st = AMCA_OK;
  std::vector<float> t(n);
  do {
    if (hipMemcpy(dp, xyzt, sizeof(double) * 4 * n, hipMemcpyHostToDevice) != hipSuccess) { st = mlp_fail(AMCA_ERR_HIP, "copy"); break; }
    hipLaunchKernelGGL(k_eval_points, dim3((n + 255) / 256), dim3(256), 0, h->stream, *d, *s, n, dp,
                       h->X, dfe);
    if ((st = forward(h, n, false)) != AMCA_OK) break;
    if (hipMemcpyAsync(t.data(), h->T, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipMemcpyAsync(Tfe, dfe, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) { st = mlp_fail(AMCA_ERR_HIP, "readback"); break; }
    for (int i = 0; i < n; i++) Ts[i] = from_target(*d, (double)t[i]);
  } while (0);
  (void)hipFree(dp);
  (void)hipFree(dfe);
  return st;
}

}  // extern "C"
