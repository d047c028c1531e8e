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
#include <random>
#include <string>
#include <vector>

#include "../../include/amca3d.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace mlpk {

constexpr int kIn = 5, kHid = 128, kLayers = 5;
constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kPad = 8;  // bf16 elements of padding per LDS row
constexpr int kThreads = 256;

// parameter vector layout: W1[5][128] b1[128] W2[128][128] b2 W3 b3 W4 b4 W5[128][1] b5[1]
constexpr int kFanIn[kLayers] = {kIn, kHid, kHid, kHid, kHid};
constexpr int kFanOut[kLayers] = {kHid, kHid, kHid, kHid, 1};

struct Layout {
  int w[kLayers], b[kLayers];
  int total;
  Layout() {
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
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
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

  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    // stage A (128 x 32) and B (32 x 128): 4096 elements each, 16 per thread
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
      float v = 0.0f;
      if (gm < M && gk < kend) v = to_f(TrA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]);
      As[mm][kk] = (__bf16)v;
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
      float v = 0.0f;
      if (gn < N && gk < kend) v = to_f(TrB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn]);
      Bs[nn][kk] = (__bf16)v;
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
  // D layout: element i of lane l is row (i/4)*8 + (l/32)*4 + i%4, column l%32
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
        e.C[o] = v * dsilu(e.Z[o]);
      } else {
        atomicAdd(&e.C[o], v);
      }
    }
  }
}

// column sums of an M x N FP32 matrix (bias gradients), atomically added
__global__ void k_colsum(const float* __restrict__ X, int M, int N, int rows_per_block,
                         float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int r0 = blockIdx.y * rows_per_block, r1 = min(M, r0 + rows_per_block);
  float s = 0.0f;
  for (int r = r0; r < r1; r++) s += X[(size_t)r * N + n];
  atomicAdd(&out[n], s);
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

using namespace mlpk;

struct amca_mlp {
  hipStream_t stream = nullptr;
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
  std::vector<void*> allocs;
  ~amca_mlp() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

thread_local std::string g_mlp_error;

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

// forward of n rows already in h->X; keeps Z and activations for a backward pass
amca_status forward(amca_mlp* h, int n) {
  const Layout& L = h->lay;
  hipStream_t s = h->stream;
  for (int l = 0; l < 4; l++) {
    EpiArgs e{nullptr, kHid, h->params + L.b[l], h->Z[l], h->Aact[l]};
    if (l == 0)
      MLP_HIP((gemm<float, float, false, false, kEpiBiasSilu>(s, h->X, kIn, h->params + L.w[0], kHid,
                                                               n, kHid, kIn, e)));
    else
      MLP_HIP((gemm<__bf16, float, false, false, kEpiBiasSilu>(s, h->Aact[l - 1], kHid,
                                                                h->params + L.w[l], kHid, n, kHid,
                                                                kHid, e)));
  }
  EpiArgs e{h->T, 1, h->params + L.b[4], nullptr, nullptr};
  MLP_HIP((gemm<__bf16, float, false, false, kEpiStore>(s, h->Aact[3], kHid, h->params + L.w[4], 1, n,
                                                         1, kHid, e)));
  return AMCA_OK;
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
  const int ksplit = std::max(1, std::min(64, n / 512));
  const int rpb = 256;
  // layer 5: dW5 = A4^T dT, db5 = sum dT, dZ4 = (dT W5^T) * silu'(Z4)
  {
    EpiArgs e{h->grads + L.w[4], 1, nullptr, nullptr, nullptr};
    MLP_HIP((gemm<__bf16, float, true, false, kEpiAtomic>(s, h->Aact[3], kHid, h->dT, 1, kHid, 1, n,
                                                           e, ksplit)));
    hipLaunchKernelGGL(k_colsum, dim3(1, (n + rpb - 1) / rpb), dim3(64), 0, s, h->dT, n, 1, rpb,
                       h->grads + L.b[4]);
    MLP_HIP(hipGetLastError());
    EpiArgs d{h->dZ[0], kHid, nullptr, h->Z[3], nullptr};
    MLP_HIP((gemm<float, float, false, true, kEpiDSilu>(s, h->dT, 1, h->params + L.w[4], 1, n, kHid,
                                                         1, d)));
  }
  int cur = 0;
  for (int l = 3; l >= 0; l--) {
    float* dz = h->dZ[cur];
    // dW_l = A_{l-1}^T dZ_l (A_{-1} = X), db_l = sum dZ_l
    EpiArgs e{h->grads + L.w[l], kHid, nullptr, nullptr, nullptr};
    if (l == 0)
      MLP_HIP((gemm<float, float, true, false, kEpiAtomic>(s, h->X, kIn, dz, kHid, kIn, kHid, n, e,
                                                            ksplit)));
    else
      MLP_HIP((gemm<__bf16, float, true, false, kEpiAtomic>(s, h->Aact[l - 1], kHid, dz, kHid, kHid,
                                                             kHid, n, e, ksplit)));
    hipLaunchKernelGGL(k_colsum, dim3((kHid + 63) / 64, (n + rpb - 1) / rpb), dim3(64), 0, s, dz, n,
                       kHid, rpb, h->grads + L.b[l]);
    MLP_HIP(hipGetLastError());
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
amca_status amca_mlp_forward_device(amca_mlp* h, int64_t n, const float* dX, float* dT) {
  if (!h || n < 1 || !dX || !dT) return mlp_fail(AMCA_ERR_INVALID, "bad arguments");
  const int chunk = 1 << 20;
  MLP_TRY(ensure_cap(h, (int)std::min<int64_t>(n, chunk)));
  for (int64_t r = 0; r < n; r += chunk) {
    const int m = (int)std::min<int64_t>(chunk, n - r);
    MLP_HIP(hipMemcpyAsync(h->X, dX + r * kIn, sizeof(float) * kIn * m, hipMemcpyDeviceToDevice,
                           h->stream));
    MLP_TRY(forward(h, m));
    MLP_HIP(hipMemcpyAsync(dT + r, h->T, sizeof(float) * m, hipMemcpyDeviceToDevice, h->stream));
  }
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
