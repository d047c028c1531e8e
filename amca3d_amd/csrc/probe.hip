// Stream-copy probe: the measured HBM copy rate of the device, and a kernel of
// known traffic (bytes read = bytes written = the buffer size) against which
// hardware byte counters are scaled.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/amca3d.h"

namespace {

__global__ void __launch_bounds__(256) k_copy_probe(const uint4* __restrict__ src,
                                                    uint4* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

thread_local std::string g_probe_error;

}  // namespace

extern "C" amca_status amca_copy_probe(int64_t bytes, int32_t reps, double* gbps, double* ms) {
  if (bytes < 16 || reps < 1 || !gbps) {
    g_probe_error = "bytes >= 16 and reps >= 1 required";
    return AMCA_ERR_INVALID;
  }
  const long long n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = nullptr;
  amca_status st = AMCA_ERR_HIP;
  do {
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) break;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) break;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) break;
    if (hipMemsetAsync(a, 1, n * 16, s) != hipSuccess) break;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    long long blocks = (n + 255) / 256;
    if (blocks > 16LL * cus) blocks = 16LL * cus;
    hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const uint4*)a, (uint4*)b, n);  // warm
    if (hipEventRecord(e0, s) != hipSuccess) break;
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, s,
                         (const uint4*)a, (uint4*)b, n);
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) break;
    const double per = (double)t / reps;
    *gbps = 2.0 * (double)n * 16.0 / (per * 1e-3) / 1e9;
    if (ms) *ms = per;
    st = AMCA_OK;
  } while (0);
  if (st != AMCA_OK) g_probe_error = "HIP error in the copy probe";
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return st;
}
