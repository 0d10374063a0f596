/*
 * probe.hip -- the memory-side atomic rate of this GPU, the ceiling of the
 * order-free raster's visibility updates (a 64-bit atomic min per covered
 * fragment, no return value).  Diagnostic only: no frame path calls it.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mop_hip.h"

namespace {

typedef __attribute__((address_space(1))) unsigned long long gu64;

/* iters atomic mins per lane; pattern 0: a random address per lane and
 * iteration (splitmix64 of the lane and step), pattern 1: the 64 lanes of a
 * wave on 64 consecutive words.  Every index is reduced modulo n. */
__global__ __launch_bounds__(256) void k_probe_atomic(unsigned long long *buf, size_t n, int iters,
                                                      int pattern) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lane = threadIdx.x & 63u, wave = gid >> 6;
  uint64_t h = (uint64_t)gid * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  for (int i = 0; i < iters; i++) {
    h ^= h >> 30;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    size_t idx = pattern == 0 ? (size_t)(h % n)
                              : (size_t)(((uint64_t)wave * (uint64_t)iters + (uint64_t)i) * 64u + lane) % n;
    __hip_atomic_fetch_min((gu64 *)(buf + idx), (unsigned long long)(h >> 8), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
}

} // namespace

extern "C" double mop_hip_probe_atomic_rate(int pattern) {
  if (pattern != 0 && pattern != 1)
    return -1.0;
  const size_t n = (size_t)64 << 20; /* 64 Mi words, 512 MiB */
  const int blocks = 256 * 32, threads = 256, iters = 64;
  unsigned long long *buf = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  double rate = -1.0;
  if (hipMalloc((void **)&buf, n * sizeof(unsigned long long)) != hipSuccess)
    return -1.0;
  if (hipStreamCreate(&s) == hipSuccess && hipEventCreate(&a) == hipSuccess &&
      hipEventCreate(&b) == hipSuccess &&
      hipMemsetAsync(buf, 0xff, n * sizeof(unsigned long long), s) == hipSuccess) {
    float best = 0.0f;
    bool ok = true;
    for (int rep = 0; rep < 3 && ok; rep++) { /* the first is a warm-up */
      hipLaunchKernelGGL(k_probe_atomic, dim3(blocks), dim3(threads), 0, s, buf, n, iters, pattern);
      ok = hipEventRecord(a, s) == hipSuccess;
      hipLaunchKernelGGL(k_probe_atomic, dim3(blocks), dim3(threads), 0, s, buf, n, iters, pattern);
      ok = ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
      float ms = 0.0f;
      ok = ok && hipEventElapsedTime(&ms, a, b) == hipSuccess && ms > 0.0f;
      if (ok && (best == 0.0f || ms < best))
        best = ms;
    }
    if (ok && best > 0.0f)
      rate = (double)blocks * threads * iters / ((double)best * 1e6);
  }
  if (a)
    (void)hipEventDestroy(a);
  if (b)
    (void)hipEventDestroy(b);
  if (s)
    (void)hipStreamDestroy(s);
  (void)hipFree(buf);
  return rate;
}
