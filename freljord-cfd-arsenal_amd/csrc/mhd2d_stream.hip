// mhd2d_stream.hip -- the measured HBM ceiling: a STREAM-style triad
// a = b + s c over three device arrays, timed with HIP events.  It has no
// counterpart in the reference; the benchmark reports the step's kernels
// against it beside the 8 TB/s specification.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "mhd2d.h"

namespace {

constexpr int TB = 256;

// U 16-byte elements per lane, TB lanes apart: every wave reads and writes
// whole 1 KiB lines, and U loads of each operand are in flight per lane
template <int U>
__global__ __launch_bounds__(TB) void k_triad(double2* __restrict__ a, const double2* __restrict__ b,
                                              const double2* __restrict__ c, double s, long long n2) {
  const long long base = (long long)blockIdx.x * (TB * U) + threadIdx.x;
  double2 x[U], y[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const long long k = base + (long long)u * TB;
    if (k < n2) {
      x[u] = b[k];
      y[u] = c[k];
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const long long k = base + (long long)u * TB;
    if (k < n2) a[k] = make_double2(x[u].x + s * y[u].x, x[u].y + s * y[u].y);
  }
}

template <int U>
void launch(double* a, const double* b, const double* c, long long n, hipStream_t st) {
  const long long n2 = n / 2;
  const long long blocks = (n2 + TB * U - 1) / (TB * U);
  hipLaunchKernelGGL(k_triad<U>, dim3((unsigned)blocks), dim3(TB), 0, st, (double2*)a, (const double2*)b,
                     (const double2*)c, 1.5, n2);
}

}  // namespace

extern "C" int mhd2d_stream_triad(int64_t n, int32_t unroll, int32_t reps, double* gbs, double* ms_per_launch) {
  if (n <= 0 || (n & 1) || reps <= 0 || !gbs) return -1;
  if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8) return -1;
  if (n / 2 / (TB * unroll) >= (1LL << 31)) return -1;
  double *a = nullptr, *b = nullptr, *c = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  int rc = -2;
  const size_t bytes = (size_t)n * sizeof(double);
  auto go = [&](void) {
    switch (unroll) {
      case 1: launch<1>(a, b, c, n, st); break;
      case 2: launch<2>(a, b, c, n, st); break;
      case 4: launch<4>(a, b, c, n, st); break;
      default: launch<8>(a, b, c, n, st); break;
    }
  };
  if (hipStreamCreate(&st) != hipSuccess) return -2;
  if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess && hipMalloc(&c, bytes) == hipSuccess &&
      hipMemsetAsync(b, 0, bytes, st) == hipSuccess && hipMemsetAsync(c, 0, bytes, st) == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    go();   // warm-up launch
    const bool rec = hipEventRecord(e0, st) == hipSuccess;
    for (int r = 0; r < reps; r++) go();
    float ms = 0.f;
    if (rec && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipGetLastError() == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f) {
      const double per = (double)ms / reps;
      *gbs = 3.0 * (double)bytes / (per * 1e-3) / 1e9;
      if (ms_per_launch) *ms_per_launch = per;
      rc = 0;
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = -2;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (c) (void)hipFree(c);
  (void)hipStreamDestroy(st);
  return rc;
}
