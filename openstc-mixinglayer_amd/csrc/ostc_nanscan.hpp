// ostc_nanscan.hpp -- where a NaN sits, for the error message of the
// residual entry points (check_nan, src/Residual.cu:35-49, which prints the
// nodes whose primitives hold a NaN when the residual is NaN).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>

namespace ostc {

static __global__ void k_nan_init(unsigned long long* first, unsigned long long n0) { *first = n0; }

// The smallest interior index t = (k my + j) mx + i whose six primitives
// (bv planes of stride sz, g ghost layers) hold a NaN, by a 64-bit atomicMin.
static __global__ void k_first_nan(const double* __restrict__ bv, long long sz, int mx, int my, int mz, int g,
                                   unsigned long long* first) {
  const long long n0 = (long long)mx * my * mz;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n0; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % mx), j = (int)((t / mx) % my), k = (int)(t / ((long long)mx * my));
    const long long p = ((long long)(k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g);
    bool bad = false;
#pragma unroll
    for (int l = 0; l < 6; ++l) bad = bad || isnan(bv[p + l * sz]);
    if (bad) atomicMin(first, (unsigned long long)t);
  }
}

// Host side: the first NaN node of one block as "node (i, j, k), bv = {...}",
// or an empty string; `word` is one device word of scratch.  Synchronous.
static inline std::string first_nan_text(const double* bv, long long sz, int mx, int my, int mz, int g,
                                         unsigned long long* word, hipStream_t st) {
  const unsigned long long n0 = (unsigned long long)mx * my * mz;
  hipLaunchKernelGGL(k_nan_init, dim3(1), dim3(1), 0, st, word, n0);
  hipLaunchKernelGGL(k_first_nan, dim3(1024), dim3(256), 0, st, bv, sz, mx, my, mz, g, word);
  unsigned long long t = n0;
  if (hipMemcpyAsync(&t, word, sizeof(t), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess || t >= n0)
    return std::string();
  const int i = (int)(t % mx), j = (int)((t / mx) % my), k = (int)(t / ((unsigned long long)mx * my));
  const long long p = ((long long)(k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g);
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int l = 0; l < 6; ++l) (void)hipMemcpyAsync(&v[l], bv + p + l * sz, sizeof(double), hipMemcpyDeviceToHost, st);
  (void)hipStreamSynchronize(st);
  char buf[256];
  snprintf(buf, sizeof(buf), "node (%d, %d, %d), bv = {%g, %g, %g, %g, %g, %g}", i, j, k, v[0], v[1], v[2], v[3], v[4],
           v[5]);
  return std::string(buf);
}

}  // namespace ostc
