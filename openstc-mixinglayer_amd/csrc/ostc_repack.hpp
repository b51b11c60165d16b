// ostc_repack.hpp -- copy of a caller-owned device metric into the library's
// SoA planes: component c (0..8) of point p is read at
// src[p * es + c * cs + off], so SoA planes (1, sz, 0) and the reference's
// 80-byte gxl::Matrix<real,3,3,1> AoS (10, 1, 1; src/gxl_lib/Matrix.hpp:53-72,
// 1-based (r, c) at data_[3r + c - 4] after a leading slot) both read in place.
#pragma once
#include <hip/hip_runtime.h>

namespace ostc {

static __global__ void k_repack_metric(const double* __restrict__ src, long long es, long long cs, long long off,
                                       double* __restrict__ dst, long long sz) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= sz) return;
#pragma unroll
  for (int c = 0; c < 9; ++c) dst[c * sz + p] = src[p * es + c * cs + off];
}

static inline void repack_metric(const double* src, long long es, long long cs, long long off, double* dst,
                                 long long sz, hipStream_t st) {
  hipLaunchKernelGGL(k_repack_metric, dim3((unsigned)((sz + 255) / 256)), dim3(256), 0, st, src, es, cs, off, dst,
                     sz);
}

}  // namespace ostc
