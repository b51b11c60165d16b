// ostc_inflow.hpp -- the inflow boundary (label 5), apply_inflow,
// src/BoundCond.cuh:235-612, shared by the air ops and the mixture kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "openstc/abi.h"

namespace ostc {

// Device layout of one inflow face's state (doubles):
// [0] inflow_type, [1] fluctuation_type, [2] delta_omega, [3] intensity,
// [4] the seed's bits, [5..10] upper rho u v w p T, [11..16] lower,
// [17..32] upper mass fractions, [33..48] lower mass fractions.
constexpr int kInflowDoubles = 50;
constexpr int kInflowUpper = 5, kInflowLower = 11, kInflowSvUpper = 17, kInflowSvLower = 33;

inline void pack_inflow(const ostc_inflow& in, double* out) {
  std::memset(out, 0, kInflowDoubles * sizeof(double));
  out[0] = in.inflow_type;
  out[1] = in.fluctuation_type;
  out[2] = in.delta_omega;
  out[3] = in.fluctuation_intensity;
  std::memcpy(&out[4], &in.seed, sizeof(double));
  for (int l = 0; l < 6; ++l) {
    out[kInflowUpper + l] = in.upper[l];
    out[kInflowLower + l] = in.lower[l];
  }
  for (int l = 0; l < OSTC_MAX_SPEC; ++l) {
    out[kInflowSvUpper + l] = in.sv_upper[l];
    out[kInflowSvLower + l] = in.sv_lower[l];
  }
}

// SplitMix64 output function (Steele, Lea & Flood, 2014).
__device__ __forceinline__ uint64_t inflow_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Two independent standard normal deviates for one face point and one
// boundary-condition pass: uniforms in (0, 1) from the top 53 bits of two
// hashes of (seed, pass, point), then the Box-Muller transform.
__device__ __forceinline__ void inflow_normals(uint64_t seed, uint64_t pass, uint64_t point, double& n1,
                                               double& n2) {
  const uint64_t key = inflow_mix(seed ^ (pass * 0xD1B54A32D192ED03ull));
  const uint64_t h1 = inflow_mix(key ^ (2 * point)), h2 = inflow_mix(key ^ (2 * point + 1));
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53, u2 = ((double)(h2 >> 11) + 0.5) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1));
  const double a = 6.283185307179586 * u2;
  n1 = r * cos(a);
  n2 = r * sin(a);
}

// Primitives (rho, u, v, w, p, T) of an inflow face point at height y, with
// the white noise of this pass; returns true when the scalars come from the
// upper stream.  `point` indexes the face point in the padded face plane.
__device__ __forceinline__ bool inflow_state(const double* F, double y, uint64_t pass, uint64_t point,
                                             double st[6]) {
  const int type = (int)F[0], fluct = (int)F[1];
  const double delta = F[2], rms = F[3];
  bool upper = true;
  double u;
  if (type == 2) {
    const double u1 = F[kInflowUpper + 1], u2 = F[kInflowLower + 1];
    u = 0.5 * (u1 + u2) + 0.5 * (u1 - u2) * tanh(2 * y / delta);
    upper = y >= 0;
  } else {
    u = F[kInflowUpper + 1];
  }
  const double* S = F + (upper ? kInflowUpper : kInflowLower);
  double v = S[2];
  const double w = S[3];
  if (fluct == 1) {
    const double vel = sqrt(u * u + v * v + w * w);
    uint64_t seed;
    memcpy(&seed, &F[4], sizeof(seed));
    if (type == 2) {
      if (y < 4 * delta && y > -4 * delta) {
        double n1, n2;
        inflow_normals(seed, pass, point, n1, n2);
        u += n1 * rms * vel;
        v += n2 * rms * vel;
      }
    } else {
      double n1, n2;
      inflow_normals(seed, pass, point, n1, n2);
      u += n1 * rms * vel;
    }
  }
  st[0] = S[0];
  st[1] = u;
  st[2] = v;
  st[3] = w;
  st[4] = S[4];
  st[5] = S[5];
  return upper;
}

// Index of a face point in its padded face plane, as the reference indexes
// its per-point random states: normal i -> (k, j), j -> (k, i), k -> (j, i).
__device__ __forceinline__ uint64_t inflow_point(int axis, int i, int j, int k, int mx, int my, int g) {
  if (axis == 0) return (uint64_t)(k + g) * (my + 2 * g) + (j + g);
  if (axis == 1) return (uint64_t)(k + g) * (mx + 2 * g) + (i + g);
  return (uint64_t)(j + g) * (mx + 2 * g) + (i + g);
}

}  // namespace ostc
