// ostc_oracle_inflow.h -- test infrastructure only (the CPU oracle).
// White noise of the inflow boundary: SplitMix64 (Steele, Lea & Flood 2014)
// hashes of (seed, pass, point) give two uniforms in (0, 1) from their top 53
// bits, the Box-Muller transform two standard normal deviates.  The
// reference draws its deviates from cuRAND states seeded by the wall clock
// (src/DigitalFilter.cu:236, src/BoundCond.cu:1385-1391); no run of it can be
// reproduced, so the product and the oracle share this generator instead.
#pragma once

#include <cmath>
#include <cstdint>

namespace orc_inflow_noise {

inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline void normal_pair(uint64_t seed, uint64_t pass, uint64_t point, double* n1, double* n2) {
  const uint64_t key = splitmix(seed ^ (pass * 0xD1B54A32D192ED03ull));
  const uint64_t h1 = splitmix(key ^ (2 * point)), h2 = splitmix(key ^ (2 * point + 1));
  const double u1 = (static_cast<double>(h1 >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = (static_cast<double>(h2 >> 11) + 0.5) * 0x1.0p-53;
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double a = 6.283185307179586 * u2;
  *n1 = r * std::cos(a);
  *n2 = r * std::sin(a);
}

}  // namespace orc_inflow_noise
