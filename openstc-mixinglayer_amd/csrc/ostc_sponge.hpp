// ostc_sponge.hpp -- host side of the sponge layer (src/SpongeLayer.cu:68-172):
// sigma at one node from its coordinates.
#pragma once

#include "openstc/abi.h"

namespace ostc {

// sponge_layer_source's sigma: the sides in the order x-, x+, y-, y+, z-, z+,
// each adding sigma_d S(xi) where the coordinate lies beyond its start,
// xi = (s - start)/(end - start) clipped at 1, S of sponge_function 0
// (6 xi^5 - 15 xi^4 + 10 xi^3, written as the reference writes it).
inline double sponge_sigma_at(const ostc_sponge& sp, double x, double y, double z) {
  double sigma = 0;
  const double s3[3] = {x, y, z};
  for (int d = 0; d < 6; ++d) {
    if (!sp.enabled[d]) continue;
    const double v = s3[d / 2], x1 = sp.start[d];
    if ((d % 2 == 0) ? !(v < x1) : !(v > x1)) continue;
    double xi = (v - x1) / (sp.end[d] - x1);
    if (xi > 1) xi = 1;
    const double xi3 = xi * xi * xi;
    const double S = 6 * xi3 * xi * xi - 15 * xi3 * xi + 10 * xi3;
    sigma += sp.sigma[d] * S;
  }
  return sigma;
}

}  // namespace ostc
