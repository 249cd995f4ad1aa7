// log Gamma(x) for x > 0 with no library call but log: the device objective's
// lgamma terms (the count-histogram tail, lgamma(alpha), lgamma(cell)).
#pragma once

#include <cmath>

#ifndef NH_HD
#ifdef __HIPCC__
#define NH_HD __host__ __device__
#else
#define NH_HD
#endif
#endif

namespace nh {

// x is shifted to z = x + n >= 10 (lgamma(x) = lgamma(z) - log(x (x+1) ...
// (x+n-1)), one log of the product: at most 10 factors, no overflow for
// x > 1e-300), then Stirling's series (z - 1/2) log z - z + log(2 pi) / 2 +
// sum B_2j / (2j (2j-1) z^(2j-1)) to j = 8 (truncation < 2e-18 at z = 10).
NH_HD inline double lgamma_pos(double x) {
  double z = x, prod = 1.0;
  while (z < 10.0) {
    prod *= z;
    z += 1.0;
  }
  const double zi = 1.0 / z, z2 = zi * zi;
  double s = -3617.0 / 122400.0;  // B16 / (16 * 15)
  s = s * z2 + 1.0 / 156.0;       // B14 / (14 * 13)
  s = s * z2 - 691.0 / 360360.0;  // B12 / (12 * 11), then B10 ... B2
  s = s * z2 + 1.0 / 1188.0;
  s = s * z2 - 1.0 / 1680.0;
  s = s * z2 + 1.0 / 1260.0;
  s = s * z2 - 1.0 / 360.0;
  s = s * z2 + 1.0 / 12.0;
  return (z - 0.5) * std::log(z) - z + 0.91893853320467274178 + s * zi - std::log(prod);
}

}  // namespace nh
