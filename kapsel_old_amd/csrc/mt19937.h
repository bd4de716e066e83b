// MT19937 (mt19937ar.c:65-171: init_genrand, genrand_int32, genrand_real3)
// and the polar Gaussian pair of Gauss2 (fluct.h:48-60), on a caller's
// state (SPM_MT_N words and the index), host code only.  The library's one
// generator: the thermostat stream (spm_api.cpp) and the particle
// initialisation's walks (udf.cpp) both draw from it, so a run continues
// the stream the initialisation leaves, as the reference's does.  The
// published algorithm with its constants; the Gaussian uses the host libm's
// log and sqrt, as the reference's stream does.
#ifndef SPM_MT19937_H
#define SPM_MT19937_H

#include <cmath>
#include <cstdint>

namespace spm {

constexpr int kMtN = 624, kMtM = 397;

inline void mt_init(uint32_t *mt, int &mti, uint32_t seed) {
  mt[0] = seed;
  for (mti = 1; mti < kMtN; ++mti) mt[mti] = 1812433253U * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (uint32_t)mti;
}

// an index of kMtN + 1 (never seeded) seeds with 5489 first, as genrand_int32
inline uint32_t mt_next(uint32_t *mt, int &mti) {
  static const uint32_t mag01[2] = {0x0U, 0x9908b0dfU};
  uint32_t y;
  if (mti >= kMtN) {
    if (mti == kMtN + 1) mt_init(mt, mti, 5489U);
    int kk;
    for (kk = 0; kk < kMtN - kMtM; kk++) {
      y = (mt[kk] & 0x80000000U) | (mt[kk + 1] & 0x7fffffffU);
      mt[kk] = mt[kk + kMtM] ^ (y >> 1) ^ mag01[y & 1U];
    }
    for (; kk < kMtN - 1; kk++) {
      y = (mt[kk] & 0x80000000U) | (mt[kk + 1] & 0x7fffffffU);
      mt[kk] = mt[kk + (kMtM - kMtN)] ^ (y >> 1) ^ mag01[y & 1U];
    }
    y = (mt[kMtN - 1] & 0x80000000U) | (mt[0] & 0x7fffffffU);
    mt[kMtN - 1] = mt[kMtM - 1] ^ (y >> 1) ^ mag01[y & 1U];
    mti = 0;
  }
  y = mt[mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680U;
  y ^= (y << 15) & 0xefc60000U;
  y ^= (y >> 18);
  return y;
}

// genrand_real3: (0, 1)
inline double mt_real3(uint32_t *mt, int &mti) { return (((double)mt_next(mt, mti)) + 0.5) * (1.0 / 4294967296.0); }

// Marsaglia's polar method: two normals from pairs in the unit disc
inline void mt_gauss2(uint32_t *mt, int &mti, double *out) {
  double x1, x2, rsq;
  do {
    x1 = 2. * mt_real3(mt, mti) - 1.;
    x2 = 2. * mt_real3(mt, mti) - 1.;
    rsq = x1 * x1 + x2 * x2;
  } while (rsq >= 1.0 || rsq == 0.0);
  const double factor = std::sqrt(-2. * std::log(rsq) / rsq);
  out[0] = factor * x1;
  out[1] = factor * x2;
}

}  // namespace spm

#endif  // SPM_MT19937_H
