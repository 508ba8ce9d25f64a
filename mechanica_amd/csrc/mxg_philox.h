// Philox4x32-10 counter-based random numbers (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Usable from host
// code and from HIP device code.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MXG_HD __host__ __device__ __forceinline__
#else
#define MXG_HD inline
#endif

namespace mxg {

struct u32x4 {
  uint32_t v[4];
};

MXG_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Ten rounds of Philox4x32 on counter ctr with key (k0, k1).
MXG_HD u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, ctr.v[0], hi0, lo0);
    philox_mulhilo(M1, ctr.v[2], hi1, lo1);
    u32x4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// Uniform float in (0, 1]: the top 24 bits of u, offset by half an ulp.
MXG_HD float u01_open_closed(uint32_t u) {
  return ((float)(u >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// Uniform float in [0, 1).
MXG_HD float u01_closed_open(uint32_t u) {
  return (float)(u >> 8) * (1.0f / 16777216.0f);
}

// 32-bit integer hash (the output permutation of O'Neill's PCG, "RXS M XS"
// on one LCG step).  Chained, it gives the per-pair DPD noise: one draw per
// pair and step at a fraction of a Philox call.
MXG_HD uint32_t mix32(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

// Key of one time step for the pair noise.
MXG_HD uint32_t step_key(uint32_t seed, uint32_t time) { return mix32(time ^ mix32(seed ^ 0x64706472u)); }

// Noise of the pair (a, b), a < b, at the step whose key is k: uniform on
// [-sqrt(3), sqrt(3)), zero mean and unit variance.
MXG_HD float pair_uniform(uint32_t a, uint32_t b, uint32_t k) {
  const uint32_t h = mix32(a ^ mix32(b ^ k));
  return 1.7320508f * (2.0f * ((float)(h >> 8) * (1.0f / 16777216.0f)) - 1.0f);
}

// Stream identifiers keep the different uses of the generator independent.
enum : uint32_t {
  kStreamOverlap = 0x6F766C70u,   // direction for pairs closer than a
  kStreamDPD = 0x64706472u,       // DPD pair noise
  kStreamSingle = 0x73696E67u,    // single-body random force
  kStreamAngle = 0x616E676Cu,     // plane of a right or straight angle
};

}  // namespace mxg
