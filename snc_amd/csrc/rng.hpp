// Counter-based and splittable generators shared by the host generators and
// the device kernels. Both are published methods restated from their papers:
//   SplitMix64   Steele, Lea, Flood, "Fast splittable pseudorandom number
//                generators", OOPSLA 2014 (the finaliser "mix13" constants).
//   Philox4x32-10  Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//                easy as 1, 2, 3", SC 2011 (multipliers and Weyl increments of
//                Section 4.1).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define SNC_HD __host__ __device__ __forceinline__
#else
#define SNC_HD inline
#endif

namespace snc_amd {

struct SplitMix64 {
  uint64_t s;
  explicit SNC_HD SplitMix64(uint64_t seed) : s(seed) {}
  SNC_HD uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // Uniform in [0, 1) with 53 random bits.
  SNC_HD double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Uniform in (0, 1]: safe for log().
  SNC_HD double uniform_pos() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
  // Uniform integer in [0, n).
  SNC_HD uint64_t below(uint64_t n) { return n ? (uint64_t)(uniform() * (double)n) % n : 0; }
};

struct Philox4 {
  uint32_t v[4];
};

// One Philox4x32-10 block.
SNC_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                             uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  Philox4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// Bernoulli threshold on a 32-bit uniform: u < p32 fires with probability
// p32 / 2^32, exactly representable, no float compare on either side.
SNC_HD uint32_t prob_to_u32(double p) {
  if (p <= 0.0) return 0u;
  if (p >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(p * 4294967296.0);
}

}  // namespace snc_amd
