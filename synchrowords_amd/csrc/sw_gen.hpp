// Seeded generator of synthetic automata (the same stream as
// synchrowords_amd/gen.py): splitmix64 seeding, xoshiro256**, Lemire's
// unbiased bounded draws; delta drawn q-major, a-minor.
#pragma once
#include <stdint.h>

namespace sw {

struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& st) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(st);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  uint64_t bounded(uint64_t n) {
    __uint128_t m = (__uint128_t)next() * n;
    uint64_t low = uint64_t(m);
    if (low < n) {
      const uint64_t t = (0 - n) % n;
      while (low < t) {
        m = (__uint128_t)next() * n;
        low = uint64_t(m);
      }
    }
    return uint64_t(m >> 64);
  }
};

inline void random_dfa(uint32_t n, uint32_t k, uint64_t seed, uint32_t* out) {
  Xoshiro r(seed);
  for (uint64_t i = 0; i < uint64_t(n) * k; ++i) out[i] = uint32_t(r.bounded(n));
}

}  // namespace sw
