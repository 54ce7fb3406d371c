// philox.h -- Philox4x32-10 counter-based generator (Salmon et al., SC'11),
// usable from host and device code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ima {

struct Philox4 {
  uint32_t v[4];
};

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b, uint32_t *lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  *lo = (uint32_t)p;
  return (uint32_t)(p >> 32);
}

__host__ __device__ inline Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0, lo1;
    const uint32_t hi0 = mulhi32(M0, ctr.v[0], &lo0);
    const uint32_t hi1 = mulhi32(M1, ctr.v[2], &lo1);
    Philox4 n;
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

// A stream of uniforms in [0, 1) for key (seed) and a 64-bit counter.
// The block's four words are separate members picked by selects: an
// array indexed by the run-time position would live in scratch memory
// inside device loops.
struct PhiloxStream {
  uint32_t k0, k1;
  uint64_t ctr;
  uint32_t b0, b1, b2, b3;
  int left;
  __host__ __device__ PhiloxStream(uint64_t seed, uint64_t counter)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), ctr(counter), b0(0), b1(0), b2(0),
        b3(0), left(0) {}
  __host__ __device__ uint32_t next32() {
    if (left == 0) {
      Philox4 c;
      c.v[0] = (uint32_t)ctr;
      c.v[1] = (uint32_t)(ctr >> 32);
      c.v[2] = 0x5EED0001u;
      c.v[3] = 0u;
      const Philox4 r = philox4x32_10(c, k0, k1);
      b0 = r.v[0];
      b1 = r.v[1];
      b2 = r.v[2];
      b3 = r.v[3];
      left = 4;
      ++ctr;
    }
    const int i = 4 - (left--);
    return i == 0 ? b0 : i == 1 ? b1 : i == 2 ? b2 : b3;
  }
  __host__ __device__ __forceinline__ double uniform() {
    const uint64_t hi = next32(), lo = next32();
    const uint64_t x = (hi << 32) | lo;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace ima
