// Philox4x32-10 counter-based random numbers (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11), the engine's replacement for the reference's threefry streams (src/lumen/_core/rng.py:43-97)
// with the same contract: every uniform is a pure function of (seed, photon id, vertex, block), in the open
// interval (0, 1) as (k + 1/2) 2^-23 for a 23-bit k, so a replayed vertex redraws exactly its numbers and a new
// block never shifts an existing one.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LH_HD __host__ __device__ __forceinline__
#else
#define LH_HD inline
#endif

struct lh_u4 {
  uint32_t x, y, z, w;
};

LH_HD uint32_t lh_mulhi(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

// ten rounds; round constants and key bumps as published (M0 = 0xD2511F53, M1 = 0xCD9E8D57, W0 = 0x9E3779B9,
// W1 = 0xBB67AE85)
LH_HD lh_u4 lh_philox(lh_u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = lh_mulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = lh_mulhi(0xCD9E8D57u, c.z);
    c = lh_u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

LH_HD float lh_uniform(uint32_t bits) { return ((float)(bits >> 9) + 0.5f) * (1.0f / 8388608.0f); }

// the four uniforms of block `block` of vertex `vertex` of photon `id`
struct lh_rng {
  uint32_t k0, k1, id_lo, id_hi;
  LH_HD void draw4(uint32_t vertex, uint32_t block, float u[4]) const {
    const lh_u4 r = lh_philox(lh_u4{id_lo, id_hi, vertex, block}, k0, k1);
    u[0] = lh_uniform(r.x);
    u[1] = lh_uniform(r.y);
    u[2] = lh_uniform(r.z);
    u[3] = lh_uniform(r.w);
  }
};

// block layout of one vertex: 0 = (path, channel, outcome, roulette), 1 = (lobe0, lobe1, lobe2, polarisation),
// 2 = (reemit0, reemit1, -, -); a facet path's draws use blocks FACET_BLOCK + (facet << 12) + attempt; the source's
// draws use vertex SOURCE_VERTEX (beyond every real vertex: the vertex guard is below 2^31)
#define LH_FACET_BLOCK 0x100u
#define LH_SOURCE_VERTEX 0xFFFFFFFFu
