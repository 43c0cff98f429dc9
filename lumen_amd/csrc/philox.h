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

// ten rounds; round constants and key bumps as published (M0 = 0xD2511F53, M1 = 0xCD9E8D57, W0 = 0x9E3779B9,
// W1 = 0xBB67AE85)
LH_HD lh_u4 lh_philox(lh_u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // each product's low and high words from one 32 x 32 -> 64 multiply (one v_mad_u64_u32 on CDNA)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
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
// 2 = (reemit0, reemit1, -, -); a facet path's draws use blocks FACET_BLOCK + (facet << 12) + attempt, and a GLISUR
// facet's own choices (the cover's shell, the Fresnel choice) GLISUR_BLOCK + facet; the source's draws use vertex
// SOURCE_VERTEX (beyond every real vertex: the vertex guard is below 2^31)
#define LH_FACET_BLOCK 0x100u
#define LH_GLISUR_BLOCK 0x400000u
#define LH_SOURCE_VERTEX 0xFFFFFFFFu
