// The navigator on the device: the next visible crossing along a ray, and the volume holding a point.
//
// Restates the reference navigator's query for volumes that lie inside their mother (no overlaps, no meshes, no
// invisible crossings): the nearer of the photon's own volume's exit and the first daughter entry, a daughter
// winning a tie (src/lumen/geometry/navigate.py:663-693 `query`, :430-463 `_node`, src/lumen/rules/navigation.py
// `entry_wins`).  Daughters are found by a stackless walk of each frame group's BVH (depth-first nodes with skip
// links, src/lumen/geometry/bvh.py:1-10); a solid is its CSG leaves' intervals along the ray combined by the postfix
// program, scanned endpoint by endpoint for the first change of membership (src/lumen/geometry/csg.py:417-449
// `first_change`).  Quadric roots are taken at the closest approach (the vertex where the roots are near it), so
// float32 never cancels the large terms (src/lumen/geometry/leaf_kernels.py:81-109 re-origination).
//
// Surface identity: the face a photon sits on (its `srf` code: volume, leaf, face) is not met again within K_ON
// bands of the start; the band is BAND_ULPS float32 ulps of the coordinates plus a floor
// (src/lumen/rules/tolerances.py:76-83), the coordinates' scale the larger of the start's and that of the ray whose
// crossing put the photon on the face (the crossing point's rounding is that ray's).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

constexpr float HUGEF = 1e30f;
constexpr float BIGF = 1e29f;
constexpr float EPS32 = 1.1920928955078125e-7f;
constexpr float BAND_ULPS = 8.0f;
constexpr float BAND_FLOOR = 1e-9f;
constexpr float K_ON = 4.0f;
constexpr int MAX_CSG_LEAVES = 8;
constexpr int WALK_PENDING = 3;  // leaves a lane collects on its walk before the wavefront tests their items
constexpr int MAX_GROUPS_WALK = 8;

// the launch's dynamic LDS: a kernel may stage the BVH nodes there (Geo::lds_nodes) after its own cells
extern __shared__ double lds_cells[];

struct Geo {
  int nvol;
  const float4 *__restrict__ xf;    // 3 per volume: rows of R^T with the offset -R^T t (world -> local)
  const int4 *__restrict__ vinfo;   // per navigator id: touch flag (a sibling's box comes near), parent, cls, sensor
  const int2 *__restrict__ vgrp;    // per navigator id: frame groups [x, y)
  const float4 *__restrict__ gxf;   // 3 per group: rows of gR^T (w: 1 if identity)
  const int2 *__restrict__ grange;  // per group: BVH nodes [x, y)
  const float4 *__restrict__ node;  // 2 per node: (lo, skip), (hi, start * 8 + count)
  const int *__restrict__ items;
  const float4 *__restrict__ leaf;  // 4 per leaf: (q, c), (A00 A01 A02 A11), (A12 A22 b0 b1), (b2, typ, p0, np)
  const float4 *__restrict__ plane; // (n, h) relative to the leaf's q
  const int4 *__restrict__ kinds;   // l0, nl, p0, pl
  const int4 *__restrict__ vkind;   // per navigator id: its kind's (l0, nl, p0, pl), read beside its transform
  const float4 *__restrict__ ixf;   // per BVH item (in items' order): its volume's 3 transform rows
  const int4 *__restrict__ ikind;   // per BVH item: its volume's kind record
  const int *__restrict__ prog;     // op << 24 | arg
  const float4 *__restrict__ vcore;  // 3 per navigator id: its daughter-free core (core_interval)
  // the reach bound (reach_limit): 2 per volume its world box padded by twice the relocation step at the geometry's
  // scale; 2 per node of a BVH over the sensor volumes' boxes (padded alike), (lo, skip), (hi, 1 at a leaf: one box)
  const float4 *__restrict__ rbox;
  const float4 *__restrict__ rnode;
  int n_rnode;  // the sensor-box BVH's nodes (0: no sensor); -1: no reach bound
  float reach;
  int n_nodes;
  int lds_nodes;     // 1: the kernel staged all BVH nodes in LDS at float4 offset lds_node_off of lds_cells
  int lds_node_off;
};

// a BVH node's two float4s, (lo, skip) and (hi, start * 8 + count): from LDS where the kernel staged them (LDS: a
// compile-time choice, so each walk compiles to LDS reads or to global reads, never to a generic load that decides
// per access).  In LDS the two halves are two arrays (every lo, then every hi): a node's halves 16 bytes apart in
// memory would put the lanes' reads of different nodes on 4 bank groups of the 32 banks, 16-byte strides on 8
template <bool LDS>
__device__ __forceinline__ void nav_node(const Geo &g, int ni, float4 &lo, float4 &hi) {
  if (LDS) {
    const float4 *s = reinterpret_cast<const float4 *>(lds_cells) + g.lds_node_off;
    lo = s[ni];
    hi = s[g.n_nodes + ni];
  } else {
    lo = g.node[2 * ni];
    hi = g.node[2 * ni + 1];
  }
}

// copy the BVH nodes into the block's LDS, the lo halves then the hi halves (every thread of the block calls it; a
// barrier follows)
__device__ inline void stage_nodes(const Geo &g) {
  if (!g.lds_nodes) return;
  float4 *dst = reinterpret_cast<float4 *>(lds_cells) + g.lds_node_off;
  for (int i = threadIdx.x; i < 2 * g.n_nodes; i += blockDim.x) {
    const int ni = i >> 1;
    dst[(i & 1) ? g.n_nodes + ni : ni] = g.node[i];
  }
}

__device__ __forceinline__ float3 mk3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float s, float3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 fma3(float s, float3 a, float3 b) {
  return mk3(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z));
}
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 unit3(float3 a) {
  const float s = dot3(a, a);
  return s > 0.0f ? mul3(rsqrtf(s), a) : a;
}
__device__ __forceinline__ float maxabs3(float3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }

// the band at a coordinate scale (src/lumen/rules/navigation.py:129-134)
__device__ __forceinline__ float band_at(float scale) { return BAND_ULPS * EPS32 * scale + BAND_FLOOR; }

__device__ __forceinline__ int srf_code(int vol, int leaf_local, int face) {
  return (vol << 10) | (leaf_local << 5) | face;
}

// world -> a frame given by its three rows (R^T | -R^T t)
__device__ __forceinline__ void rows_to_local(const float4 *__restrict__ rows, float3 x, float3 d, float3 &xl,
                                              float3 &dl) {
  const float4 r0 = rows[0], r1 = rows[1], r2 = rows[2];
  xl = mk3(fmaf(r0.x, x.x, fmaf(r0.y, x.y, fmaf(r0.z, x.z, r0.w))),
           fmaf(r1.x, x.x, fmaf(r1.y, x.y, fmaf(r1.z, x.z, r1.w))),
           fmaf(r2.x, x.x, fmaf(r2.y, x.y, fmaf(r2.z, x.z, r2.w))));
  dl = mk3(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
           r2.x * d.x + r2.y * d.y + r2.z * d.z);
}

// world -> a volume's frame
__device__ __forceinline__ void to_local(const Geo &g, int v, float3 x, float3 d, float3 &xl, float3 &dl) {
  const float4 r0 = g.xf[3 * v], r1 = g.xf[3 * v + 1], r2 = g.xf[3 * v + 2];
  xl = mk3(fmaf(r0.x, x.x, fmaf(r0.y, x.y, fmaf(r0.z, x.z, r0.w))),
           fmaf(r1.x, x.x, fmaf(r1.y, x.y, fmaf(r1.z, x.z, r1.w))),
           fmaf(r2.x, x.x, fmaf(r2.y, x.y, fmaf(r2.z, x.z, r2.w))));
  dl = mk3(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
           r2.x * d.x + r2.y * d.y + r2.z * d.z);
}

__device__ __forceinline__ float3 rows_normal_to_world(const float4 *__restrict__ rows, float3 n) {
  const float4 r0 = rows[0], r1 = rows[1], r2 = rows[2];
  return mk3(r0.x * n.x + r1.x * n.y + r2.x * n.z, r0.y * n.x + r1.y * n.y + r2.y * n.z,
             r0.z * n.x + r1.z * n.y + r2.z * n.z);
}

__device__ __forceinline__ float3 normal_to_world(const Geo &g, int v, float3 n) {
  const float4 r0 = g.xf[3 * v], r1 = g.xf[3 * v + 1], r2 = g.xf[3 * v + 2];
  return mk3(r0.x * n.x + r1.x * n.y + r2.x * n.z, r0.y * n.x + r1.y * n.y + r2.y * n.z,
             r0.z * n.x + r1.z * n.y + r2.z * n.z);
}

struct LeafIv {
  float lo[2], hi[2];
  int flo[2], fhi[2];
};

// A torus leaf (type 2; swept radius R, tube radius r, unit axis w through its centre): along y + t d, |d| = 1, with
// s = t - tc measured from the point nearest the centre, g(s) = s^2 + |p0|^2 + R^2 - r^2 - 2 R rho(s), rho(s)^2 =
// A s^2 + 2 B s + C the squared distance from the axis; g <= 0 inside.  g'' = 2 - 2 R (A C - B^2) / rho^3 changes
// sign at most twice (where rho = cbrt(R (A C - B^2)): closed form), so the bounding sphere's chord splits into at
// most three convex or concave pieces, each with one extremum at most (bisection on g'), giving at most six
// monotone pieces with one root each (bisection on the membership): up to four roots, two intervals, the chord's
// start outside.  The reference's root isolation (geometry/leaf_kernels.py:212-248 torus_set), in FP32.
constexpr int TORUS_BISECT = 30;

struct TorusLine {
  float R, r2, A, B, C, k;  // k = |p0|^2 + R^2 - r^2

  __device__ __forceinline__ float rho(float s) const { return sqrtf(fmaxf(fmaf(fmaf(A, s, 2.0f * B), s, C), 0.0f)); }
  __device__ __forceinline__ float g(float s) const { return fmaf(s, s, k) - 2.0f * R * rho(s); }
  __device__ __forceinline__ float dg(float s) const {
    const float rh = fmaxf(rho(s), 1e-30f);
    return 2.0f * s - 2.0f * R * fmaf(A, s, B) / rh;
  }
};

// the roots of g on [xl, xr] where g is monotone (at most one: where the membership changes)
__device__ __forceinline__ bool torus_root(const TorusLine &L, float xl, float xr, float &root) {
  const bool il = L.g(xl) <= 0.0f, ir = L.g(xr) <= 0.0f;
  if (il == ir || !(xr > xl)) return false;
  float lo = xl, hi = xr;
  for (int k = 0; k < TORUS_BISECT; ++k) {
    const float mid = 0.5f * (lo + hi);
    if ((L.g(mid) <= 0.0f) == il) lo = mid; else hi = mid;
  }
  root = 0.5f * (lo + hi);
  return true;
}

// the extremum of g on [xl, xr] where g' is monotone, or -HUGEF when g' keeps its sign there; g' read just inside
// the ends (at a knot on the axis, rho = 0, g' jumps and the piece's own one-sided value is wanted)
__device__ __forceinline__ float torus_extremum(const TorusLine &L, float xl, float xr) {
  const float in = 1e-6f * (xr - xl);
  const float gl = L.dg(xl + in), gr = L.dg(xr - in);
  if (!(xr > xl) || (gl < 0.0f) == (gr < 0.0f)) return -HUGEF;
  float lo = xl, hi = xr;
  const bool nl = gl < 0.0f;
  for (int k = 0; k < TORUS_BISECT; ++k) {
    const float mid = 0.5f * (lo + hi);
    if ((L.dg(mid) < 0.0f) == nl) lo = mid; else hi = mid;
  }
  return 0.5f * (lo + hi);
}

// the torus's intervals [a0, b0], [a1, b1] along y + t d (empty: a > b)
__device__ inline void torus_intervals(float R, float r, float3 w, float3 y, float3 d, float &a0, float &b0,
                                       float &a1, float &b1) {
  a0 = a1 = HUGEF;
  b0 = b1 = -HUGEF;
  const float tc = -dot3(y, d);
  const float3 p0 = fma3(tc, d, y);
  const float pp = dot3(p0, p0), rb = R + r;
  if (pp >= rb * rb) return;
  const float sb = sqrtf(rb * rb - pp) * (1.0f + 1e-6f) + 1e-6f * rb;  // ends strictly outside the torus
  const float pz = dot3(p0, w), dz = dot3(d, w);
  const float3 pq = fma3(-pz, w, p0), dq = fma3(-dz, w, d);
  TorusLine L{R, r * r, dot3(dq, dq), dot3(pq, dq), dot3(pq, pq), pp + R * R - r * r};
  // the inflection points, clamped into the chord [-sb, sb]
  float k1 = -sb, k2 = -sb;
  const float det = fmaxf(L.A * L.C - L.B * L.B, 0.0f);
  if (L.A > 1e-12f) {
    const float ri = cbrtf(R * det);
    const float disc = L.B * L.B - L.A * (L.C - ri * ri);
    if (disc >= 0.0f) {
      const float sq = sqrtf(disc);
      k1 = fminf(fmaxf((-L.B - sq) / L.A, -sb), sb);
      k2 = fminf(fmaxf((-L.B + sq) / L.A, -sb), sb);
    }
  }
  float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;  // the roots in order (registers: no indexed array)
  int nr = 0;
  auto push = [&](float v) {
    r0 = nr == 0 ? v : r0;
    r1 = nr == 1 ? v : r1;
    r2 = nr == 2 ? v : r2;
    r3 = nr == 3 ? v : r3;
    nr = min(nr + 1, 4);
  };
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float xl = i == 0 ? -sb : i == 1 ? k1 : k2, xr = i == 0 ? k1 : i == 1 ? k2 : sb;
    const float e = torus_extremum(L, xl, xr);
    float rt;
    if (e > -HUGEF) {
      if (torus_root(L, xl, e, rt)) push(rt);
      if (torus_root(L, e, xr, rt)) push(rt);
    } else if (torus_root(L, xl, xr, rt)) {
      push(rt);
    }
  }
  if (nr >= 2) { a0 = tc + r0; b0 = tc + r1; }
  if (nr >= 4) { a1 = tc + r2; b1 = tc + r3; }
}

// a leaf's (at most two) intervals [lo, hi) along o + t d in its kind's frame; face 0 the quadric, k + 1 plane k.
// TOR: the scene has torus leaves (a kernel feature: the root isolation costs registers every other kernel keeps)
template <bool TOR>
__device__ inline void leaf_intervals(const Geo &g, int li, float3 o, float3 d, LeafIv &iv) {
  const float4 L0 = g.leaf[4 * li], L1 = g.leaf[4 * li + 1], L2 = g.leaf[4 * li + 2], L3 = g.leaf[4 * li + 3];
  const int typ = __float_as_int(L3.y), p0 = __float_as_int(L3.z), np = __float_as_int(L3.w);
  const float3 y = sub3(o, mk3(L0.x, L0.y, L0.z));
  float tp0 = -HUGEF, tp1 = HUGEF;
  int fp0 = -1, fp1 = -1;
  bool empty = false;
  for (int k = 0; k < np; ++k) {
    const float4 P = g.plane[p0 + k];
    const float nd = P.x * d.x + P.y * d.y + P.z * d.z;
    const float rhs = P.w - (P.x * y.x + P.y * y.y + P.z * y.z);
    if (fabsf(nd) < 1e-12f) {
      empty |= rhs < 0.0f;
    } else {
      const float tp = rhs / nd;
      if (nd < 0.0f) {
        if (tp > tp0) { tp0 = tp; fp0 = k + 1; }
      } else {
        if (tp < tp1) { tp1 = tp; fp1 = k + 1; }
      }
    }
  }
  // the quadric's pieces [a0, b0] and [a1, b1] (a piece with a > b is empty); qa/qb: the end is a root
  float a0 = -HUGEF, b0 = HUGEF, a1 = HUGEF, b1 = -HUGEF;
  bool qa0 = false, qb0 = false, qa1 = false, qb1 = false;
  if (TOR && typ == 2) {
    torus_intervals(L1.x, L1.y, mk3(L2.z, L2.w, L3.x), y, d, a0, b0, a1, b1);
    qa0 = qb0 = qa1 = qb1 = true;
  } else if (typ == 1) {
    const float A00 = L1.x, A01 = L1.y, A02 = L1.z, A11 = L1.w, A12 = L2.x, A22 = L2.y;
    const float3 b = mk3(L2.z, L2.w, L3.x);
    const float c = L0.w;
    const float sa = fmaxf(fmaxf(fmaxf(fabsf(A00), fabsf(A01)), fmaxf(fabsf(A02), fabsf(A11))),
                           fmaxf(fabsf(A12), fabsf(A22)));
    auto Av = [&](float3 v) {
      return mk3(A00 * v.x + A01 * v.y + A02 * v.z, A01 * v.x + A11 * v.y + A12 * v.z,
                 A02 * v.x + A12 * v.y + A22 * v.z);
    };
    const float3 Ad = Av(d);
    const float a = dot3(d, Ad);
    const float bd = dot3(b, d);
    const bool lin = fabsf(a) <= 16.0f * EPS32 * sa;
    float tc = -dot3(y, d);
    float3 yc = fma3(tc, d, y);
    float3 Ay = Av(yc);
    float beta = dot3(Ay, d) + bd;
    float gam = dot3(yc, Ay) + 2.0f * dot3(b, yc) + c;
    if (!lin && a * gam > 0.5f * beta * beta) {  // the roots lie near the vertex: take them there
      tc -= beta / a;
      yc = fma3(tc, d, y);
      Ay = Av(yc);
      beta = dot3(Ay, d) + bd;
      gam = dot3(yc, Ay) + 2.0f * dot3(b, yc) + c;
    }
    if (lin) {
      if (beta > 0.0f) {
        a0 = -HUGEF; b0 = tc - gam / (2.0f * beta); qb0 = true;
      } else if (beta < 0.0f) {
        a0 = tc - gam / (2.0f * beta); b0 = HUGEF; qa0 = true;
      } else if (gam > 0.0f) {
        a0 = HUGEF; b0 = -HUGEF;
      }
    } else {
      const float D = beta * beta - a * gam;
      if (D < 0.0f) {
        if (a > 0.0f) { a0 = HUGEF; b0 = -HUGEF; }
      } else {
        const float sq = sqrtf(D);
        const float qq = -(beta + copysignf(sq, beta));
        float r1 = qq / a, r2 = qq != 0.0f ? gam / qq : r1;
        const float rl = fminf(r1, r2), rh = fmaxf(r1, r2);
        if (a > 0.0f) {
          a0 = tc + rl; b0 = tc + rh; qa0 = qb0 = true;
        } else {
          a0 = -HUGEF; b0 = tc + rl; qb0 = true;
          a1 = tc + rh; b1 = HUGEF; qa1 = true;
        }
      }
    }
  }
  // clip each piece by the planes' slab
  {
    const float lo = fmaxf(a0, tp0), hi = fminf(b0, tp1);
    const bool ok = lo < hi && !empty;
    iv.lo[0] = ok ? lo : HUGEF;
    iv.hi[0] = ok ? hi : HUGEF;
    iv.flo[0] = (a0 > tp0 && qa0) ? 0 : fp0;
    iv.fhi[0] = (b0 < tp1 && qb0) ? 0 : fp1;
  }
  {
    const float lo = fmaxf(a1, tp0), hi = fminf(b1, tp1);
    const bool ok = lo < hi && !empty;
    iv.lo[1] = ok ? lo : HUGEF;
    iv.hi[1] = ok ? hi : HUGEF;
    iv.flo[1] = (a1 > tp0 && qa1) ? 0 : fp0;
    iv.fhi[1] = (b1 < tp1 && qb1) ? 0 : fp1;
  }
}

// a leaf's membership of a point (its kind's frame)
template <bool TOR>
__device__ inline bool leaf_inside(const Geo &g, int li, float3 x) {
  const float4 L0 = g.leaf[4 * li], L3 = g.leaf[4 * li + 3];
  const int typ = __float_as_int(L3.y), p0 = __float_as_int(L3.z), np = __float_as_int(L3.w);
  const float3 y = sub3(x, mk3(L0.x, L0.y, L0.z));
  for (int k = 0; k < np; ++k) {
    const float4 P = g.plane[p0 + k];
    if (P.x * y.x + P.y * y.y + P.z * y.z > P.w) return false;
  }
  if (typ == 1) {
    const float4 L1 = g.leaf[4 * li + 1], L2 = g.leaf[4 * li + 2];
    const float3 Ay = mk3(L1.x * y.x + L1.y * y.y + L1.z * y.z, L1.y * y.x + L1.w * y.y + L2.x * y.z,
                          L1.z * y.x + L2.x * y.y + L2.y * y.z);
    const float q = dot3(y, Ay) + 2.0f * (L2.z * y.x + L2.w * y.y + L3.x * y.z) + L0.w;
    if (q > 0.0f) return false;
  } else if (TOR && typ == 2) {  // |y|^2 - 2 R rho + R^2 - r^2 <= 0
    const float4 L1 = g.leaf[4 * li + 1], L2 = g.leaf[4 * li + 2];
    const float yy = dot3(y, y), z = dot3(y, mk3(L2.z, L2.w, L3.x));
    const float rho = sqrtf(fmaxf(yy - z * z, 0.0f));
    if (yy - 2.0f * L1.x * rho + L1.x * L1.x - L1.y * L1.y > 0.0f) return false;
  }
  return true;
}

// postfix program over leaf bits (bit k: leaf l0 + k is in)
__device__ __forceinline__ bool run_program(const Geo &g, int p0, int pl, uint32_t bits) {
  uint32_t st = 0;
  int sp = 0;
  for (int k = 0; k < pl; ++k) {
    const int w = g.prog[p0 + k];
    const int op = w >> 24, arg = w & 0xFFFFFF;
    if (op == 0) {
      st = (st & ~(1u << sp)) | (((bits >> arg) & 1u) << sp);
      ++sp;
    } else {
      const uint32_t b = (st >> (sp - 1)) & 1u, a = (st >> (sp - 2)) & 1u;
      const uint32_t r = op == 1 ? (a | b) : op == 2 ? (a & b) : (a & ~b & 1u);
      sp -= 1;
      st = (st & ~(1u << (sp - 1))) | (r << (sp - 1));
    }
  }
  return (st & 1u) != 0u;
}

template <bool TOR>
__device__ inline bool solid_inside(const Geo &g, int4 K, float3 x) {
  if (K.y == 1) return leaf_inside<TOR>(g, K.x, x);
  uint32_t bits = 0;
  for (int k = 0; k < K.y && k < 32; ++k) bits |= (uint32_t)leaf_inside<TOR>(g, K.x + k, x) << k;
  return run_program(g, K.z, K.w, bits);
}

struct Change {
  float t;
  int leaf, face;  // leaf within its kind; -1: none
  float3 n;        // the face's normal there, in the solid's frame (not normalised)
};

// the normal of leaf li's face `face` (0 the quadric: A y + b at y = o + t d - q; k >= 1 plane k) in its frame; the
// leaf's data was just read by its interval test, so these reads hit the cache
template <bool TOR>
__device__ __forceinline__ float3 leaf_normal(const Geo &g, int li, int face, float3 o, float3 d, float t) {
  const float4 L0 = g.leaf[4 * li], L1 = g.leaf[4 * li + 1], L2 = g.leaf[4 * li + 2], L3 = g.leaf[4 * li + 3];
  if (face == 0) {
    const float3 y = sub3(fma3(t, d, o), mk3(L0.x, L0.y, L0.z));
    if (TOR && __float_as_int(L3.y) == 2) {  // a torus: y - R (the point's direction across the axis)
      const float3 w = mk3(L2.z, L2.w, L3.x);
      const float3 q = fma3(-dot3(y, w), w, y);
      const float rho = fmaxf(sqrtf(dot3(q, q)), 1e-30f);
      return fma3(-L1.x / rho, q, y);
    }
    return mk3(L1.x * y.x + L1.y * y.y + L1.z * y.z + L2.z, L1.y * y.x + L1.w * y.y + L2.x * y.z + L2.w,
               L1.z * y.x + L2.x * y.y + L2.y * y.z + L3.x);
  }
  const float4 P = g.plane[__float_as_int(L3.z) + face - 1];
  return mk3(P.x, P.y, P.z);
}


// the first change of the solid's membership along the ray after t = 0, from the side `want_from` (true: the
// point starts inside and the change is an exit; false: an entry).  Endpoints on the face `skip_code` within
// `skip_t` of the start count as passed.  `vol` names the volume in the surface codes.
template <bool TOR>
__device__ inline Change first_change(const Geo &g, int4 K, int vol, float3 o, float3 d, bool exit_query,
                                      int skip_code, float skip_t) {
  Change ch{HUGEF, -1, -1, mk3(0.f, 0.f, 1.f)};
  if (K.y == 1) {  // one leaf: its intervals are the solid's
    LeafIv iv;
    leaf_intervals<TOR>(g, K.x, o, d, iv);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float lo = iv.lo[p], hi = iv.hi[p];
      if (lo >= HUGEF) continue;
      if (srf_code(vol, 0, iv.flo[p]) == skip_code && fabsf(lo) <= skip_t) lo = -BIGF;
      if (srf_code(vol, 0, iv.fhi[p]) == skip_code && fabsf(hi) <= skip_t) hi = -BIGF;
      if (exit_query) {
        if (lo <= 0.0f && 0.0f < hi && hi < ch.t) { ch.t = hi; ch.leaf = 0; ch.face = iv.fhi[p]; }
      } else {
        if (lo > 0.0f && lo < hi && lo < ch.t) { ch.t = lo; ch.leaf = 0; ch.face = iv.flo[p]; }
      }
    }
    if (exit_query && ch.leaf < 0)  // the start is outside its own solid (rounding): leave it where it is, the
      return Change{0.0f, 0, 0, mul3(-1.0f, d)};  // normal facing the photon (no face is met)
    if (ch.leaf >= 0) ch.n = leaf_normal<TOR>(g, K.x, ch.face, o, d, ch.t);
    return ch;
  }
  // several leaves: scan the endpoints in order for the first change of the program's value
  const int nl = K.y < MAX_CSG_LEAVES ? K.y : MAX_CSG_LEAVES;
  LeafIv ivs[MAX_CSG_LEAVES];
  for (int k = 0; k < nl; ++k) {
    leaf_intervals<TOR>(g, K.x + k, o, d, ivs[k]);
    for (int p = 0; p < 2; ++p) {
      if (ivs[k].lo[p] >= HUGEF) continue;
      if (srf_code(vol, k, ivs[k].flo[p]) == skip_code && fabsf(ivs[k].lo[p]) <= skip_t) ivs[k].lo[p] = -BIGF;
      if (srf_code(vol, k, ivs[k].fhi[p]) == skip_code && fabsf(ivs[k].hi[p]) <= skip_t) ivs[k].hi[p] = -BIGF;
    }
  }
  auto member = [&](float t) {
    uint32_t bits = 0;
    for (int k = 0; k < nl; ++k) {
      const bool in = (ivs[k].lo[0] <= t && t < ivs[k].hi[0]) || (ivs[k].lo[1] <= t && t < ivs[k].hi[1]);
      bits |= (uint32_t)in << k;
    }
    return run_program(g, K.z, K.w, bits);
  };
  const bool start = member(0.0f);
  if (exit_query && !start) return Change{0.0f, 0, 0, mul3(-1.0f, d)};  // outside its own solid: leave it now
  if (!exit_query && start) return ch;  // inside a daughter already (rounding): no entry ahead
  float cur = 0.0f;
  for (int it = 0; it < 4 * nl + 5; ++it) {
    float nxt = HUGEF;
    int nl_ = -1, nf = -1;
    for (int k = 0; k < nl; ++k)
      for (int p = 0; p < 2; ++p) {
        if (ivs[k].lo[p] > cur && ivs[k].lo[p] < nxt) { nxt = ivs[k].lo[p]; nl_ = k; nf = ivs[k].flo[p]; }
        if (ivs[k].hi[p] > cur && ivs[k].hi[p] < nxt && ivs[k].hi[p] < HUGEF) {
          nxt = ivs[k].hi[p]; nl_ = k; nf = ivs[k].fhi[p];
        }
      }
    if (nl_ < 0 || nxt >= HUGEF) break;
    // a change only where the value stays changed a band further on: two leaves meeting on a shared face (a
    // polycone's sections, a union of touching pieces) may leave a gap of a few ulps between their intervals
    if (member(nxt) != start && member(nxt + band_at(maxabs3(o) + fabsf(nxt))) != start)
      return Change{nxt, nl_, nf, leaf_normal<TOR>(g, K.x + nl_, nf, o, d, nxt)};
    cur = nxt;
  }
  if (exit_query) return Change{0.0f, 0, 0, mul3(-1.0f, d)};
  return ch;
}

struct Cross {
  float t;
  int vol_after;  // navigator id entered; -1: no crossing within reach
  int srf;        // the crossed face's code
  int own;        // 1: the exit of the photon's own volume
  float3 n;       // unit normal facing the photon
};

// the next visible crossing from x along d in navigator id `vol`, within tmax; `last` the face the photon sits on
// the stretch [lo, hi] of the ray xl + t dl (the volume's frame) inside the volume's daughter-free core: a box
// (lo, hi about the centre c, +-HUGEF where open) cut by a cylinder about the z axis through c (radius^2 r2, HUGEF
// none): records (c, r2), (lo, code), (hi, leaf); code 0 no core, 1 this form, 2 any other leaf (its intervals)
__device__ inline void core_interval(const Geo &g, int vol, float3 xl, float3 dl, float &lo, float &hi) {
  const float4 C1 = g.vcore[3 * vol + 1];
  const int code = __float_as_int(C1.w);
  lo = HUGEF;
  hi = -HUGEF;
  if (code == 0) return;
  const float4 C0 = g.vcore[3 * vol], C2 = g.vcore[3 * vol + 2];
  if (code == 2) {
    LeafIv iv;
    leaf_intervals<false>(g, __float_as_int(C2.w), xl, dl, iv);
    lo = iv.lo[0];
    hi = iv.hi[0];
    return;
  }
  const float3 y = sub3(xl, mk3(C0.x, C0.y, C0.z));
  const float ix = 1.0f / (fabsf(dl.x) > 1e-30f ? dl.x : copysignf(1e-30f, dl.x));
  const float iy = 1.0f / (fabsf(dl.y) > 1e-30f ? dl.y : copysignf(1e-30f, dl.y));
  const float iz = 1.0f / (fabsf(dl.z) > 1e-30f ? dl.z : copysignf(1e-30f, dl.z));
  const float tx0 = (C1.x - y.x) * ix, tx1 = (C2.x - y.x) * ix;
  const float ty0 = (C1.y - y.y) * iy, ty1 = (C2.y - y.y) * iy;
  const float tz0 = (C1.z - y.z) * iz, tz1 = (C2.z - y.z) * iz;
  float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
  float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
  if (C0.w < HUGEF) {  // the cylinder: (y + t d)_xy^2 <= r2
    const float a = dl.x * dl.x + dl.y * dl.y, b = y.x * dl.x + y.y * dl.y;
    const float c = y.x * y.x + y.y * y.y - C0.w;
    const float disc = b * b - a * c;
    if (a > 0.0f && disc > 0.0f) {
      const float q = -(b + copysignf(sqrtf(disc), b));
      const float r0 = q / a, r1 = q != 0.0f ? c / q : r0;
      t0 = fmaxf(t0, fminf(r0, r1));
      t1 = fminf(t1, fmaxf(r0, r1));
    } else if (!(a == 0.0f && c <= 0.0f)) {  // misses the cylinder (or runs along it outside)
      t1 = -HUGEF;
    }
  }
  if (t0 < t1) {
    lo = t0;
    hi = t1;
  }
}

template <bool TOR>
__device__ inline int locate(const Geo &g, float3 x, int start = -1);

// `scale`: the coordinate scale of the ray that put the start on face `last` (its crossing's rounding is within
// that ray's band, not the start's own: a crossing near the origin reached from far away)
template <bool LDS, bool TOR>
__device__ inline Cross query_t(const Geo &g, float3 x, float3 d, int vol, int last, float tmax, float scale = 0.0f) {
  const float skip_t = K_ON * band_at(fmaxf(maxabs3(x), scale));
  Cross best{tmax, -1, -1, 0, mk3(0.f, 0.f, 1.f)};
  int best_leaf = -1, best_face = -1, best_vol = -1;
  // the ray's stretch [core_lo, core_hi] inside the volume's daughter-free core: a box the ray crosses only there
  // holds nothing it can hit (the parameter t is the same in every frame: the transforms are rigid)
  float core_lo = HUGEF, core_hi = -HUGEF;
  if (vol < g.nvol) {
    float3 xl, dl;
    to_local(g, vol, x, d, xl, dl);
    core_interval(g, vol, xl, dl, core_lo, core_hi);
    const Change ch = first_change<TOR>(g, g.vkind[vol], vol, xl, dl, true, last, skip_t);
    if (ch.leaf >= 0 && ch.t <= best.t) {
      best.t = ch.t;
      best.vol_after = g.vinfo[vol].y;
      best.own = 1;
      best_leaf = ch.leaf;
      best_face = ch.face;
      best_vol = vol;
      best.n = normal_to_world(g, vol, ch.n);
    }
  }
  const int2 gr = g.vgrp[vol];
  for (int gi = gr.x; gi < gr.y; ++gi) {
    const float4 q0 = g.gxf[3 * gi], q1 = g.gxf[3 * gi + 1], q2 = g.gxf[3 * gi + 2];
    float3 xg = x, dg = d;
    if (q0.w == 0.0f) {
      xg = mk3(dot3(mk3(q0.x, q0.y, q0.z), x), dot3(mk3(q1.x, q1.y, q1.z), x), dot3(mk3(q2.x, q2.y, q2.z), x));
      dg = mk3(dot3(mk3(q0.x, q0.y, q0.z), d), dot3(mk3(q1.x, q1.y, q1.z), d), dot3(mk3(q2.x, q2.y, q2.z), d));
    }
    const float3 inv = mk3(1.0f / (fabsf(dg.x) > 1e-30f ? dg.x : copysignf(1e-30f, dg.x)),
                           1.0f / (fabsf(dg.y) > 1e-30f ? dg.y : copysignf(1e-30f, dg.y)),
                           1.0f / (fabsf(dg.z) > 1e-30f ? dg.z : copysignf(1e-30f, dg.z)));
    const int2 nr = g.grange[gi];
    const float3 xi = mk3(xg.x * inv.x, xg.y * inv.y, xg.z * inv.z);  // the slabs as one fma each
    int ni = nr.x;
    // the walk in two loops (traverse collecting up to WALK_PENDING leaves, then test their items): a wavefront
    // runs the costly item tests once for all lanes that found leaves, instead of on every node step where any one
    // lane sits at a leaf, and a lane that found a leaf walks on while others look for theirs.  A node step is
    // branch-free: a missed box or a leaf moves on to the skip link, an inner box to its first child.
    for (;;) {
      int npend = 0, pl0 = 0, pl1 = 0, pl2 = 0;  // the leaves found (start * 8 + count), the latest in pl0
      while (ni < nr.y && npend < WALK_PENDING) {
        float4 lo, hi;
        nav_node<LDS>(g, ni, lo, hi);
        const float tx0 = fmaf(lo.x, inv.x, -xi.x), tx1 = fmaf(hi.x, inv.x, -xi.x);
        const float ty0 = fmaf(lo.y, inv.y, -xi.y), ty1 = fmaf(hi.y, inv.y, -xi.y);
        const float tz0 = fmaf(lo.z, inv.z, -xi.z), tz1 = fmaf(hi.z, inv.z, -xi.z);
        const float te = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        const float tl = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), best.t));
        const int sc = __float_as_int(hi.w);
        // non-short-circuit: the compares as a mask, no branch per node (3 % of the tank's forward)
        const bool in = !(te > tl) & !((te >= core_lo) & (tl <= core_hi));
        const bool leaf = in & ((sc & 7) != 0);
        pl2 = leaf ? pl1 : pl2;
        pl1 = leaf ? pl0 : pl1;
        pl0 = leaf ? sc : pl0;
        npend += leaf ? 1 : 0;
        ni = (in && !leaf) ? ni + 1 : __float_as_int(lo.w);
      }
      if (npend == 0) break;
      for (int q = npend - 1; q >= 0; --q) {  // oldest first: the walk's order
        const int sc = q == 2 ? pl2 : q == 1 ? pl1 : pl0;
        const int pst = sc >> 3, pcnt = sc & 7;
        for (int k = 0; k < pcnt; ++k) {  // the item's transform and kind by item index: no wait on its volume id
          const int it = pst + k;
          const int u = g.items[it];
          float3 xl, dl;
          rows_to_local(g.ixf + 3 * it, x, d, xl, dl);
          const Change ch = first_change<TOR>(g, g.ikind[it], u, xl, dl, false, last, skip_t);
          if (ch.leaf >= 0 && (ch.t < best.t || (ch.t == best.t && best.own))) {
            best.t = ch.t;
            best.vol_after = u;
            best.own = 0;
            best_leaf = ch.leaf;
            best_face = ch.face;
            best_vol = u;
            best.n = rows_normal_to_world(g.ixf + 3 * it, ch.n);
          }
        }
      }
    }
  }
  if (best_vol >= 0) {  // the winner's normal, computed by its own test: unit, facing the photon
    best.srf = srf_code(best_vol, best_leaf, best_face);
    best.n = unit3(best.n);
    if (dot3(best.n, d) > 0.0f) best.n = mul3(-1.0f, best.n);
  }
  // an exit into the mother at a point another volume holds (touching siblings, a daughter flush with its
  // mother's face, whose exit is the mother's too): the crossing goes on into the volume holding the point K_ON
  // bands past the face (the reference's relocation past the face met, navigate.py:526-555), the crossing point
  // moved there and the face behind the photon
  // An entry whose point past the face a daughter of the entered volume holds (a daughter flush with its mother's
  // face) goes on into that daughter alike: faces within K_ON bands are crossed as one
  // Looked for only where it can happen: past an exit, outside the mother (a daughter flush with its face), or in
  // a sibling whose box comes near (vinfo.x, the touch flag); past an entry, in the entered volume's daughters
  if (best.vol_after >= 0 && best.vol_after < g.nvol) {
    const float step = K_ON * band_at(maxabs3(x) + best.t);
    const float3 xp = fma3(best.t + step, d, x);
    int v2 = best.vol_after;
    if (best.own) {
      float3 xl, dl;
      to_local(g, best.vol_after, xp, d, xl, dl);
      if (!solid_inside<TOR>(g, g.vkind[best.vol_after], xl)) v2 = locate<TOR>(g, xp);
      else if (g.vinfo[vol].x & 1) v2 = locate<TOR>(g, xp, best.vol_after);
    } else {
      const int2 gr = g.vgrp[best.vol_after];
      if (gr.x < gr.y) v2 = locate<TOR>(g, xp, best.vol_after);
    }
    if (v2 != vol && v2 != best.vol_after) {
      best.vol_after = v2;
      best.t += step;
      best.srf = -1;
    }
  }
  return best;
}

// The reach bound (the reference's geometry/reach.py:1-40, 174-222): the limit for the query of a step whose free
// path is s, from x along d in volume vol, not sitting on a face, within tmax.  Every crossing the query can return
// lies in the volume's padded box B (its own exit, a daughter's entry, or a point the relocation step past one), and
// the ray leaves the convex B once, at T; the crossing deposits only where it enters a sensor volume or leaves the
// one the photon is in, at a point of that volume's box.  So where s < T and the segment [0, T] meets no sensor
// volume's padded box, the query may stop a few bands past s: every crossing at t <= s is found as before (the
// limit only prunes candidates beyond it), and one beyond s, a bulk event either way, deposits nothing.  The boxes'
// padding (twice the relocation step at the geometry's scale) holds the float32 slab tests' rounding.  Otherwise,
// or for a start outside B, the limit is tmax.  The sensor boxes are walked any-hit: the first leaf box met ends it.
__device__ inline float reach_limit(const Geo &g, float3 x, float3 d, int vol, float s, float tmax) {
  const float4 lo = g.rbox[2 * vol], hi = g.rbox[2 * vol + 1];
  if (!(x.x >= lo.x && x.x <= hi.x && x.y >= lo.y && x.y <= hi.y && x.z >= lo.z && x.z <= hi.z)) return tmax;
  const float3 inv = mk3(1.0f / (fabsf(d.x) > 1e-30f ? d.x : copysignf(1e-30f, d.x)),
                         1.0f / (fabsf(d.y) > 1e-30f ? d.y : copysignf(1e-30f, d.y)),
                         1.0f / (fabsf(d.z) > 1e-30f ? d.z : copysignf(1e-30f, d.z)));
  const float T = fminf(fminf(fmaxf((lo.x - x.x) * inv.x, (hi.x - x.x) * inv.x),
                              fmaxf((lo.y - x.y) * inv.y, (hi.y - x.y) * inv.y)),
                        fmaxf((lo.z - x.z) * inv.z, (hi.z - x.z) * inv.z));
  const float lim = s + (K_ON + 2.0f) * band_at(maxabs3(x) + s);
  if (!(s < T) || !(lim < tmax)) return tmax;
  const float3 xi = mk3(x.x * inv.x, x.y * inv.y, x.z * inv.z);
  int ni = 0;
  while (ni < g.n_rnode) {
    const float4 a = g.rnode[2 * ni], b = g.rnode[2 * ni + 1];
    const float tx0 = fmaf(a.x, inv.x, -xi.x), tx1 = fmaf(b.x, inv.x, -xi.x);
    const float ty0 = fmaf(a.y, inv.y, -xi.y), ty1 = fmaf(b.y, inv.y, -xi.y);
    const float tz0 = fmaf(a.z, inv.z, -xi.z), tz1 = fmaf(b.z, inv.z, -xi.z);
    const float te = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    const float tl = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), T));
    const bool in = !(te > tl);
    if (in && __float_as_int(b.w) != 0) return tmax;  // a sensor volume's box on the segment
    ni = in ? ni + 1 : __float_as_int(a.w);
  }
  return lim;
}

template <bool TOR>
__device__ inline Cross query(const Geo &g, float3 x, float3 d, int vol, int last, float tmax) {
  return query_t<false, TOR>(g, x, d, vol, last, tmax);
}

// whether a point (the volume's frame) lies in the volume's daughter-free core: then no daughter holds it.  The core
// is the solid shrunk by its daughters' depth plus a margin, so a rounding at its boundary cannot hide a daughter
__device__ inline bool core_holds(const Geo &g, int vol, float3 xl) {
  const float4 C1 = g.vcore[3 * vol + 1];
  const int code = __float_as_int(C1.w);
  if (code == 0) return false;
  const float4 C0 = g.vcore[3 * vol], C2 = g.vcore[3 * vol + 2];
  if (code == 2) return leaf_inside<false>(g, __float_as_int(C2.w), xl);  // a core leaf is never a torus
  const float3 y = sub3(xl, mk3(C0.x, C0.y, C0.z));
  if (y.x < C1.x || y.x > C2.x || y.y < C1.y || y.y > C2.y || y.z < C1.z || y.z > C2.z) return false;
  return !(C0.w < HUGEF && y.x * y.x + y.y * y.y > C0.w);
}

// the deepest navigator id holding x (the outside when no volume does), searched from `start` (-1: the outside)
template <bool TOR>
__device__ inline int locate(const Geo &g, float3 x, int start) {
  int cur = start < 0 ? g.nvol : start;
  for (int depth = 0; depth < 64; ++depth) {
    if (cur < g.nvol) {  // inside the current volume's daughter-free core: no daughter to look for
      float3 xl, dl;
      to_local(g, cur, x, mk3(0.f, 0.f, 1.f), xl, dl);
      if (core_holds(g, cur, xl)) return cur;
    }
    int found = -1;
    const int2 gr = g.vgrp[cur];
    for (int gi = gr.x; gi < gr.y && found < 0; ++gi) {
      const float4 q0 = g.gxf[3 * gi], q1 = g.gxf[3 * gi + 1], q2 = g.gxf[3 * gi + 2];
      float3 xg = x;
      if (q0.w == 0.0f)
        xg = mk3(dot3(mk3(q0.x, q0.y, q0.z), x), dot3(mk3(q1.x, q1.y, q1.z), x), dot3(mk3(q2.x, q2.y, q2.z), x));
      const int2 nr = g.grange[gi];
      int ni = nr.x;
      while (ni < nr.y && found < 0) {
        const float4 lo = g.node[2 * ni], hi = g.node[2 * ni + 1];
        const int skip = __float_as_int(lo.w);
        if (xg.x < lo.x || xg.x > hi.x || xg.y < lo.y || xg.y > hi.y || xg.z < lo.z || xg.z > hi.z) {
          ni = skip;
          continue;
        }
        const int sc = __float_as_int(hi.w);
        const int cnt = sc & 7;
        if (cnt > 0) {
          for (int k = 0; k < cnt; ++k) {
            const int u = g.items[(sc >> 3) + k];
            float3 xl, dl;
            to_local(g, u, x, mk3(0.f, 0.f, 1.f), xl, dl);
            if (solid_inside<TOR>(g, g.vkind[u], xl)) { found = u; break; }
          }
          ni = skip;
        } else {
          ni += 1;
        }
      }
    }
    if (found < 0) return cur;
    cur = found;
  }
  return cur;
}

}  // namespace lh
