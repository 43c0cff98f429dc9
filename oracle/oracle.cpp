// THE ORACLE -- test infrastructure, not the product.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it, and only as the checker (or as the CPU baseline that is timed); the product path
// (lumen_amd) never calls it and has no CPU fallback.
//
// A C++17/OpenMP restatement of the reference's photon loop, in FP64, written from the reference's algorithm:
//   loop and fates        src/lumen/run.py:293-368 (_loop), src/lumen/transport/item.py:20-36
//   one step              src/lumen/transport/step.py:318-359 (_step_context), :362-394 (_surface_event),
//                         :397-436 (_bulk), :464-492 (_after), :495-513 (_advance), :516-571 (step)
//   estimator             src/lumen/transport/estimator.py:50-153 (proposal, choose, defensive, edge_floor)
//   media                 src/lumen/media/bulk.py:61-110, src/lumen/media/laws.py:97-181 (tables, GROUPVEL)
//   wavelength shifter    src/lumen/media/emission.py:120-159 (EmissionRecord.sample), src/lumen/media/spectra.py:37-137
//                         (g4_arrays, covered, cumulative, invert, on_intervals, kept), src/lumen/rules/optics.py:319-350
//   surfaces              src/lumen/surfaces/types.py:211-296, :352-389 (chained_cuts, _none, _real, kernel),
//                         src/lumen/surfaces/fresnel.py:25-45, src/lumen/surfaces/outcomes.py:36-75,
//                         src/lumen/surfaces/polar.py:39-118
//   lobes                 src/lumen/media/phase.py:45-75 (Rayleigh), :98-190 (Mie: G4OpMieHG's two Henyey-Greenstein
//                         lobes), src/lumen/transport/medium.py:96-133 (channel_draw), src/lumen/lobes.py:60-78 (Lambert),
//                         src/lumen/surfaces/facets.py:318-397 (unified facet path)
//   sources               src/lumen/sources/point.py:73-86, src/lumen/sources/cherenkov.py:33-66
//   readout               src/lumen/run.py:483-528 (_moment_adds: a photon's deposits per sensor summed, then
//                         squared), src/lumen/readout/fill.py:88-213
//   navigator referee     src/lumen/compare/oracle.py:1-24, :208 (FP64 labels and the first change along a ray),
//                         here by brute force over every daughter (no BVH), each solid's CSG leaves evaluated in FP64.
// Random numbers: Philox4x32-10 (Salmon et al., SC'11) with the same (seed, photon id, vertex, block) contract as the
// engine, restated here independently.
//
// Parity status per function (the checks that pin it live in tests/test_oracle.py):
//   philox            pinned: Random123 known-answer vectors
//   tables, GROUPVEL  pinned: closed forms (Water Rayleigh lengths at the RINDEX knots, n and L_abs at 400 nm)
//   navigator         pinned: closed-form ray-box / ray-cylinder / ray-sphere-cap distances
//   facet path        pinned: Geant4 11.3.0 histograms (tests/golden/g4_unified_metal_*.npz), two-sample chi2
//   Rayleigh, Lambert pinned: their closed-form angular densities (chi2 on the cosine)
//   Fresnel           pinned: textbook normal-incidence and Brewster-angle values, energy conservation
//   whole loop        pinned: the tutorial's published arrivals (example_tank), the specular-sphere closed form
//   Cherenkov         pinned: Frank-Tamm photon yield per metre in closed form; and (tests/test_oracle_emission_jones.py)
//                     every photon on its cone cos = 1/(beta n(lambda)) under a dispersive index (1e-12), the weighted
//                     spectrum against Frank-Tamm's dN/dlambda (chi2), the start uniform along the track (KS) at
//                     t = s/(beta c), the polarisation transverse, in the track's plane, against the track; the cone
//                     toy's ring shares and times against the geometry's closed form
//   Jones steps       pinned (tests/test_oracle_emission_jones.py): reflection and refraction recombined from the
//                     Fresnel amplitudes satisfy Maxwell's boundary conditions (tangential E, H; normal D) for complex
//                     states; total internal reflection's s-p phase (Born and Wolf); normal incidence gives r E, t E;
//                     the mirror is a perfect conductor's (E + E' along h, no loss, transverse)
//   moment error      pinned: brute force on a small run
//   WLS emission      pinned (tests/test_oracle_wls.py): the trapezoid sampler's interval masses (chi2), the truncation
//                     at the absorbed energy with weight yield (1 - (1 - F)^100), the live/reference reweighting
//                     (importance-weighted mean energy), the exponential delay (KS), an enclosing sensor: Q/N = yield
//   Mie lobes         pinned (tests/test_oracle_mie.py): the scattering angle against the mixture's closed-form CDF in
//                     equal-mass bins (chi2), the Henyey-Greenstein moments <P_l> = r g_f^l + (1 - r)(-g_b)^l, the
//                     Geant4 10.7.4 / 11.2.2 single-scatter histograms (tests/golden/g4_*_mie*_hist.npz), the density's
//                     normalisation and the live/drawn weight's unit mean
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/lumen_hip.h"

namespace orc {

// ------------------------------------------------------------------------------------------ the scalar types
// Counted: a double that counts the floating-point operations made with it (add, sub, mul, div, compare and each
// math function as one; fma as two), for the algorithmic work per photon.
thread_local long long g_ops = 0;

struct Counted {
  double v;
  Counted() : v(0) {}
  Counted(double x) : v(x) {}
  explicit operator double() const { return v; }
};
inline Counted operator+(Counted a, Counted b) { ++g_ops; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { ++g_ops; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { ++g_ops; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { ++g_ops; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted &operator+=(Counted &a, Counted b) { ++g_ops; a.v += b.v; return a; }
inline Counted &operator-=(Counted &a, Counted b) { ++g_ops; a.v -= b.v; return a; }
inline Counted &operator*=(Counted &a, Counted b) { ++g_ops; a.v *= b.v; return a; }
inline Counted &operator/=(Counted &a, Counted b) { ++g_ops; a.v /= b.v; return a; }
inline bool operator<(Counted a, Counted b) { ++g_ops; return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { ++g_ops; return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { ++g_ops; return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { ++g_ops; return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { ++g_ops; return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { ++g_ops; return a.v != b.v; }

inline double val(double x) { return x; }
inline double val(float x) { return x; }
inline double val(Counted x) { return x.v; }
#define ORC_F1(name, f)                                  \
  inline double name(double x) { return f(x); }           \
  inline float name(float x) { return f(x); }             \
  inline Counted name(Counted x) { ++g_ops; return Counted(f(x.v)); }
ORC_F1(SQRT, std::sqrt)
ORC_F1(CBRT, std::cbrt)
ORC_F1(EXP, std::exp)
ORC_F1(LOG, std::log)
ORC_F1(SIN, std::sin)
ORC_F1(COS, std::cos)
ORC_F1(ACOS, std::acos)
ORC_F1(FABS, std::fabs)
inline double ATAN2(double y, double x) { return std::atan2(y, x); }
inline float ATAN2(float y, float x) { return std::atan2(y, x); }
inline Counted ATAN2(Counted y, Counted x) { ++g_ops; return Counted(std::atan2(y.v, x.v)); }
template <class R> inline R MIN(R a, R b) { return b < a ? b : a; }
template <class R> inline R MAX(R a, R b) { return a < b ? b : a; }
template <class R> inline R CLIP01(R x) { return MIN<R>(MAX<R>(x, R(0.0)), R(1.0)); }

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> unit(V3<R> a) {
  R s = dot(a, a);
  if (s > R(0.0)) return (R(1.0) / SQRT(s)) * a;
  return a;
}

// ------------------------------------------------------------------------------------------ Philox4x32-10
inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// integer operations of one Philox call, counted for the work model: per round 2 32x32->64 multiplies (as 2 ops
// each: low and high), 4 xors, 2 key adds
constexpr long long PHILOX_INT_OPS = 10 * (4 + 4 + 2);
thread_local long long g_int_ops = 0;

struct Rng {
  uint32_t k0, k1, id_lo, id_hi;
  void draw4(uint32_t vertex, uint32_t block, double u[4]) const {
    uint32_t c[4] = {id_lo, id_hi, vertex, block};
    philox(c, k0, k1);
    g_int_ops += PHILOX_INT_OPS;
    for (int i = 0; i < 4; ++i) u[i] = ((double)(c[i] >> 9) + 0.5) * (1.0 / 8388608.0);
  }
};
constexpr uint32_t FACET_BLOCK = 0x100u, SOURCE_VERTEX = 0xFFFFFFFFu;

// ------------------------------------------------------------------------------------------ constants
constexpr double HC = 1239.8419843320026, C_VACUUM = 0.299792458, TWO_PI = 6.283185307179586;
constexpr double PI = 3.141592653589793, HALF_PI = 1.5707963267948966;
constexpr double FLOOR32 = 1e-30, TINY_NORMAL = 1.1754943508222875e-38, DEFENSIVE = 0.01;
constexpr double FINE_STRUCTURE = 1.0 / 137.035999084;
constexpr double ZERO_STEP = 1e-9;  // m: a crossing this close to the start is a zero-length step (FP64)
constexpr double ON_FACE = 1e-9;    // m: the face a photon sits on is not met again this close
constexpr double ON_FACE_TORUS = 1e-7;  // m: the same for a torus face (its grazing roots: sqrt(eps) conditioned)
constexpr int ZERO_STEPS_BEFORE_PUSH = 10, ZERO_STEPS_BEFORE_KILL = 25;
constexpr double PUSH = 1e-7;       // m
enum { STRAIGHT = 0, REFRACT, SPECULAR, BACKSCATTER, GLOSSY, DIFFUSE, NOUT };
enum { COL_R = 0, COL_T, COL_EFF, COL_SPIKE, COL_LOBE, COL_BACK, COL_SIGMA, COL_POLISH, NCOL };

// ------------------------------------------------------------------------------------------ the scene (FP64)
struct Leaf {
  int typ;
  double q[3], A[9], b[3], c;
  std::vector<double> n;  // 3 per plane, relative to q
  std::vector<double> h;
};
struct Kind {
  std::vector<Leaf> leaves;
  std::vector<int> op, arg;
};
struct Scene {
  int n;
  std::vector<double> R, t;  // local -> world
  std::vector<int> kind, parent, cls, sensor;
  std::vector<std::vector<int>> kids;  // per navigator id, its daughters (brute force: no BVH)
  std::vector<Kind> kinds;
  double reach;
  // the navigator arrays' BVHs (geometry/bvh.py:1-10: depth first, skip links), for the BVH walk mode
  std::vector<int> vg0, vg1, g0, g1, skip, start, count, items;
  std::vector<double> gR, lo, hi;
  bool bvh = false;  // true: daughters by the BVH walk (the reference's Navigator._node); false: brute force
};

Scene scene_of(const lh_geometry *G) {
  Scene s;
  s.n = G->n_vol;
  s.R.assign(G->R, G->R + 9 * s.n);
  s.t.assign(G->t, G->t + 3 * s.n);
  s.kind.assign(G->kind, G->kind + s.n);
  s.parent.assign(G->parent, G->parent + s.n + 1);
  s.cls.assign(G->cls, G->cls + s.n + 1);
  s.sensor.assign(G->sensor, G->sensor + s.n + 1);
  s.kids.assign(s.n + 1, {});
  for (int v = 0; v < s.n; ++v) s.kids[s.parent[v]].push_back(v);
  s.kinds.resize(G->n_kinds);
  for (int k = 0; k < G->n_kinds; ++k) {
    Kind &K = s.kinds[k];
    for (int j = 0; j < G->kind_nl[k]; ++j) {
      const int l = G->kind_l0[k] + j;
      Leaf L;
      L.typ = G->leaf_typ[l];
      for (int a = 0; a < 3; ++a) L.q[a] = G->leaf_q[3 * l + a], L.b[a] = G->leaf_b[3 * l + a];
      for (int a = 0; a < 9; ++a) L.A[a] = G->leaf_A[9 * l + a];
      L.c = G->leaf_c[l];
      for (int p = 0; p < G->leaf_np[l]; ++p) {
        const int pi = G->leaf_p0[l] + p;
        const float *nn = G->plane_n + 3 * pi;
        const double o[3] = {G->leaf_o[3 * l], G->leaf_o[3 * l + 1], G->leaf_o[3 * l + 2]};
        // n.(x - o) <= h  ->  n.(x - q) <= h - n.(q - o)
        L.n.insert(L.n.end(), {nn[0], nn[1], nn[2]});
        L.h.push_back(G->plane_h[pi] - (nn[0] * (L.q[0] - o[0]) + nn[1] * (L.q[1] - o[1]) + nn[2] * (L.q[2] - o[2])));
      }
      K.leaves.push_back(L);
    }
    for (int j = 0; j < G->kind_pl[k]; ++j) {
      K.op.push_back(G->prog_op[G->kind_p0[k] + j]);
      K.arg.push_back(G->prog_arg[G->kind_p0[k] + j]);
    }
  }
  s.reach = G->reach;
  s.vg0.assign(G->vg0, G->vg0 + s.n + 1);
  s.vg1.assign(G->vg1, G->vg1 + s.n + 1);
  s.g0.assign(G->g0, G->g0 + G->n_groups);
  s.g1.assign(G->g1, G->g1 + G->n_groups);
  s.gR.assign(G->gR, G->gR + 9 * G->n_groups);
  s.skip.assign(G->skip, G->skip + G->n_nodes);
  s.start.assign(G->start, G->start + G->n_nodes);
  s.count.assign(G->count, G->count + G->n_nodes);
  s.lo.assign(G->lo, G->lo + 3 * G->n_nodes);
  s.hi.assign(G->hi, G->hi + 3 * G->n_nodes);
  s.items.assign(G->items, G->items + G->n_items);
  return s;
}

template <class R> inline void to_local(const Scene &s, int v, V3<R> x, V3<R> d, V3<R> &xl, V3<R> &dl) {
  const double *M = &s.R[9 * v], *t = &s.t[3 * v];
  const V3<R> y = x - v3<R>(R(t[0]), R(t[1]), R(t[2]));
  xl = v3<R>(R(M[0]) * y.x + R(M[3]) * y.y + R(M[6]) * y.z, R(M[1]) * y.x + R(M[4]) * y.y + R(M[7]) * y.z,
             R(M[2]) * y.x + R(M[5]) * y.y + R(M[8]) * y.z);
  dl = v3<R>(R(M[0]) * d.x + R(M[3]) * d.y + R(M[6]) * d.z, R(M[1]) * d.x + R(M[4]) * d.y + R(M[7]) * d.z,
             R(M[2]) * d.x + R(M[5]) * d.y + R(M[8]) * d.z);
}
template <class R> inline V3<R> to_world_dir(const Scene &s, int v, V3<R> n) {
  const double *M = &s.R[9 * v];
  return v3<R>(R(M[0]) * n.x + R(M[1]) * n.y + R(M[2]) * n.z, R(M[3]) * n.x + R(M[4]) * n.y + R(M[5]) * n.z,
               R(M[6]) * n.x + R(M[7]) * n.y + R(M[8]) * n.z);
}

// a leaf's intervals along y + t d, y relative to its q: each (lo, hi, face_lo, face_hi); face 0 the quadric
template <class R> struct Iv {
  R lo, hi;
  int flo, fhi;
};
// A solid torus about the unit axis w through the leaf's q (swept radius Rt = A[0], tube radius rt = A[1]): along
// y + t d (|d| = 1), from the point nearest the centre (s = t - tc), g(s) = s^2 + |p0|^2 + Rt^2 - rt^2 - 2 Rt rho(s)
// with rho(s)^2 = a s^2 + 2 b s + c; g <= 0 inside.  geometry/leaf_kernels.py:212-248 `torus_set`: the chord of the
// bounding sphere split at g's inflection points (rho^3 = Rt (a c - b^2), closed form), each piece's extremum by
// bisection on g', each monotone piece's root by bisection on the membership; the bracket's start is outside, so the
// roots alternate entry and exit.
constexpr int TORUS_BISECT_F64 = 80;
template <class R> struct TorusG {
  R Rt, a, b, c, k;
  R rho(R s) const { return SQRT(MAX<R>(a * s * s + R(2.0) * b * s + c, R(0.0))); }
  R g(R s) const { return s * s + k - R(2.0) * Rt * rho(s); }
  R dg(R s) const { return R(2.0) * s - R(2.0) * Rt * (a * s + b) / MAX<R>(rho(s), R(1e-300)); }
};
template <class R> int torus_roots(R Rt, R rt, V3<R> w, V3<R> y, V3<R> d, R roots[4]) {
  const R tc = -dot(y, d);
  const V3<R> p0 = y + tc * d;
  const R pp = dot(p0, p0), rb = Rt + rt;
  if (pp >= rb * rb) return 0;
  // the chord a little past the bounding sphere, so its ends lie strictly outside (an in-plane ray meets the
  // sphere on the torus's outer equator)
  const R sb = SQRT(rb * rb - pp) * R(1.0 + 1e-9) + R(1e-12) * rb;
  const V3<R> pq = p0 - dot(p0, w) * w, dq = d - dot(d, w) * w;
  const TorusG<R> G{Rt, dot(dq, dq), dot(pq, dq), dot(pq, pq), pp + Rt * Rt - rt * rt};
  std::vector<R> knots = {-sb};
  const R det = MAX<R>(G.a * G.c - G.b * G.b, R(0.0));
  if (G.a > R(1e-24)) {
    const R ri = CBRT(Rt * det), disc = G.b * G.b - G.a * (G.c - ri * ri);
    if (disc >= R(0.0)) {
      for (R sgn : {R(-1.0), R(1.0)}) {
        const R s = (-G.b + sgn * SQRT(disc)) / G.a;
        if (s > knots.back() && s < sb) knots.push_back(s);
      }
    }
  }
  knots.push_back(sb);
  std::vector<R> mono = {knots[0]};  // the monotone pieces' ends: each piece's extremum added
  for (size_t i = 0; i + 1 < knots.size(); ++i) {
    R lo = knots[i], hi = knots[i + 1];
    // g' just inside the piece: at a knot on the axis (rho = 0) g' jumps, and its one-sided values are the piece's
    const R in = R(1e-12) * (hi - lo);
    const bool nl = G.dg(lo + in) < R(0.0);
    if (nl != (G.dg(hi - in) < R(0.0))) {
      for (int it = 0; it < TORUS_BISECT_F64; ++it) {
        const R mid = R(0.5) * (lo + hi);
        if ((G.dg(mid) < R(0.0)) == nl) lo = mid; else hi = mid;
      }
      mono.push_back(R(0.5) * (lo + hi));
    }
    mono.push_back(knots[i + 1]);
  }
  int n = 0;
  for (size_t i = 0; i + 1 < mono.size() && n < 4; ++i) {
    R lo = mono[i], hi = mono[i + 1];
    const bool il = G.g(lo) <= R(0.0);
    if (il == (G.g(hi) <= R(0.0)) || !(hi > lo)) continue;
    for (int it = 0; it < TORUS_BISECT_F64; ++it) {
      const R mid = R(0.5) * (lo + hi);
      if ((G.g(mid) <= R(0.0)) == il) lo = mid; else hi = mid;
    }
    roots[n++] = tc + R(0.5) * (lo + hi);
  }
  return n;
}

template <class R>
int leaf_iv(const Leaf &L, V3<R> o, V3<R> d, Iv<R> out[2]) {
  const R INF = R(1e300);
  const V3<R> y = o - v3<R>(R(L.q[0]), R(L.q[1]), R(L.q[2]));
  R t0 = -INF, t1 = INF;
  int f0 = -1, f1 = -1;
  for (size_t p = 0; p < L.h.size(); ++p) {
    const V3<R> n = v3<R>(R(L.n[3 * p]), R(L.n[3 * p + 1]), R(L.n[3 * p + 2]));
    const R nd = dot(n, d), rhs = R(L.h[p]) - dot(n, y);
    if (FABS(nd) < R(1e-15)) {
      if (rhs < R(0.0)) return 0;
      continue;
    }
    const R tp = rhs / nd;
    if (nd < R(0.0)) {
      if (tp > t0) { t0 = tp; f0 = (int)p + 1; }
    } else {
      if (tp < t1) { t1 = tp; f1 = (int)p + 1; }
    }
  }
  R qa[2] = {-INF, INF}, qb[2] = {INF, -INF};
  bool ra[2] = {false, false}, rb[2] = {false, false};
  if (L.typ == 2) {
    R rts[4];
    const int n = torus_roots<R>(R(L.A[0]), R(L.A[1]), v3<R>(R(L.b[0]), R(L.b[1]), R(L.b[2])), y, d, rts);
    qa[0] = INF, qb[0] = -INF;
    if (n >= 2) { qa[0] = rts[0]; qb[0] = rts[1]; ra[0] = rb[0] = true; }
    if (n >= 4) { qa[1] = rts[2]; qb[1] = rts[3]; ra[1] = rb[1] = true; }
  } else if (L.typ == 1) {
    auto Av = [&](V3<R> v) {
      return v3<R>(R(L.A[0]) * v.x + R(L.A[1]) * v.y + R(L.A[2]) * v.z, R(L.A[3]) * v.x + R(L.A[4]) * v.y + R(L.A[5]) * v.z,
                   R(L.A[6]) * v.x + R(L.A[7]) * v.y + R(L.A[8]) * v.z);
    };
    const V3<R> bb = v3<R>(R(L.b[0]), R(L.b[1]), R(L.b[2]));
    const R a = dot(d, Av(d));
    const R beta = dot(Av(y), d) + dot(bb, d);
    const R gam = dot(y, Av(y)) + R(2.0) * dot(bb, y) + R(L.c);
    double amax = 0;
    for (int i = 0; i < 9; ++i) amax = std::max(amax, std::fabs(L.A[i]));
    if (FABS(a) <= R(1e-13 * amax)) {
      if (beta > R(0.0)) { qa[0] = -INF; qb[0] = -gam / (R(2.0) * beta); rb[0] = true; }
      else if (beta < R(0.0)) { qa[0] = -gam / (R(2.0) * beta); qb[0] = INF; ra[0] = true; }
      else if (gam > R(0.0)) { qa[0] = INF; qb[0] = -INF; }
    } else {
      const R D = beta * beta - a * gam;
      if (D < R(0.0)) {
        if (a > R(0.0)) { qa[0] = INF; qb[0] = -INF; }
      } else {
        const R sq = SQRT(D);
        const R qq = beta > R(0.0) ? -(beta + sq) : -(beta - sq);
        R r1 = qq / a, r2 = qq != R(0.0) ? gam / qq : r1;
        if (r2 < r1) std::swap(r1, r2);
        if (a > R(0.0)) { qa[0] = r1; qb[0] = r2; ra[0] = rb[0] = true; }
        else { qa[0] = -INF; qb[0] = r1; rb[0] = true; qa[1] = r2; qb[1] = INF; ra[1] = true; }
      }
    }
  }
  int k = 0;
  for (int p = 0; p < 2; ++p) {
    const R lo = MAX<R>(qa[p], t0), hi = MIN<R>(qb[p], t1);
    if (lo < hi) {
      out[k].lo = lo;
      out[k].hi = hi;
      out[k].flo = (qa[p] > t0 && ra[p]) ? 0 : f0;
      out[k].fhi = (qb[p] < t1 && rb[p]) ? 0 : f1;
      ++k;
    }
  }
  return k;
}

inline bool run_prog(const Kind &K, const std::vector<bool> &in) {
  std::vector<bool> st;
  for (size_t j = 0; j < K.op.size(); ++j) {
    if (K.op[j] == 0) {
      st.push_back(in[K.arg[j]]);
    } else {
      const bool b = st.back();
      st.pop_back();
      const bool a = st.back();
      st.pop_back();
      st.push_back(K.op[j] == 1 ? (a || b) : K.op[j] == 2 ? (a && b) : (a && !b));
    }
  }
  return st.back();
}

template <class R> struct Change {
  R t;
  int leaf, face;
};

// the first change of a solid's membership along the ray after t = 0 (compare/oracle.py: FP64 leaf roots, the CSG
// value scanned between consecutive endpoints); `exit_query`: from inside, else from outside.  Endpoints of the face
// the photon sits on (volume `vol`, code `skip`) within ON_FACE of the start count as passed.
template <class R>
Change<R> first_change(const Scene &s, int vol, V3<R> o, V3<R> d, bool exit_query, long long skip, double on_face_f32) {
  const Kind &K = s.kinds[s.kind[vol]];
  const int nl = (int)K.leaves.size();
  std::vector<std::array<Iv<R>, 2>> ivs(nl);
  std::vector<int> cnt(nl);
  for (int l = 0; l < nl; ++l) {
    cnt[l] = leaf_iv(K.leaves[l], o, d, ivs[l].data());
    // a torus root met at a grazing angle is conditioned to sqrt(eps) of the tube, not eps: its face is passed
    // within ON_FACE_TORUS
    // in float32 (the engine's arithmetic) the engine's band rule: K_ON bands at the crossing ray's scale
    const R on_face = sizeof(R) == 4 ? R(std::max(on_face_f32, K.leaves[l].typ == 2 ? ON_FACE_TORUS : 0.0))
                                     : R(K.leaves[l].typ == 2 ? ON_FACE_TORUS : ON_FACE);
    for (int p = 0; p < cnt[l]; ++p) {
      Iv<R> &iv = ivs[l][p];
      const long long clo = ((long long)vol << 10) | (l << 5) | iv.flo, chi = ((long long)vol << 10) | (l << 5) | iv.fhi;
      if (clo == skip && FABS(iv.lo) <= on_face) iv.lo = R(-1e299);
      if (chi == skip && FABS(iv.hi) <= on_face) iv.hi = R(-1e299);
    }
  }
  auto member = [&](R t) {
    std::vector<bool> in(nl);
    for (int l = 0; l < nl; ++l) {
      bool b = false;
      for (int p = 0; p < cnt[l]; ++p) b = b || (ivs[l][p].lo <= t && t < ivs[l][p].hi);
      in[l] = b;
    }
    return run_prog(K, in);
  };
  const bool start = member(R(0.0));
  if (exit_query && !start) return {R(0.0), 0, 0};
  if (!exit_query && start) return {R(1e300), -1, -1};
  R cur = R(0.0);
  for (int it = 0; it < 4 * nl + 5; ++it) {
    R nxt = R(1e300);
    int bl = -1, bf = -1;
    for (int l = 0; l < nl; ++l)
      for (int p = 0; p < cnt[l]; ++p) {
        const Iv<R> &iv = ivs[l][p];
        if (iv.lo > cur && iv.lo < nxt) { nxt = iv.lo; bl = l; bf = iv.flo; }
        if (iv.hi > cur && iv.hi < nxt && iv.hi < R(1e299)) { nxt = iv.hi; bl = l; bf = iv.fhi; }
      }
    if (bl < 0) break;
    // a change only where the value stays changed a little further on (touching leaves' rounding gaps; FP64)
    if (member(nxt) != start && member(nxt + R(1e-12) * (R(1.0) + FABS(nxt))) != start) return {nxt, bl, bf};
    cur = nxt;
  }
  if (exit_query) return {R(0.0), 0, 0};
  return {R(1e300), -1, -1};
}

template <class R> struct Cross {
  R t;
  int vol_after;
  long long srf;
  V3<R> n;
};

template <class R> V3<R> normal_at(const Scene &s, int v, int leaf, int face, V3<R> x, V3<R> d, R t) {
  V3<R> xl, dl;
  to_local(s, v, x, d, xl, dl);
  const Leaf &L = s.kinds[s.kind[v]].leaves[leaf];
  V3<R> nl;
  if (face == 0 && L.typ == 2) {  // a torus: y less the swept circle's point nearest it
    const V3<R> y = (xl + t * dl) - v3<R>(R(L.q[0]), R(L.q[1]), R(L.q[2]));
    const V3<R> w = v3<R>(R(L.b[0]), R(L.b[1]), R(L.b[2]));
    const V3<R> q = y - dot(y, w) * w;
    nl = y - (R(L.A[0]) / MAX<R>(SQRT(dot(q, q)), R(1e-300))) * q;
  } else if (face == 0) {
    const V3<R> y = (xl + t * dl) - v3<R>(R(L.q[0]), R(L.q[1]), R(L.q[2]));
    nl = v3<R>(R(L.A[0]) * y.x + R(L.A[1]) * y.y + R(L.A[2]) * y.z + R(L.b[0]),
               R(L.A[3]) * y.x + R(L.A[4]) * y.y + R(L.A[5]) * y.z + R(L.b[1]),
               R(L.A[6]) * y.x + R(L.A[7]) * y.y + R(L.A[8]) * y.z + R(L.b[2]));
  } else {
    nl = v3<R>(R(L.n[3 * (face - 1)]), R(L.n[3 * (face - 1) + 1]), R(L.n[3 * (face - 1) + 2]));
  }
  V3<R> n = unit(to_world_dir(s, v, nl));
  if (dot(n, d) > R(0.0)) n = R(-1.0) * n;
  return n;
}

// the next visible crossing: the own exit, then every daughter (a daughter wins a tie)
inline int locate_from(const Scene &s, int cur, V3<double> x);

// `scale`: the coordinate scale of the ray that put the start on face `last` (float32: its band bounds the start's
// rounding off that face)
template <class R>
Cross<R> query(const Scene &s, V3<R> x, V3<R> d, int vol, long long last, R tmax, double scale = 0.0) {
  const double xa = std::max(std::max(std::fabs((double)val(x.x)), std::fabs((double)val(x.y))),
                             std::fabs((double)val(x.z)));
  const double on_face_f32 = 4.0 * (8.0 * 1.1920928955078125e-7 * std::max(xa, scale) + 1e-9);
  Cross<R> best{tmax, -1, -1, v3<R>(R(0.0), R(0.0), R(1.0))};
  int bv = -1, bl = -1, bf = -1;
  bool own = false;
  if (vol < s.n) {
    V3<R> xl, dl;
    to_local(s, vol, x, d, xl, dl);
    const Change<R> ch = first_change(s, vol, xl, dl, true, last, on_face_f32);
    if (ch.leaf >= 0 && ch.t <= best.t) {
      best.t = ch.t; best.vol_after = s.parent[vol]; bv = vol; bl = ch.leaf; bf = ch.face; own = true;
    }
  }
  auto visit = [&](int u) {
    V3<R> xl, dl;
    to_local(s, u, x, d, xl, dl);
    const Change<R> ch = first_change(s, u, xl, dl, false, last, on_face_f32);
    if (ch.leaf >= 0 && (ch.t < best.t || (ch.t == best.t && own))) {
      best.t = ch.t; best.vol_after = u; bv = u; bl = ch.leaf; bf = ch.face; own = false;
    }
  };
  if (!s.bvh) {
    for (int u : s.kids[vol]) visit(u);
  } else {  // each frame group's BVH, stackless: a node the ray's window misses is skipped with its subtree
    for (int gi = s.vg0[vol]; gi < s.vg1[vol]; ++gi) {
      const double *M = &s.gR[9 * gi];
      const V3<R> xg = v3<R>(R(M[0]) * x.x + R(M[3]) * x.y + R(M[6]) * x.z, R(M[1]) * x.x + R(M[4]) * x.y + R(M[7]) * x.z,
                             R(M[2]) * x.x + R(M[5]) * x.y + R(M[8]) * x.z);
      const V3<R> dg = v3<R>(R(M[0]) * d.x + R(M[3]) * d.y + R(M[6]) * d.z, R(M[1]) * d.x + R(M[4]) * d.y + R(M[7]) * d.z,
                             R(M[2]) * d.x + R(M[5]) * d.y + R(M[8]) * d.z);
      const R ix = R(1.0) / (FABS(dg.x) > R(1e-30) ? dg.x : R(1e-30)), iy = R(1.0) / (FABS(dg.y) > R(1e-30) ? dg.y : R(1e-30)),
              iz = R(1.0) / (FABS(dg.z) > R(1e-30) ? dg.z : R(1e-30));
      int ni = s.g0[gi];
      while (ni < s.g1[gi]) {
        const R tx0 = (R(s.lo[3 * ni]) - xg.x) * ix, tx1 = (R(s.hi[3 * ni]) - xg.x) * ix;
        const R ty0 = (R(s.lo[3 * ni + 1]) - xg.y) * iy, ty1 = (R(s.hi[3 * ni + 1]) - xg.y) * iy;
        const R tz0 = (R(s.lo[3 * ni + 2]) - xg.z) * iz, tz1 = (R(s.hi[3 * ni + 2]) - xg.z) * iz;
        const R te = MAX<R>(MAX<R>(MIN<R>(tx0, tx1), MIN<R>(ty0, ty1)), MAX<R>(MIN<R>(tz0, tz1), R(0.0)));
        const R tl = MIN<R>(MIN<R>(MAX<R>(tx0, tx1), MAX<R>(ty0, ty1)), MIN<R>(MAX<R>(tz0, tz1), best.t));
        if (te > tl) {
          ni = s.skip[ni];
          continue;
        }
        if (s.count[ni] > 0) {
          for (int k = 0; k < s.count[ni]; ++k) visit(s.items[s.start[ni] + k]);
          ni = s.skip[ni];
        } else {
          ni += 1;
        }
      }
    }
  }
  if (bv >= 0) {
    best.srf = ((long long)bv << 10) | (bl << 5) | bf;
    // an own exit at t = 0 is a start outside its own solid (rounding): no face is met, the normal faces the photon
    best.n = (own && best.t == R(0.0)) ? R(-1.0) * d : normal_at(s, bv, bl, bf, x, d, best.t);
  }
  // an exit into the mother at a point another volume holds (touching siblings, a daughter flush with its mother's
  // face, whose exit is the mother's too): the crossing goes on into the volume holding the point a little past the
  // face (src/lumen/geometry/navigate.py:526-555: K_ON float32 bands at the crossing's scale, so faces closer than
  // that are crossed as one, in FP64 as in float32), the crossing point moved there
  // (an entry whose point past the face a daughter of the entered volume holds goes on into that daughter alike)
  if (best.vol_after >= 0 && best.vol_after < s.n) {
    const double step = 4.0 * (8.0 * 1.1920928955078125e-7 * (xa + (double)val(best.t)) + 1e-9);
    const V3<double> xp = v3(val(x.x), val(x.y), val(x.z)) + ((double)val(best.t) + step) * v3(val(d.x), val(d.y), val(d.z));
    const int v2 = locate_from(s, own ? s.n : best.vol_after, xp);
    if (v2 != vol && v2 != best.vol_after) {
      best.vol_after = v2;
      best.t = best.t + R(step);
      best.srf = -1;
    }
  }
  return best;
}

inline bool inside(const Scene &s, int v, V3<double> x) {
  V3<double> xl, dl;
  to_local<double>(s, v, x, v3<double>(0, 0, 1), xl, dl);
  const Kind &K = s.kinds[s.kind[v]];
  std::vector<bool> in(K.leaves.size());
  for (size_t l = 0; l < K.leaves.size(); ++l) {
    const Leaf &L = K.leaves[l];
    const double y[3] = {xl.x - L.q[0], xl.y - L.q[1], xl.z - L.q[2]};
    bool b = true;
    for (size_t p = 0; p < L.h.size(); ++p)
      b = b && (L.n[3 * p] * y[0] + L.n[3 * p + 1] * y[1] + L.n[3 * p + 2] * y[2] <= L.h[p]);
    if (L.typ == 1) {
      double q = L.c;
      for (int i = 0; i < 3; ++i) {
        q += 2.0 * L.b[i] * y[i];
        for (int j = 0; j < 3; ++j) q += y[i] * L.A[3 * i + j] * y[j];
      }
      b = b && q <= 0.0;
    } else if (L.typ == 2) {
      const double yy = y[0] * y[0] + y[1] * y[1] + y[2] * y[2], z = y[0] * L.b[0] + y[1] * L.b[1] + y[2] * L.b[2];
      b = b && yy - 2.0 * L.A[0] * std::sqrt(std::max(yy - z * z, 0.0)) + L.A[0] * L.A[0] - L.A[1] * L.A[1] <= 0.0;
    }
    in[l] = b;
  }
  return run_prog(K, in);
}

inline int locate_from(const Scene &s, int cur, V3<double> x) {
  for (;;) {
    int found = -1;
    for (int u : s.kids[cur])
      if (inside(s, u, x)) { found = u; break; }
    if (found < 0) return cur;
    cur = found;
  }
}

inline int locate(const Scene &s, V3<double> x) { return locate_from(s, s.n, x); }

// ------------------------------------------------------------------------------------------ optics (FP64)
struct Optics {
  const int32_t *L;
  const float *P;
  double tab(int off, int n, double x, int *i_out = nullptr, double *f_out = nullptr) const {
    int i = 0;
    double f = 0.0;
    if (n > 1 && x > P[off]) {
      if (x >= P[off + n - 1]) {
        i = n - 1;
      } else {
        while (i < n - 2 && P[off + i + 1] <= x) ++i;
        f = (x - P[off + i]) / ((double)P[off + i + 1] - P[off + i]);
      }
    }
    if (i_out) *i_out = i, *f_out = f;
    const double v0 = P[off + n + i];
    return f == 0.0 ? v0 : v0 + f * ((double)P[off + n + i + 1] - v0);
  }
  int cls_medium(int c) const { return c < 0 ? -1 : L[L[7] + c]; }
  int row_of(int a, int b) const { return L[L[8] + a * L[2] + b]; }
  const int32_t *medium(int m) const { return L + L[6] + 25 * m; }
  const int32_t *row(int r) const { return L + L[9] + 18 * r; }
};

inline double rate(double len) { return len >= TINY_NORMAL ? 1.0 / len : 1.0 / TINY_NORMAL; }

template <class R> struct Med {
  R n, ng, mu_abs, mu_ray, mu_mie, mu_wls;  // the channels in Geant4's order: Rayleigh, Mie, WLS
};

template <class R> Med<R> medium_at(const Optics &o, int m, double lam) {
  Med<R> r{R(1.0), R(1.0), R(0.0), R(0.0), R(0.0), R(0.0)};
  if (m < 0) return r;
  const int32_t *M = o.medium(m);
  if (M[8] & 1) return r;
  const double E = HC / lam;
  r.n = R(o.tab(M[0], M[1], E));
  int i;
  double f;
  o.tab(M[2], M[3], E, &i, &f);  // the group velocity (1 / n_g) linear in energy between GROUPVEL knots
  const R g0 = R(1.0) / R(o.P[M[2] + M[3] + i]);
  const R inv = f == 0.0 ? g0 : g0 + R(f) * (R(1.0) / R(o.P[M[2] + M[3] + i + 1]) - g0);
  r.ng = R(1.0) / inv;
  if (M[5] > 0) r.mu_abs = R(rate(o.tab(M[4], M[5], E)));
  if (M[7] > 0) r.mu_ray = R(rate(o.tab(M[6], M[7], E)));
  if (M[8] & 8) r.mu_mie = R(rate(o.tab(M[17], M[18], E)));
  if (M[8] & 2) r.mu_wls = R(rate(o.tab(M[9], M[10], E)));
  return r;
}

// G4OpWLS's emission (src/lumen/media/emission.py:120-159 `EmissionRecord.sample`; the sampler of
// src/lumen/media/spectra.py:37-137: g4_arrays, covered, cumulative, invert, on_intervals, kept), FP64
struct Spec {
  std::vector<double> e, d, c;  // interval starts+end (eV), densities (1/eV), CDF at the starts
};
inline Spec spec_of(const float *E, const float *S, int n, double cover) {
  Spec q;
  q.e.assign(E, E + n);
  double tot = 0.0, zero = 0.0;
  for (int k = 0; k + 1 < n; ++k) {
    const double h = (double)E[k + 1] - E[k];
    tot += 0.5 * ((double)S[k] + S[k + 1]) * h;
    if ((double)S[k] + S[k + 1] <= 0.0 && h > 0.0) zero += h;
  }
  double c = 0.0;
  for (int k = 0; k + 1 < n; ++k) {
    const double h = (double)E[k + 1] - E[k], m = (double)S[k] + S[k + 1];
    double d = tot > 0.0 ? 0.5 * m / tot : 0.0;
    if (zero > 0.0) d = (1.0 - cover) * d + ((m <= 0.0 && h > 0.0) ? cover / zero : 0.0);
    q.d.push_back(d);
    q.c.push_back(c);
    c += d * h;
  }
  return q;
}
inline int spec_interval(const Spec &q, double e) {  // the last interval starting at or below e (clipped)
  int k = 0;
  for (int j = 0; j + 1 < (int)q.e.size(); ++j)
    if (q.e[j] <= e) k = j;
  return k;
}
inline double spec_cdf(const Spec &q, double e) {
  if (e < q.e[0]) return 0.0;
  const int k = spec_interval(q, e);
  const double u = q.c[k] + q.d[k] * std::min(std::max(e - q.e[k], 0.0), q.e[k + 1] - q.e[k]);
  return std::min(std::max(u, 0.0), 1.0);
}
inline double spec_invert(const Spec &q, double r) {
  int k = 0;
  for (int j = 0; j < (int)q.c.size(); ++j)
    if (q.c[j] <= r) k = j;
  const double x = q.d[k] > 0.0 ? std::max(r - q.c[k], 0.0) / q.d[k] : 0.0;
  return q.e[k] + std::min(std::max(x, 0.0), q.e[k + 1] - q.e[k]);
}
inline double spec_density(const Spec &q, double e) {
  if (!(e >= q.e.front() && e <= q.e.back())) return 0.0;
  return q.d[spec_interval(q, e)];
}
inline double wls_kept(double f) {
  return -std::expm1(100.0 * std::log1p(-std::min(std::max(f, 0.0), 1.0 - 1.1920928955078125e-07)));
}
inline double safe_ratio(double a, double b) { return b != 0.0 ? a / b : 0.0; }
struct WlsDraw {
  double lam, delay, w, wr;
};
inline WlsDraw wls_emit(const Optics &live, const Optics &ref, const int32_t *M, double lam_in, double u_lam,
                        double u_time) {
  const int off = M[11], n = M[12];
  const Spec q = spec_of(ref.P + off, ref.P + off + n, n, DEFENSIVE);
  const Spec l = spec_of(live.P + off, live.P + off + n, n, 0.0);
  const double e_in = HC / lam_in;
  const double f_q = spec_cdf(q, e_in);
  const double e = spec_invert(q, u_lam * f_q);
  const double qd = spec_density(q, e), dens_q = safe_ratio(qd, f_q);
  const double f_l = spec_cdf(l, e_in);
  const double y = live.P[M[13] + M[14]], y_r = ref.P[M[13] + M[14]], tau = live.P[M[15] + M[16]];
  WlsDraw r;
  r.lam = HC / e;
  r.w = safe_ratio(y * wls_kept(f_l) * safe_ratio(spec_density(l, e), f_l), dens_q);
  r.wr = safe_ratio(y_r * wls_kept(f_q) * safe_ratio(qd, f_q), dens_q);
  r.delay = (M[8] & 4) ? -tau * std::log(u_time) : tau;
  return r;
}

inline double index_at(const Optics &o, int m, double lam) {
  if (m < 0) return 1.0;
  const int32_t *M = o.medium(m);
  if (M[8] & 1) return 1.0;
  return o.tab(M[0], M[1], HC / lam);
}

template <class R> void basis(V3<R> n, V3<R> &t1, V3<R> &t2) {
  const R sign = n.z >= R(0.0) ? R(1.0) : R(-1.0);
  const R a = R(-1.0) / (sign + n.z);
  const R b = n.x * n.y * a;
  t1 = v3<R>(R(1.0) + sign * n.x * n.x * a, sign * b, -sign * n.x);
  t2 = v3<R>(b, sign + n.y * n.y * a, -n.y);
}
template <class R> V3<R> to_world(V3<R> l, V3<R> n) {
  V3<R> t1, t2;
  basis(n, t1, t2);
  return l.x * t1 + l.y * t2 + l.z * n;
}

struct C2 {  // complex in double (Fresnel, Jones)
  double re, im;
};
inline C2 cm(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C2 cd(C2 a, C2 b) {
  const double d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
inline double ca2(C2 a) { return a.re * a.re + a.im * a.im; }

struct Fres {
  C2 rs, rp, ts, tp;
};
inline Fres fresnel_amps(double c, double n1, double n2) {
  c = std::min(std::max(c, 0.0), 1.0);
  const double z = n2 * n2 - n1 * n1 * (1 - c * c);
  C2 q = z >= 0 ? C2{std::sqrt(z), 0.0} : C2{0.0, std::sqrt(-z)};  // Im q >= 0: the decaying wave
  if (std::fabs(z) < 1e-24) q = {0, 0};
  const C2 ds{n1 * c + q.re, q.im}, dp{n2 * n2 * c + n1 * q.re, n1 * q.im};
  Fres f;
  const bool oks = std::sqrt(ca2(ds)) > 1e-12, okp = std::sqrt(ca2(dp)) > 1e-12;
  f.rs = oks ? cd({n1 * c - q.re, -q.im}, ds) : C2{-1, 0};
  f.rp = okp ? cd({n2 * n2 * c - n1 * q.re, -n1 * q.im}, dp) : C2{1, 0};
  f.ts = oks ? cd({2 * n1 * c, 0}, ds) : C2{0, 0};
  f.tp = okp ? cd({2 * n1 * n2 * c, 0}, dp) : C2{0, 0};
  return f;
}

// Fresnel amplitudes from real n1 into a complex index nr + i nk (src/lumen/surfaces/fresnel.py:25-45 `amplitudes`
// with a complex n2; the photocathode model src/lumen/surfaces/types.py:262-274 `_complex`)
inline C2 csqrt_decaying(C2 z) {
  const double m = std::sqrt(z.re * z.re + z.im * z.im);
  if (m < 1e-24) return {0.0, 0.0};
  double re = std::sqrt(std::max(0.5 * (m + z.re), 0.0)), im = std::sqrt(std::max(0.5 * (m - z.re), 0.0));
  if (z.im < 0.0) im = -im;
  if (im < 0.0) re = -re, im = -im;
  return {re, im};
}
inline Fres fresnel_amps_complex(double c, double n1, double nr, double nk) {
  c = std::min(std::max(c, 0.0), 1.0);
  const C2 n2{nr, nk}, n2sq = cm(n2, n2);
  const C2 q = csqrt_decaying({n2sq.re - n1 * n1 * (1.0 - c * c), n2sq.im});
  const C2 ds{n1 * c + q.re, q.im}, dp{n2sq.re * c + n1 * q.re, n2sq.im * c + n1 * q.im};
  Fres f;
  const bool oks = std::sqrt(ca2(ds)) > 1e-12, okp = std::sqrt(ca2(dp)) > 1e-12;
  f.rs = oks ? cd({n1 * c - q.re, -q.im}, ds) : C2{-1, 0};
  f.rp = okp ? cd({n2sq.re * c - n1 * q.re, n2sq.im * c - n1 * q.im}, dp) : C2{1, 0};
  f.ts = oks ? cd({2 * n1 * c, 0}, ds) : C2{0, 0};
  f.tp = okp ? cd(cm({2 * n1 * c, 0}, n2), dp) : C2{0, 0};
  return f;
}

template <class R> struct Surf {
  R p[NOUT];
  R detect, sigma;
  Fres amps;
  int flags;
};

template <class R> V3<R> s_axis(V3<R> wi, V3<R> n) {
  const V3<R> c = cross(wi, n);
  if (dot(c, c) > R(1e-28)) return unit(c);
  V3<R> t1, t2;
  basis(wi, t1, t2);
  return t1;
}

template <class R>
Surf<R> surface_eval(const Optics &o, int row, double lam, V3<R> wi, V3<R> n, double n1, double n2, V3<R> er, V3<R> ei) {
  Surf<R> s;
  const int32_t *Rw = o.row(row);
  s.flags = Rw[1];
  const double E = HC / lam;
  auto col = [&](int c) { return R(o.tab(Rw[2 + 2 * c], Rw[3 + 2 * c], E)); };
  R c1 = CLIP01<R>(col(COL_R)), c2 = CLIP01<R>(col(COL_R) + col(COL_T));
  R r = c1, t = c2 - c1, a = R(1.0) - c2;
  s.sigma = (Rw[1] & 16) ? R(1.0) - col(COL_POLISH) : col(COL_SIGMA);  // GLISUR: the facet radius 1 - polish
  for (int k = 0; k < NOUT; ++k) s.p[k] = R(0.0);
  s.amps = Fres{{-1, 0}, {1, 0}, {1, 0}, {1, 0}};
  if (Rw[0] == 0) {
    const R ps = col(COL_SPIKE), pl = col(COL_LOBE), pb = col(COL_BACK);
    const R k1 = CLIP01<R>(ps), k2 = CLIP01<R>(ps + pl), k3 = CLIP01<R>(ps + pl + pb);
    const bool normal = Rw[1] & 1;
    s.p[SPECULAR] = r * (normal ? k2 : k1);
    s.p[GLOSSY] = normal ? R(0.0) : r * (k2 - k1);
    s.p[BACKSCATTER] = r * (k3 - k2);
    s.p[DIFFUSE] = r * (R(1.0) - k3);
  } else {
    const bool cx = Rw[0] == 2;  // a complex index (nr, nk in the p_lobe, p_back slots): its reflectance is the cut
    s.amps = cx ? fresnel_amps_complex(val(-dot(wi, n)), n1, val(col(COL_LOBE)), val(col(COL_BACK)))
                : fresnel_amps(val(-dot(wi, n)), n1, n2);
    const V3<R> sx = s_axis(wi, n), px = cross(wi, sx);
    const R es2 = dot(er, sx) * dot(er, sx) + dot(ei, sx) * dot(ei, sx);
    const R ep2 = dot(er, px) * dot(er, px) + dot(ei, px) * dot(ei, px);
    const R F = CLIP01<R>(R(ca2(s.amps.rs)) * es2 + R(ca2(s.amps.rp)) * ep2);
    if (cx) {
      c1 = F;
      c2 = CLIP01<R>(F + col(COL_T));
      r = c1;
      t = c2 - c1;
      a = R(1.0) - c2;
      s.p[SPECULAR] = r;
    } else {
      s.p[SPECULAR] = r * F;
      s.p[REFRACT] = r * (R(1.0) - F);
    }
  }
  s.p[STRAIGHT] = t;
  if (Rw[1] & 2) s.p[STRAIGHT] = s.p[REFRACT] = R(0.0);
  s.detect = a * col(COL_EFF);
  if (Rw[1] & 8) {  // per normal (src/lumen/surfaces/types.py:576-592): over A_n, the absorbed share at normal incidence
    R an = a;
    if (Rw[0] == 2) {
      const Fres f0 = fresnel_amps_complex(1.0, n1, val(col(COL_LOBE)), val(col(COL_BACK)));
      an = R(1.0) - CLIP01<R>(R(ca2(f0.rs)) + col(COL_T));
    }
    s.detect = an > R(0.0) ? s.detect / an : R(0.0);
  }
  return s;
}

template <class R> int choose(const R *q, int n, double u) {
  R tot = R(0.0);
  for (int k = 0; k < n; ++k) tot += q[k];
  if (!(tot > R(0.0))) return -1;
  R c = R(0.0);
  for (int k = 0; k < n; ++k) {
    c += q[k];
    if (c > R(u) * tot && q[k] > R(0.0)) return k;
  }
  for (int k = n - 1; k >= 0; --k)
    if (q[k] > R(0.0)) return k;
  return -1;
}

template <class R> void pol_normalise(V3<R> &er, V3<R> &ei) {
  const R s = dot(er, er) + dot(ei, ei);
  if (s > R(FLOOR32)) {
    const R r = R(1.0) / SQRT(s);
    er = r * er;
    ei = r * ei;
  } else {
    er = ei = v3<R>(R(0.0), R(0.0), R(0.0));
  }
}
template <class R> void pol_mirror(V3<R> &er, V3<R> &ei, V3<R> h) {
  er = R(2.0) * dot(er, h) * h - er;
  ei = R(2.0) * dot(ei, h) * h - ei;
}
template <class R> void pol_recombine(V3<R> &er, V3<R> &ei, V3<R> wi, V3<R> s, C2 as, C2 ap, V3<R> k) {
  const V3<R> pin = cross(wi, s);
  const C2 es{val(dot(er, s)), val(dot(ei, s))}, ep{val(dot(er, pin)), val(dot(ei, pin))};
  const C2 cs = cm(as, es), cp = cm(ap, ep);
  const V3<R> p = cross(k, s);
  er = R(cs.re) * s + R(cp.re) * p;
  ei = R(cs.im) * s + R(cp.im) * p;
  pol_normalise(er, ei);
}

// ------------------------------------------------------------------------------------------ the run
struct Run {
  const Scene *s;
  Optics live, ref;
  bool has_ref;
  const lh_source *src;
  int nsrc;
  std::vector<long long> ends;
  uint32_t k0, k1;
  double roulette, max_dist;
  int max_vert;
  int S, K, Kl, Kt;  // sensors; charge cells per sensor = Kl wavelength (QE) knots x Kt theta (CE) knots
  long long stride;
  int nt;  // the time histogram's nodes t0 + j dt (0: none), its cells from time_at of each event's tally
  bool jump;  // the diffusion jump where a medium has a grid
  bool keep_zero;  // photons without live weight walk on
  double t0, dt;
  long long time_at;
};

// the diffusion jump's clearance grids by medium (oracle_set_jump): nodes lo + h i, bounds clear (float32)
struct JumpGridO {
  double lo[3], h;
  int n[3];
  std::vector<float> clear;
};
static std::vector<JumpGridO> g_jump;
inline double jump_clearance(int medium, V3<double> x) {  // the nearest node's bound less the distance to it
  if (medium < 0 || medium >= (int)g_jump.size() || g_jump[medium].clear.empty()) return 0.0;
  const JumpGridO &G = g_jump[medium];
  const double f[3] = {(x.x - G.lo[0]) / G.h, (x.y - G.lo[1]) / G.h, (x.z - G.lo[2]) / G.h};
  int i[3];
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!(f[a] > -0.5 && f[a] < G.n[a] - 0.5)) return 0.0;
    i[a] = std::min(std::max((int)std::floor(f[a] + 0.5), 0), G.n[a] - 1);
    const double dd = (a == 0 ? x.x : a == 1 ? x.y : x.z) - (G.lo[a] + G.h * i[a]);
    d2 += dd * dd;
  }
  return (double)G.clear[((size_t)i[0] * G.n[1] + i[1]) * G.n[2] + i[2]] - std::sqrt(d2);
}
template <class R> inline double jump_clearance(int medium, V3<R> x) {
  return jump_clearance(medium, v3(val(x.x), val(x.y), val(x.z)));
}

struct Tally {
  std::vector<double> t;   // n_events * stride
  double h[LH_HEALTH_N] = {0};
};

template <class R> struct Photon {
  V3<R> x, d, er, ei;
  R t, w, wr;
  double lam, w0;
  int vol, cls, nz, vertex, event;
  long long srf;
  double xs;  // the scale |x0| + t of the ray whose crossing put the photon on face srf
  Rng rng;
};

template <class R> V3<R> facet_path(const Photon<R> &p, double sigma, V3<R> wi, V3<R> n, V3<R> &er, V3<R> &ei) {
  V3<R> t1, t2;
  basis(n, t1, t2);
  const V3<R> mwi = R(-1.0) * wi;
  V3<R> V = v3<R>(dot(mwi, t1), dot(mwi, t2), dot(mwi, n));
  V3<R> w = wi;
  const double fmax = std::min(1.0, 4.0 * sigma);
  auto arc = [&](R a) {
    const R sv = SQRT(V.x * V.x + V.y * V.y);
    const R ratio = V.z * COS(a) / MAX<R>(sv * SIN(a), R(FLOOR32));
    return ACOS(-MIN<R>(ratio, R(1.0)));
  };
  for (int k = 0; k < 256; ++k) {
    R a = R(0.0);
    double u2 = 0.5;
    for (int t = 0; t < 4096; ++t) {
      double u[4];
      p.rng.draw4(p.vertex, FACET_BLOCK + ((uint32_t)k << 12) + (uint32_t)t, u);
      const R al = R(sigma) * SQRT(R(-2.0) * LOG(R(u[0])));
      if (al < R(HALF_PI) && R(u[1]) * al * R(PI) < MIN<R>(SIN(al), R(fmax)) * arc(al)) {
        a = al;
        u2 = u[2];
        break;
      }
    }
    const R psi = (R(2.0 * u2) - R(1.0)) * arc(a);
    const R ph = ATAN2(V.y, V.x) + psi;
    const V3<R> m = v3<R>(SIN(a) * COS(ph), SIN(a) * SIN(ph), COS(a));
    const V3<R> Lo = R(2.0) * dot(V, m) * m - V;
    const V3<R> wo = to_world(Lo, n);
    pol_mirror(er, ei, unit(wo - w));
    w = wo;
    V = R(-1.0) * Lo;
    if (Lo.z >= R(0.0)) break;
  }
  return w;
}

// ------------------------------------------------------------------------------------------ GLISUR facet paths
// src/lumen/surfaces/facets.py:14-40, :226-300, :341-365, :431-530 (GetFacetNormal glisur; DielectricMetal and
// DielectricDielectric's facet loops, G4OpBoundaryProcess 11.3.0).  Blocks: a facet's attempts FACET_BLOCK + (k << 12)
// + t (the ball point), its own GLISUR_BLOCK + k (u0: the cover's shell choice, u1: the Fresnel choice)
constexpr uint32_t GLISUR_BLOCK = 0x400000u;
constexpr double INDEX_MATCH = 128.0 * 1.1920928955078125e-7;  // src/lumen/rules/choices.py:98

// M = z + r s with s uniform in the unit ball (on the unit sphere when `shell`), redrawn until V . M > 0; unit
inline V3<double> glisur_facet(const Rng &g, uint32_t vertex, int k, double r, V3<double> V, bool shell) {
  V3<double> M = v3(0.0, 0.0, 1.0);
  for (int t = 0; t < 4096; ++t) {
    double u[4];
    g.draw4(vertex, FACET_BLOCK + ((uint32_t)k << 12) + (uint32_t)t, u);
    const double z = 1.0 - 2.0 * u[1], ph = TWO_PI * u[2], sxy = std::sqrt(std::max(1.0 - z * z, 0.0));
    const double rad = r * (shell ? 1.0 : std::cbrt(u[0]));
    const V3<double> c = v3(rad * sxy * std::cos(ph), rad * sxy * std::sin(ph), 1.0 + rad * z);
    if (dot(V, c) > 0.0) { M = c; break; }
  }
  return unit(M);
}

// the log density per solid angle of the unit facet m given V (local frame): p_m = 2 q (3 c^2 + q^2) / (3 Vol), c =
// m_z, q^2 = r^2 - sin^2, Vol the ball's volume on V's side (facets.py `_glisur_log_density`); -inf off the support
inline double glisur_lp(double r, V3<double> V, V3<double> m) {
  const double c = m.z, q2 = r * r - (m.x * m.x + m.y * m.y);
  if (!(q2 > 0.0 && c > 0.0 && dot(V, m) > 0.0)) return -INFINITY;
  const double d = std::max(r - V.z, 0.0);
  const double vol = 4.0 / 3.0 * PI * r * r * r - PI * d * d * (2.0 * r + V.z) / 3.0;
  return std::log(2.0 / 3.0) + 0.5 * std::log(q2) + std::log(3.0 * c * c + q2) - std::log(vol);
}

// the cover mixture's log density (1 - e) p_m + e p_shell, p_shell = 2 r (c^2 + q^2) / (q Area) (facets.py
// `_glisur_log_proposal`)
inline double glisur_lq(double r, V3<double> V, V3<double> m, double e) {
  const double lp = glisur_lp(r, V, m);
  if (!std::isfinite(lp)) return -INFINITY;
  const double c = m.z, q2 = r * r - (m.x * m.x + m.y * m.y);
  const double area = V.z >= r ? 4.0 * PI * r * r : TWO_PI * r * (r + std::min(V.z, r));
  const double ls = std::log(2.0 * r) + std::log(c * c + q2) - 0.5 * std::log(q2) - std::log(area);
  const double a = std::log1p(-e) + lp, b = std::log(e) + ls, mx = std::max(a, b);
  return mx + std::log(std::exp(a - mx) + std::exp(b - mx));
}

// one facet's log weight: the live density over the density drawn from; a facet that rounds off the drawn density's
// support weighs 0 (src/lumen/transport/surface_event.py:124-129)
inline void glisur_weigh(double &logw, double r_draw, double r_live, bool cover, V3<double> V, V3<double> m) {
  const double ld = cover ? glisur_lq(r_draw, V, m, DEFENSIVE) : glisur_lp(r_draw, V, m);
  logw = std::isfinite(ld) ? logw + glisur_lp(r_live, V, m) - ld : -INFINITY;
}

struct GlisurPath {
  V3<double> w;
  bool far;
  double logw;  // the sum of the facets' log(live density / drawn density)
};

// Geant4's DielectricMetal loop with GLISUR facets: a facet for the current direction, the mirror reflection about it,
// repeated while the direction points into the surface; the state mirrored about each facet.  Drawn at width
// r_draw (from the cover mixture when `cover`); logw accumulates log p_m(r_live) / drawn density per facet
inline GlisurPath glisur_metal_path(const Rng &g, uint32_t vertex, double r_draw, double r_live, bool cover,
                                    V3<double> wi, V3<double> n, V3<double> &er, V3<double> &ei) {
  V3<double> t1, t2;
  basis(n, t1, t2);
  auto local = [&](V3<double> v) { return v3(dot(v, t1), dot(v, t2), dot(v, n)); };
  V3<double> w = wi;
  GlisurPath out{wi, false, 0.0};
  for (int k = 0; k < 256; ++k) {
    double u[4];
    g.draw4(vertex, GLISUR_BLOCK + (uint32_t)k, u);
    const V3<double> V = local(-1.0 * w);
    const bool shell = cover && u[0] < DEFENSIVE;
    const V3<double> m = glisur_facet(g, vertex, k, r_draw, V, shell);
    if (cover || r_live != r_draw) glisur_weigh(out.logw, r_draw, r_live, cover, V, m);
    const V3<double> mw = to_world(m, n);
    const V3<double> wo = w - 2.0 * dot(w, mw) * mw;
    pol_mirror(er, ei, mw);
    w = wo;
    if (dot(w, n) >= 0.0) break;
  }
  out.w = w;
  return out;
}

// Geant4's DielectricDielectric loop with GLISUR facets: a facet about the current side's normal, reflected or
// refracted by the Fresnel shares of the photon's state there (its own uniform), the state recombined; a refraction
// that still points back at the surface swaps the side; ends when the direction leaves the surface on either side.
// `g4conv`: Geant4's own conventions for the state between facets (the in-plane part as a magnitude; at total
// reflection the real vector's mirror), only for the comparison with Geant4's histograms
inline GlisurPath glisur_dielectric_path(const Rng &g, uint32_t vertex, double r_draw, double r_live, bool cover,
                                         V3<double> wi, V3<double> n, double n1, double n2, V3<double> &er,
                                         V3<double> &ei, bool g4conv) {
  GlisurPath out{wi, false, 0.0};
  V3<double> w = wi;
  double side = 1.0;
  for (int k = 0; k < 256; ++k) {
    const V3<double> N = side * n;
    const double na = side > 0 ? n1 : n2, nb = side > 0 ? n2 : n1;
    V3<double> t1, t2;
    basis(N, t1, t2);
    const V3<double> V = v3(-dot(w, t1), -dot(w, t2), -dot(w, N));
    double u[4];
    g.draw4(vertex, GLISUR_BLOCK + (uint32_t)k, u);
    const bool shell = cover && u[0] < DEFENSIVE;
    const V3<double> ml = glisur_facet(g, vertex, k, r_draw, V, shell);
    if (cover || r_live != r_draw) glisur_weigh(out.logw, r_draw, r_live, cover, V, ml);
    const V3<double> m = to_world(ml, N);
    const double c1 = std::min(std::max(-dot(w, m), 0.0), 1.0);
    const Fres f = fresnel_amps(c1, na, nb);
    const V3<double> s_ax = s_axis(w, m), p_in = cross(w, s_ax);
    C2 es{dot(er, s_ax), dot(ei, s_ax)}, ep{dot(er, p_in), dot(ei, p_in)};
    const bool tir = (na / nb) * (na / nb) * (1.0 - c1 * c1) >= 1.0;
    if (g4conv && !tir) ep = {std::sqrt(ca2(ep)), 0.0};
    const double R = std::min(std::max(ca2(f.rs) * ca2(es) + ca2(f.rp) * ca2(ep), 0.0), 1.0) /
                     std::max(ca2(es) + ca2(ep), 1e-300);
    const bool matched = std::fabs(na - nb) <= INDEX_MATCH * std::max(na, nb);
    const bool refr = !tir && u[1] >= R;
    V3<double> w1;
    if (refr) {
      if (matched) {
        w1 = w;
      } else {
        const double eta = na / nb, s2t = eta * eta * (1.0 - c1 * c1);
        w1 = unit(eta * w + (eta * c1 - std::sqrt(std::max(1.0 - s2t, 0.0))) * m);
      }
    } else {
      w1 = w - 2.0 * dot(w, m) * m;
    }
    if (!(refr && matched)) {  // recombine the state on the facet's s axis into the new direction
      C2 as = refr ? f.ts : f.rs, ap = refr ? f.tp : f.rp;
      if (g4conv && tir && !refr) as = {-1.0, 0.0}, ap = {1.0, 0.0};
      const C2 cs = cm(as, es), cp = cm(ap, ep);
      const V3<double> p_out = cross(w1, s_ax);
      er = cs.re * s_ax + cp.re * p_out;
      ei = cs.im * s_ax + cp.im * p_out;
      pol_normalise(er, ei);
    }
    const double d = dot(w1, N);
    const bool end = refr ? d <= 0.0 : d >= 0.0;
    out.far = out.far != refr;
    w = w1;
    if (end) break;
    if (refr) side = -side;
  }
  out.w = w;
  return out;
}

// Rayleigh, polarised (phase.py `rayleigh_sample`, `dipole_cos`): the dipole axis er with probability |er|^2, else
// ei; c = cos(k', axis) from 3/4 (1 - c^2) by its inverse CDF
template <class R> V3<R> rayleigh_dir(V3<R> er, V3<R> ei, const double ul[4]) {
  const R a2 = dot(er, er), n2 = a2 + dot(ei, ei);
  const V3<R> axis = unit(R(ul[2]) * n2 < a2 ? er : ei);
  const R c = MIN<R>(MAX<R>(R(2.0) * COS((ACOS(MIN<R>(MAX<R>(R(1.0 - 2.0 * ul[0]), R(-1.0)), R(1.0))) - R(TWO_PI)) / R(3.0)),
                            R(-1.0)), R(1.0));
  const R sn = SQRT(MAX<R>(R(1.0) - c * c, R(0.0)));
  return unit(to_world(v3<R>(sn * COS(R(TWO_PI * ul[1])), sn * SIN(R(TWO_PI * ul[1])), c), axis));
}

// G4OpMieHG (src/lumen/media/phase.py:98-190), FP64: the forward lobe HG(g_f) with probability r, else the backward
// one, mirrored (theta -> pi - theta); the lobe constants are the medium's words 19-24 (one value each).  The
// cosine is Geant4's inverse CDF as PostStepDoIt writes it: mu = 2u(1+g)^2(1-g+gu)/(1-g+2gu)^2 - 1 (2u - 1 at g = 0).
struct MieConsts {
  double gf, gb, r;
};
inline MieConsts mie_consts(const Optics &o, const int32_t *M) {
  return {o.P[M[19] + M[20]], o.P[M[21] + M[22]], o.P[M[23] + M[24]]};
}
// the forward ratio the lobe is drawn at: the defensive mixture over (r, 1 - r) (phase.py `_mie_cover`,
// estimator.py `defensive_q`): a ratio of 0 (1) draws the absent lobe with the share e
inline double mie_cover(double r) {
  if (r > 0.0 && r < 1.0) return r;
  const double eps = DEFENSIVE / (1.0 - DEFENSIVE);
  return r <= 0.0 ? eps / (eps + 1.0) : 1.0 / (1.0 + eps);
}
inline double hg_mu(double g, double u) {
  if (g == 0.0) return 2.0 * u - 1.0;
  const double d = 1.0 - g + 2.0 * g * u;
  return std::min(std::max(2.0 * u * (1.0 + g) * (1.0 + g) * (1.0 - g + g * u) / (d * d) - 1.0, -1.0), 1.0);
}
// HG(mu; g) per steradian: (1 - g^2) / (4 pi (1 + g^2 - 2 g mu)^(3/2))
inline double hg_pdf(double g, double mu) {
  const double q = 1.0 + g * g - 2.0 * g * mu;
  return (1.0 - g * g) / (4.0 * PI * q * std::sqrt(q));
}
inline double mie_pdf(const MieConsts &c, double r, double mu) {
  return r * hg_pdf(c.gf, mu) + (1.0 - r) * hg_pdf(-c.gb, mu);
}
template <class R> V3<R> mie_dir(V3<R> wi, const MieConsts &c, double rc, const double ul[4]) {
  const bool fwd = ul[0] <= rc;
  double mu = hg_mu(fwd ? c.gf : c.gb, ul[1]);
  if (!fwd) mu = -mu;
  const double sn = std::sqrt(std::max(1.0 - mu * mu, 0.0));
  return unit(to_world(v3<R>(R(sn * std::cos(TWO_PI * ul[2])), R(sn * std::sin(TWO_PI * ul[2])), R(mu)), wi));
}

// Lambertian about n (lobes.py `_cosine`)
template <class R> V3<R> lambert_dir(V3<R> n, const double ul[4]) {
  const R rr = SQRT(R(ul[0]));
  return to_world(v3<R>(rr * COS(R(TWO_PI * ul[1])), rr * SIN(R(TWO_PI * ul[1])), SQRT(MAX<R>(R(1.0 - ul[0]), R(0.0)))), n);
}

template <class R> struct Moments {  // a photon's deposits summed per sensor (exact, unbounded)
  std::vector<std::pair<int, double>> m;
  void add(int s, double v) {
    for (auto &e : m)
      if (e.first == s) { e.second += v; return; }
    m.push_back({s, v});
  }
};

template <class R> bool emit(const Run &r, long long gi, Photon<R> &p, Tally &T) {
  const int j = (int)(std::upper_bound(r.ends.begin(), r.ends.end(), gi) - r.ends.begin());
  const lh_source &S = r.src[j];
  const long long local = gi - (j > 0 ? r.ends[j - 1] : 0);
  const unsigned long long id = (unsigned long long)(S.first_id + local);
  p.rng = Rng{r.k0, r.k1, (uint32_t)id, (uint32_t)(id >> 32)};
  double u[4];
  p.rng.draw4(SOURCE_VERTEX, 0, u);
  p.x = v3<R>(R(S.pos[0]), R(S.pos[1]), R(S.pos[2]));
  p.t = R(S.time);
  p.w0 = S.weight;
  p.er = p.ei = v3<R>(R(0.0), R(0.0), R(0.0));
  p.event = S.event;
  if (S.kind == LH_SRC_CHERENKOV) {
    const int32_t *M = r.live.medium(S.medium);
    const double e_lo = r.live.P[M[0]], e_hi = r.live.P[M[0] + M[1] - 1];
    const double lam_min = HC / e_hi, lam_max = HC / e_lo, zq = 1.0 / lam_min - 1.0 / lam_max;
    p.lam = 1.0 / (1.0 / lam_min - u[0] * zq);
    const double n = r.live.tab(M[0], M[1], HC / p.lam), bn = S.beta * n;
    const bool em = bn > 1.0;
    const double s2 = em ? 1.0 - 1.0 / (bn * bn) : 0.0;
    p.w0 = TWO_PI * FINE_STRUCTURE * S.charge * S.charge * S.length * s2 * zq * 1e9 * S.weight;
    const double cc = em ? 1.0 / bn : 1.0, sc = std::sqrt(std::max(1.0 - cc * cc, 0.0));
    const V3<R> axis = v3<R>(R(S.dir[0]), R(S.dir[1]), R(S.dir[2]));
    p.d = unit(to_world(v3<R>(R(sc * std::cos(TWO_PI * u[2])), R(sc * std::sin(TWO_PI * u[2])), R(cc)), axis));
    const double xs = u[1] * S.length;
    p.x = p.x + R(xs) * axis;
    p.t = R(S.time + xs / (S.beta * C_VACUUM));
    p.er = unit(dot(axis, p.d) * p.d - axis);
  } else {
    p.lam = S.lam_hi > S.lam ? S.lam + u[2] * (S.lam_hi - S.lam) : S.lam;
    if (S.kind == LH_SRC_BEAM) {
      p.d = unit(v3<R>(R(S.dir[0]), R(S.dir[1]), R(S.dir[2])));
    } else {
      const double c = 1.0 - 2.0 * u[0], sn = std::sqrt(std::max(1.0 - c * c, 0.0));
      p.d = v3<R>(R(sn * std::cos(TWO_PI * u[1])), R(sn * std::sin(TWO_PI * u[1])), R(c));
    }
  }
  p.w = p.wr = R(1.0);
  p.vol = locate(*r.s, v3<double>(val(p.x.x), val(p.x.y), val(p.x.z)));
  p.cls = r.s->cls[p.vol];
  p.srf = -1;
  p.xs = 0.0;
  p.nz = 0;
  p.vertex = 0;
  T.h[LH_H_PHOTONS] += 1;
  T.h[LH_H_START_WEIGHT] += p.w0;
  if (p.vol >= r.s->n || r.live.cls_medium(p.cls) < 0) {
    T.h[LH_H_LOST] += p.w0;
    return false;
  }
  return true;
}

inline double ratio(double a, double b) { return b != 0.0 ? a / b : 0.0; }
template <class R> R ratio(R a, R b) { return b != R(0.0) ? a / b : R(0.0); }

template <class R> bool step(const Run &r, Photon<R> &p, Tally &T, Moments<R> &mom) {
  const Scene &s = *r.s;
  if (p.nz >= ZERO_STEPS_BEFORE_PUSH) {
    if (p.nz == ZERO_STEPS_BEFORE_PUSH) T.h[LH_H_PUSHED] += val(p.w) * p.w0;
    p.x = p.x + R(PUSH) * p.d;
    p.vol = locate(s, v3<double>(val(p.x.x), val(p.x.y), val(p.x.z)));
    p.cls = s.cls[p.vol];
    p.srf = -1;
  }
  double u[4], ul[4];
  p.rng.draw4(p.vertex, 0, u);
  p.rng.draw4(p.vertex, 1, ul);
  V3<R> er0 = p.er, ei0 = p.ei;
  if (val(dot(er0, er0) + dot(ei0, ei0)) < 0.5) {  // unpolarised: a random linear state (polar.py `fill`)
    V3<R> t1, t2;
    basis(p.d, t1, t2);
    er0 = R(std::cos(TWO_PI * ul[3])) * t1 + R(std::sin(TWO_PI * ul[3])) * t2;
    ei0 = v3<R>(R(0.0), R(0.0), R(0.0));
  }
  double *tal = &T.t[(size_t)p.event * r.stride];
  // a record at a sensor: the tallies, the charge cells (theta about volume `vol`'s axis at the point xc), the time
  // histogram and the photon's moment
  auto deposit = [&](int sensor, int vol, V3<R> xc, double v, double time) {
    T.h[LH_H_RECORDS] += 1;
    if (v != 0.0) {
      tal[sensor] += v;
      tal[r.S + sensor] += v * time;
      // the charge cells: hat weights on the QE wavelength knots and on the CE knots of the cathode position
      // angle (readout/grid.py:161-185 `hats`, `cells`, `charge_cells`; sensors/features.py `theta`: the polar
      // angle of the crossing point about the sensor volume's +z axis), the wavelength's knot major
      int il = 0, it = 0;
      double fl = 0.0, ft = 0.0, accept = 1.0;
      if (r.Kl > 1) {
        r.live.tab(r.live.L[10], r.Kl, p.lam, &il, &fl);
        accept *= r.live.tab(r.live.L[10], r.Kl, p.lam);
      }
      if (r.Kt > 1) {
        V3<R> xl, dl;
        to_local(s, vol, xc, p.d, xl, dl);
        const double th = std::atan2(std::sqrt(val(xl.x) * val(xl.x) + val(xl.y) * val(xl.y)), val(xl.z));
        r.live.tab(r.live.L[15], r.Kt, th, &it, &ft);
        accept *= r.live.tab(r.live.L[15], r.Kt, th);
      }
      double *cc = &tal[3 * r.S + (long long)sensor * r.K + il * r.Kt + it];
      cc[0] += v * (1 - fl) * (1 - ft);
      if (fl > 0) cc[r.Kt] += v * fl * (1 - ft);
      if (ft > 0) cc[1] += v * (1 - fl) * ft;
      if (fl > 0 && ft > 0) cc[r.Kt + 1] += v * fl * ft;
      if (r.nt > 0) {  // the time histogram (readout/grid.py:190-225 `_time_weights`, `time_cells`)
        const double dq = v * accept;  // weighted by the acceptance: the detected light's time shape
        const double x = std::min(std::max((time - r.t0) / r.dt, -4.0), r.nt + 4.0);
        const double jf = std::floor(x), g = x - jf;
        const int j = (int)jf;
        const double wk[4] = {-g * (g - 1) * (g - 2) / 6, (g + 1) * (g - 1) * (g - 2) / 2,
                              -(g + 1) * g * (g - 2) / 2, (g + 1) * g * (g - 1) / 6};
        double late = 0, early = 0;
        for (int k = 0; k < 4; ++k) {
          const int nn = j - 1 + k;
          if (nn >= r.nt) late += wk[k];
          else if (nn < 0) early += wk[k];
          else tal[r.time_at + (long long)sensor * r.nt + nn] += dq * wk[k];
        }
        double *ov = tal + r.time_at + (long long)r.S * r.nt + 4LL * sensor;
        ov[0] += dq * late;
        ov[1] += dq * late * time;
        ov[2] += dq * early;
        ov[3] += dq * early * time;
      }
      mom.add(sensor, v);
    }
  };
  if (r.live.L[17] && p.vertex == 0) {  // emitted inside a sensor volume that counts by volume (step.py:574-612)
    const int br = r.live.L[r.live.L[17] + p.cls], sensor = s.sensor[p.vol];
    if (br >= 0 && sensor >= 0) {  // the entry's detected share at the emission (normal incidence, indices 1)
      const Surf<R> se = surface_eval<R>(r.live, br, p.lam, p.d, R(-1.0) * p.d, 1.0, 1.0, er0, ei0);
      deposit(sensor, p.vol, p.x, val(p.w * se.detect) * p.w0, val(p.t));
      T.h[LH_H_ABSORBED_SURFACE] += val(p.w) * p.w0;
      p.w = p.wr = R(0.0);
      p.vertex += 1;
      T.h[LH_H_VERTICES] += 1;
      return false;
    }
  }
  const int m = r.live.cls_medium(p.cls);
  const Med<R> ml = medium_at<R>(r.live, m, p.lam);
  const Med<R> mr = r.has_ref ? medium_at<R>(r.ref, m, p.lam) : ml;
  const R mu = (ml.mu_ray + ml.mu_mie) + ml.mu_wls, mu_r = (mr.mu_ray + mr.mu_mie) + mr.mu_wls;  // total rates
  const R s_free = -LOG(R(u[0])) / MAX<R>(mu_r, R(FLOOR32));
  const Cross<R> cr = query<R>(s, p.x, p.d, p.vol, p.srf, R(r.max_dist), p.xs);
  const bool hit = cr.vol_after >= 0;
  const R dist = cr.t;
  const bool zero = dist <= R(ZERO_STEP);
  const int cb = p.cls, ca = hit ? s.cls[cr.vol_after] : -1;
  int row = hit ? r.live.row_of(cb, ca) : -1;
  const bool cont = hit && (zero || ca == cb) && row < 0;  // no surface: zero length, or one class on both sides
  if (cont) row = 0;
  const bool scatter = mu_r > R(0.0) && s_free < dist;
  const R reach_live = EXP(-mu * dist), reach_ref = r.has_ref ? EXP(-mu_r * dist) : reach_live;
  const R att = EXP(-ml.mu_abs * dist), att_r = r.has_ref ? EXP(-mr.mu_abs * dist) : att;
  const R v_g = R(C_VACUUM) / ml.ng;
  Surf<R> sl{}, sr{};
  R pdraw[NOUT], pprop[NOUT];
  int o = -1;
  const bool at_if = hit && row >= 0;
  if (at_if) {
    const int m2 = r.live.cls_medium(ca);
    sl = surface_eval<R>(r.live, row, p.lam, p.d, cr.n, val(ml.n), index_at(r.live, m2, p.lam), er0, ei0);
    sr = r.has_ref ? surface_eval<R>(r.ref, row, p.lam, p.d, cr.n, val(mr.n), index_at(r.ref, m2, p.lam), er0, ei0)
                   : sl;
    if (cont) {
      for (int k = 0; k < NOUT; ++k) sl.p[k] = sr.p[k] = R(0.0);
      sl.p[STRAIGHT] = sr.p[STRAIGHT] = R(1.0);
      sl.detect = sr.detect = R(0.0);
    }
    const int support = (sr.flags >> 8) & 63;
    const bool edge = (sr.flags & 4) && !cont;
    int ncov = 0;
    for (int k = 0; k < NOUT; ++k) {
      pdraw[k] = sr.p[k];
      if (edge && ((support >> k) & 1) && !(sr.p[k] > R(0.0))) ++ncov;
    }
    if (ncov)
      for (int k = 0; k < NOUT; ++k)
        if (((support >> k) & 1) && !(sr.p[k] > R(0.0))) pdraw[k] = R(DEFENSIVE / (1 - DEFENSIVE) / ncov);
    R tot = R(0.0);
    for (int k = 0; k < NOUT; ++k) tot += pdraw[k];
    for (int k = 0; k < NOUT; ++k) pprop[k] = tot > R(0.0) ? pdraw[k] / tot : R(0.0);
    o = choose<R>(pprop, NOUT, u[2]);
    // the sensor entered, or across an exit sensor's border the region left (step.py:544-547 `_sensor_dep`)
    const int vs = (r.live.L[16] && r.live.L[r.live.L[16] + cb * r.live.L[2] + ca]) ? p.vol : cr.vol_after;
    const int sensor = s.sensor[vs];
    if (sensor >= 0 && !cont)
      deposit(sensor, vs, p.x + dist * p.d, val(p.w * reach_live * att * sl.detect) * p.w0, val(p.t + dist / v_g));
  }
  bool ok;
  R w_new = R(0.0), wr_new = R(0.0);
  if (scatter) {  // the channel from the reference rates (Rayleigh, Mie, WLS), step.py:397-436 `_bulk`
    // estimator.py `choose`: the first channel whose running sum exceeds u times the total (the first if none)
    const R xc = R(u[1]) * mu_r;
    const int ch = xc < mr.mu_ray ? 0 : xc < mr.mu_ray + mr.mu_mie ? 1 : xc < mu_r ? 2 : 0;
    const bool wls = ch == 2;
    const R mc = ch == 0 ? ml.mu_ray : ch == 1 ? ml.mu_mie : ml.mu_wls;
    const R mc_r = ch == 0 ? mr.mu_ray : ch == 1 ? mr.mu_mie : mr.mu_wls;
    const R w_path = ratio<R>(mc * EXP(-mu * s_free), mc_r * EXP(-mu_r * s_free));
    const R att_s = EXP(-ml.mu_abs * s_free);
    const R w_at = p.w * w_path;
    const R att_s_r = r.has_ref ? EXP(-mr.mu_abs * s_free) : att_s;
    p.x = p.x + s_free * p.d;
    p.t = p.t + s_free / v_g;
    if (!wls) {
      R w_lobe = R(1.0);
      V3<R> dn;
      if (ch == 0) {
        dn = rayleigh_dir(er0, ei0, ul);
      } else {  // Mie: drawn at the reference's lobes under the cover of its ratio, weighted live / drawn density
        const int32_t *M = r.live.medium(m);
        const MieConsts lv = mie_consts(r.live, M), lr = mie_consts(r.has_ref ? r.ref : r.live, M);
        const double rc = mie_cover(lr.r);
        dn = mie_dir(p.d, lr, rc, ul);
        const double c = std::min(std::max(val(dot(p.d, dn)), -1.0), 1.0);
        const double drawn = mie_pdf(lr, rc, c);
        w_lobe = R(drawn > 0.0 ? mie_pdf(lv, lv.r, c) / drawn : 0.0);
      }
      w_new = w_at * att_s * w_lobe;
      wr_new = p.wr * att_s_r;
      T.h[LH_H_ABSORBED_BULK] += val(w_at * (R(1.0) - att_s)) * p.w0;
      er0 = er0 - dot(dn, er0) * dn;
      ei0 = ei0 - dot(dn, ei0) * dn;
      pol_normalise(er0, ei0);
      p.d = dn;
      p.er = er0;
      p.ei = ei0;
      if (ch == 0 && r.jump) {  // the diffusion jump (transport/fastpath.py:166-224 `diffuse`)
        const double c = jump_clearance(m, p.x);
        const double mur = val(mr.mu_ray);
        if (c > 0.0 && mur > 0.0 && val(dot(p.ei, p.ei)) == 0.0) {
          const double budget = std::ceil(std::min(c * c * mur * mur, 1073741824.0));
          V3<R> y = v3<R>(R(0.0), R(0.0), R(0.0)), d = p.d, ee = p.er;
          const V3<R> zero = v3<R>(R(0.0), R(0.0), R(0.0));
          R L = R(0.0);
          int nsc = 0;
          for (double j = 0; j < budget; j += 1.0) {
            double u4[4];
            p.rng.draw4((uint32_t)p.vertex | 0x80000000u, (uint32_t)j, u4);
            const R sj = R(-std::log(u4[0]) / mur);
            const R by = dot(y, d);
            const R te = -by + SQRT(MAX<R>(by * by - (dot(y, y) - R(c * c)), R(0.0)));  // leaves the ball
            if (!(sj < te)) {
              y = y + te * d;
              L = L + te;
              break;
            }
            const double ul3[4] = {u4[1], u4[2], u4[3], 0.0};
            const V3<R> k = rayleigh_dir(ee, zero, ul3);
            V3<R> e2 = ee - dot(k, ee) * k, z2 = zero;
            pol_normalise(e2, z2);
            y = y + sj * d;
            L = L + sj;
            d = k;
            ee = e2;
            ++nsc;
          }
          const R wj = EXP(-ml.mu_abs * L) * EXP(R((double)nsc) * LOG(ml.mu_ray / mr.mu_ray) - (ml.mu_ray - mr.mu_ray) * L);
          T.h[LH_H_ABSORBED_BULK] += val(w_new * (R(1.0) - wj)) * p.w0;
          w_new = w_new * wj;
          wr_new = wr_new * EXP(-mr.mu_abs * L);
          p.x = p.x + y;
          p.d = d;
          p.er = ee;
          p.t = p.t + L / v_g;
          if (val(L) > 0.0) p.srf = -1;
        }
      }
    } else {  // G4OpWLS: isotropic, unpolarised, at the record's wavelength after its delay
      double ur[4];
      p.rng.draw4(p.vertex, 2, ur);
      const WlsDraw em = wls_emit(r.live, r.has_ref ? r.ref : r.live, r.live.medium(m), p.lam, ur[0], ur[1]);
      w_new = w_at * att_s * R(em.w);
      wr_new = p.wr * att_s_r * R(em.wr);
      T.h[LH_H_ABSORBED_BULK] += val(w_at * (R(1.0) - att_s * R(em.w))) * p.w0;
      const double c = 1.0 - 2.0 * ul[0], sn = std::sqrt(std::max(1.0 - c * c, 0.0));
      p.d = unit(to_world(v3<R>(R(sn * std::cos(TWO_PI * ul[1])), R(sn * std::sin(TWO_PI * ul[1])), R(c)), p.d));
      p.er = p.ei = v3<R>(R(0.0), R(0.0), R(0.0));
      p.lam = em.lam;
      p.t = p.t + R(em.delay);
    }
    if (!(s_free <= R(ZERO_STEP) && p.srf >= 0)) p.srf = -1;
    ok = true;
  } else if (hit) {
    const R w_at = p.w * ratio<R>(reach_live, reach_ref);
    R s_cont = R(0.0);
    if (at_if)
      for (int k = 0; k < NOUT; ++k) s_cont += sl.p[k];
    T.h[LH_H_ABSORBED_BULK] += val(w_at * (R(1.0) - att)) * p.w0;
    if (at_if) T.h[LH_H_ABSORBED_SURFACE] += val(w_at * att * (R(1.0) - s_cont)) * p.w0;
    else T.h[LH_H_LOST] += val(w_at * att) * p.w0;
    ok = at_if && o >= 0;
    if (ok) {
      const R w_if = w_at * att * ratio<R>(sl.p[o], pprop[o]);
      wr_new = p.wr * (reach_ref != R(0.0) ? R(1.0) : R(0.0)) * att_r * ratio<R>(pdraw[o], pprop[o]);
      const V3<R> wi = p.d, n = cr.n;
      V3<R> dn = wi;
      bool far = false;
      R w_lobe = R(1.0);  // a GLISUR facet path's live / drawn density
      if (o == STRAIGHT) {
        far = true;
      } else if (o == REFRACT) {
        const double eta = val(ml.n) / index_at(r.live, r.live.cls_medium(ca), p.lam);
        const R ci = -dot(wi, n), s2t = R(eta * eta) * (R(1.0) - ci * ci);
        if (s2t >= R(1.0)) dn = wi + R(2.0) * ci * n;
        else dn = unit(R(eta) * wi + (R(eta) * ci - SQRT(MAX<R>(R(1.0) - s2t, R(0.0)))) * n);
        pol_recombine(er0, ei0, wi, s_axis(wi, n), sl.amps.ts, sl.amps.tp, unit(dn));
        far = true;
      } else if (o == SPECULAR) {
        dn = wi - R(2.0) * dot(wi, n) * n;
        pol_recombine(er0, ei0, wi, s_axis(wi, n), sl.amps.rs, sl.amps.rp, unit(dn));
      } else if (o == BACKSCATTER) {
        dn = R(-1.0) * wi;
        er0 = R(-1.0) * er0;
        ei0 = R(-1.0) * ei0;
      } else if (o == GLOSSY && (sr.flags & 16)) {  // GLISUR's facet path at r = 1 - polish (types.py:460-467)
        auto dbl = [](V3<R> v) { return v3(val(v.x), val(v.y), val(v.z)); };
        auto back = [](V3<double> v) { return v3<R>(R(v.x), R(v.y), R(v.z)); };
        V3<double> e_r = dbl(er0), e_i = dbl(ei0);
        const bool cover = sr.flags & 64;
        const double rd = val(sr.sigma), rl = val(sl.sigma);
        const GlisurPath P =
            (sr.flags & 32) ? glisur_dielectric_path(p.rng, (uint32_t)p.vertex, rd, rl, cover, dbl(wi), dbl(n),
                                                     val(ml.n), index_at(r.live, r.live.cls_medium(ca), p.lam), e_r,
                                                     e_i, false)
                            : glisur_metal_path(p.rng, (uint32_t)p.vertex, rd, rl, cover, dbl(wi), dbl(n), e_r, e_i);
        dn = back(P.w);
        er0 = back(e_r);
        ei0 = back(e_i);
        far = P.far;
        w_lobe = R(std::exp(P.logw));
      } else if (o == GLOSSY) {
        dn = facet_path(p, val(sr.sigma), wi, n, er0, ei0);
      } else {
        dn = lambert_dir(n, ul);
        pol_mirror(er0, ei0, unit(dn - wi));
      }
      w_new = w_if * w_lobe;
      T.h[LH_H_REWEIGHTED] += val(w_at * att * s_cont - w_new) * p.w0;
      p.xs = std::max(std::max(std::fabs((double)val(p.x.x)), std::fabs((double)val(p.x.y))),
                      std::fabs((double)val(p.x.z))) + (double)val(dist);
      p.x = p.x + dist * p.d;
      p.t = p.t + dist / v_g;
      p.d = unit(dn);
      p.er = er0;
      p.ei = ei0;
      p.srf = cr.srf;
      if (far) {
        p.vol = cr.vol_after;
        p.cls = ca;
      }
    }
  } else {
    const R d_end = MIN<R>(dist, R(r.max_dist));
    const R w_far = p.w * ratio<R>(EXP(-mu * d_end), EXP(-mu_r * d_end));
    const R att_far = EXP(-ml.mu_abs * d_end);
    T.h[r.max_dist < r.s->reach ? LH_H_DISTANCE_CUT : LH_H_ESCAPED] += val(w_far * att_far) * p.w0;
    T.h[LH_H_ABSORBED_BULK] += val(w_far * (R(1.0) - att_far)) * p.w0;
    ok = false;
  }
  bool alive = ok;
  if (ok && wr_new < R(r.roulette)) {
    const R ps = wr_new / R(r.roulette);
    if (R(u[3]) < ps) {
      const R wk = w_new / ps;
      T.h[LH_H_ROULETTED] += val(w_new - wk) * p.w0;
      w_new = wk;
      wr_new = wr_new / ps;
    } else {
      T.h[LH_H_ROULETTED] += val(w_new) * p.w0;
      w_new = wr_new = R(0.0);
      alive = false;
    }
  }
  // a photon left without live weight deposits nothing more: ended unless the controls keep it (the engine's rule;
  // the reference walks it on, which changes only the records and vertices counts)
  if (alive && val(w_new) == 0.0 && !r.keep_zero) alive = false;
  p.nz = (hit && !scatter && zero) ? p.nz + 1 : 0;
  p.w = w_new;
  p.wr = wr_new;
  p.vertex += 1;
  T.h[LH_H_VERTICES] += 1;
  if (alive && p.nz >= ZERO_STEPS_BEFORE_KILL) {
    T.h[LH_H_ABANDONED] += val(p.w) * p.w0;
    alive = false;
  }
  if (alive && p.vertex >= r.max_vert) {
    T.h[LH_H_TRUNCATED] += val(p.w) * p.w0;
    alive = false;
  }
  return alive;
}

template <class R> void trace_range(const Run &r, long long a, long long b, Tally &T) {
  for (long long gi = a; gi < b; ++gi) {
    Photon<R> p;
    Moments<R> mom;
    bool alive = emit<R>(r, gi, p, T);
    while (alive) alive = step<R>(r, p, T, mom);
    double *tal = &T.t[(size_t)p.event * r.stride];
    for (auto &e : mom.m) tal[2 * r.S + e.first] += e.second * e.second;
  }
}

}  // namespace orc

using namespace orc;

#define ORC_API extern "C" __attribute__((visibility("default")))

static int make_run(Run &r, const Scene &s, const int32_t *layout, const float *values, const float *ref,
                    const lh_source *src, int nsrc, uint64_t seed, const lh_controls *ctl) {
  r.s = &s;
  r.live = Optics{layout, values};
  r.ref = Optics{layout, ref ? ref : values};
  r.has_ref = ref != nullptr && ref != values;
  r.src = src;
  r.nsrc = nsrc;
  long long acc = 0;
  r.ends.resize(nsrc);
  for (int i = 0; i < nsrc; ++i) r.ends[i] = (acc += src[i].n);
  r.k0 = (uint32_t)seed;
  r.k1 = (uint32_t)(seed >> 32);
  r.roulette = ctl ? ctl->roulette : 0.01;
  r.max_dist = (ctl && ctl->max_distance > 0) ? ctl->max_distance : s.reach;
  r.max_vert = (ctl && ctl->max_vertices > 0) ? ctl->max_vertices : (1 << 22);
  r.jump = ctl && ctl->jump;
  r.keep_zero = ctl && ctl->keep_zero_weight;
  r.S = layout[4];
  r.Kl = layout[5];
  r.Kt = layout[14];
  r.K = r.Kl * r.Kt;
  r.nt = layout[11];
  float t0f, dtf;
  std::memcpy(&t0f, &layout[12], 4);
  std::memcpy(&dtf, &layout[13], 4);
  r.t0 = t0f;
  r.dt = dtf;
  r.time_at = 3LL * r.S + (long long)r.S * r.K;
  r.stride = r.time_at + (r.nt > 0 ? (long long)r.S * (r.nt + 4) : 0);
  return 0;
}

// Trace every photon of the sources in FP64 on `threads` OpenMP threads; add into tally (n_events * stride) and
// health (LH_HEALTH_N), both host arrays.  `use_bvh`: daughters by the BVH walk (the reference's navigator, the CPU
// baseline's mode); 0: by brute force over every daughter (the independent referee).
template <class R>
static int trace_all(const lh_geometry *G, const int32_t *layout, const float *values, const float *ref,
                     const lh_source *src, int nsrc, uint64_t seed, const lh_controls *ctl, double *tally, int n_events,
                     double *health, int threads, int use_bvh);

ORC_API int oracle_trace(const lh_geometry *G, const int32_t *layout, const float *values, const float *ref,
                         const lh_source *src, int nsrc, uint64_t seed, const lh_controls *ctl, double *tally,
                         int n_events, double *health, int threads, int use_bvh) {
  return trace_all<double>(G, layout, values, ref, src, nsrc, seed, ctl, tally, n_events, health, threads, use_bvh);
}

// The same loop with the photon's state and arithmetic in float32 (the reference's CPU path traces in FP32; tallies
// stay FP64): the CPU baseline's FP32 figure.  Not a checker: its navigator is the FP64 referee's, in float32.
ORC_API int oracle_trace_f32(const lh_geometry *G, const int32_t *layout, const float *values, const float *ref,
                             const lh_source *src, int nsrc, uint64_t seed, const lh_controls *ctl, double *tally,
                             int n_events, double *health, int threads, int use_bvh) {
  return trace_all<float>(G, layout, values, ref, src, nsrc, seed, ctl, tally, n_events, health, threads, use_bvh);
}

template <class R>
static int trace_all(const lh_geometry *G, const int32_t *layout, const float *values, const float *ref,
                     const lh_source *src, int nsrc, uint64_t seed, const lh_controls *ctl, double *tally, int n_events,
                     double *health, int threads, int use_bvh) {
  Scene s = scene_of(G);
  s.bvh = use_bvh != 0;
  Run r;
  make_run(r, s, layout, values, ref, src, nsrc, seed, ctl);
  const long long total = r.ends.empty() ? 0 : r.ends.back();
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<Tally> parts;
#pragma omp parallel
  {
#pragma omp single
    parts.resize(omp_get_num_threads());
    Tally &T = parts[omp_get_thread_num()];
    T.t.assign((size_t)n_events * r.stride, 0.0);
#pragma omp for schedule(dynamic, 1024)
    for (long long gi = 0; gi < total; ++gi) trace_range<R>(r, gi, gi + 1, T);
  }
  for (auto &T : parts) {
    for (size_t i = 0; i < T.t.size(); ++i) tally[i] += T.t[i];
    for (int k = 0; k < LH_HEALTH_N; ++k) health[k] += T.h[k];
  }
  return 0;
}

// The algorithmic work of the first n photons: floating-point operations (Counted) and Philox integer operations,
// and the vertices they took (single thread), with the daughters found by the BVH walk as the reference finds them.
ORC_API int oracle_count_ops(const lh_geometry *G, const int32_t *layout, const float *values, const lh_source *src,
                             int nsrc, uint64_t seed, const lh_controls *ctl, long long n, double *out3) {
  Scene s = scene_of(G);
  s.bvh = true;  // the reference's algorithm walks the BVH
  Run r;
  make_run(r, s, layout, values, nullptr, src, nsrc, seed, ctl);
  Tally T;
  T.t.assign((size_t)64 * r.stride, 0.0);
  const long long total = std::min(n, r.ends.empty() ? 0LL : r.ends.back());
  g_ops = 0;
  g_int_ops = 0;
  for (long long gi = 0; gi < total; ++gi) {
    Photon<Counted> p;
    Moments<Counted> mom;
    if (src[0].event >= 64) return 1;
    bool alive = emit<Counted>(r, gi, p, T);
    while (alive) alive = step<Counted>(r, p, T, mom);
  }
  out3[0] = (double)g_ops;
  out3[1] = (double)g_int_ops;
  out3[2] = T.h[LH_H_VERTICES];
  return 0;
}

// The navigator referee: the next visible crossing of n rays (x, d: n*3; vol, last per ray), FP64.  Outputs t,
// the volume entered, the normal (n*3).
ORC_API int oracle_query(const lh_geometry *G, int n, const double *x, const double *d, const int32_t *vol,
                         const int64_t *last, double tmax, double *t, int32_t *vol_after, double *normal,
                         int64_t *srf) {
  const Scene s = scene_of(G);
  for (int i = 0; i < n; ++i) {
    const Cross<double> c = query<double>(s, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]),
                                          unit(v3(d[3 * i], d[3 * i + 1], d[3 * i + 2])), vol[i], last[i], tmax);
    t[i] = c.t;
    vol_after[i] = c.vol_after;
    normal[3 * i] = c.n.x, normal[3 * i + 1] = c.n.y, normal[3 * i + 2] = c.n.z;
    srf[i] = c.srf;
  }
  return 0;
}

ORC_API int oracle_locate(const lh_geometry *G, int n, const double *x, int32_t *vol) {
  const Scene s = scene_of(G);
  for (int i = 0; i < n; ++i) vol[i] = locate(s, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
  return 0;
}

ORC_API void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// n facet paths off a flat surface (normal +z) at incidence wi with width sigma; outgoing directions (n*3)
ORC_API void oracle_facet_paths(double sigma, const double *wi, long long n, uint64_t seed, double *out) {
  for (long long i = 0; i < n; ++i) {
    Photon<double> p;
    p.rng = Rng{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, (uint32_t)(i >> 32)};
    p.vertex = 0;
    V3<double> er = v3(0.0, 1.0, 0.0), ei = v3(0.0, 0.0, 0.0);
    const V3<double> w = facet_path(p, sigma, unit(v3(wi[0], wi[1], wi[2])), v3(0.0, 0.0, 1.0), er, ei);
    out[3 * i] = w.x, out[3 * i + 1] = w.y, out[3 * i + 2] = w.z;
  }
}

// Fresnel amplitudes (re, im of r_s, r_p, t_s, t_p) at incidence cosine c from n1 into n2
ORC_API void oracle_fresnel(double c, double n1, double n2, double out8[8]) {
  const Fres f = fresnel_amps(c, n1, n2);
  const C2 a[4] = {f.rs, f.rp, f.ts, f.tp};
  for (int i = 0; i < 4; ++i) out8[2 * i] = a[i].re, out8[2 * i + 1] = a[i].im;
}

// n draws of the Rayleigh direction for the state (er, ei) and of the Lambertian direction about +z
ORC_API void oracle_lobes(const double *er, const double *ei, long long n, uint64_t seed, double *ray, double *lam) {
  const Rng g{(uint32_t)seed, (uint32_t)(seed >> 32), 7u, 0u};
  for (long long i = 0; i < n; ++i) {
    double ul[4];
    g.draw4((uint32_t)i, (uint32_t)(i >> 32), ul);
    const V3<double> r = rayleigh_dir(v3(er[0], er[1], er[2]), v3(ei[0], ei[1], ei[2]), ul);
    const V3<double> l = lambert_dir(v3(0.0, 0.0, 1.0), ul);
    ray[3 * i] = r.x, ray[3 * i + 1] = r.y, ray[3 * i + 2] = r.z;
    lam[3 * i] = l.x, lam[3 * i + 1] = l.y, lam[3 * i + 2] = l.z;
  }
}

// the diffusion jump's clearance grid of medium m (clear: n[0] n[1] n[2] float32, x-major); m < 0 clears them all
ORC_API void oracle_set_jump(int m, const double *lo, double h, const int *n, const float *clear) {
  if (m < 0) {
    g_jump.clear();
    return;
  }
  if ((int)g_jump.size() <= m) g_jump.resize(m + 1);
  JumpGridO &G = g_jump[m];
  for (int a = 0; a < 3; ++a) G.lo[a] = lo[a], G.n[a] = n[a];
  G.h = h;
  G.clear.assign(clear, clear + (size_t)n[0] * n[1] * n[2]);
}

// n Mie directions about wi drawn at the constants (gf, gb, r) under the cover of r, and the weight of the live
// constants (lf, lb, lr) over the drawn density: out (n*4) = direction, weight
ORC_API void oracle_mie(const double *wi, double gf, double gb, double r, double lf, double lb, double lr,
                        long long n, uint64_t seed, double *out) {
  const Rng g{(uint32_t)seed, (uint32_t)(seed >> 32), 11u, 0u};
  const MieConsts drawn{gf, gb, r}, live{lf, lb, lr};
  const double rc = mie_cover(r);
  const V3<double> w = unit(v3(wi[0], wi[1], wi[2]));
  for (long long i = 0; i < n; ++i) {
    double ul[4];
    g.draw4((uint32_t)i, (uint32_t)(i >> 32), ul);
    const V3<double> d = mie_dir(w, drawn, rc, ul);
    const double c = std::min(std::max(dot(w, d), -1.0), 1.0);
    out[4 * i] = d.x, out[4 * i + 1] = d.y, out[4 * i + 2] = d.z;
    out[4 * i + 3] = mie_pdf(live, live.r, c) / mie_pdf(drawn, rc, c);
  }
}

// the Mie mixture density per steradian at the cosines mu (n of them)
ORC_API void oracle_mie_pdf(double gf, double gb, double r, long long n, const double *mu, double *out) {
  const MieConsts c{gf, gb, r};
  for (long long i = 0; i < n; ++i) out[i] = mie_pdf(c, r, mu[i]);
}

// n re-emissions by medium m's wavelength shifter of photons absorbed at lam_in (uniforms u_lam, u_time: n each);
// out (n*4): the emitted wavelength (nm), the delay (ns), the weight and the reference weight
ORC_API int oracle_wls_emit(const int32_t *layout, const float *values, const float *ref, int m, double lam_in,
                            long long n, const double *u_lam, const double *u_time, double *out) {
  const Optics live{layout, values}, rf{layout, ref ? ref : values};
  const int32_t *M = live.medium(m);
  if (!(M[8] & 2)) return 1;
  for (long long i = 0; i < n; ++i) {
    const WlsDraw e = wls_emit(live, rf, M, lam_in, u_lam[i], u_time[i]);
    out[4 * i] = e.lam, out[4 * i + 1] = e.delay, out[4 * i + 2] = e.w, out[4 * i + 3] = e.wr;
  }
  return 0;
}

// The first n photons of the sources as emitted (before any step): out (n*16) = x (3), d (3), t, lambda, the start
// weight w0, the polarisation er (3) and ei (3), the volume located at x
ORC_API int oracle_emit(const lh_geometry *G, const int32_t *layout, const float *values, const lh_source *src,
                        int nsrc, uint64_t seed, long long n, double *out) {
  Scene s = scene_of(G);
  Run r;
  make_run(r, s, layout, values, nullptr, src, nsrc, seed, nullptr);
  const long long total = std::min(n, r.ends.empty() ? 0LL : r.ends.back());
  Tally T;
  T.t.assign((size_t)64 * r.stride, 0.0);
  for (long long gi = 0; gi < total; ++gi) {
    Photon<double> p;
    emit<double>(r, gi, p, T);
    double *o = out + 16 * gi;
    const double v[16] = {p.x.x, p.x.y, p.x.z, p.d.x, p.d.y, p.d.z, p.t, p.lam, p.w0,
                          p.er.x, p.er.y, p.er.z, p.ei.x, p.ei.y, p.ei.z, (double)p.vol};
    for (int k = 0; k < 16; ++k) o[k] = v[k];
  }
  return (int)total;
}

// The Jones step of a specular reflection or a refraction: the state e = (er, ei) of a photon along wi meeting a
// surface of normal n, recombined on the s axis with the amplitudes (a_s, a_p) (amps = re, im of each) into the
// direction k; out = the new (er, ei), normalised
ORC_API void oracle_pol_recombine(const double wi[3], const double n[3], const double amps[4], const double k[3],
                                  const double e[6], double out[6]) {
  V3<double> er = v3(e[0], e[1], e[2]), ei = v3(e[3], e[4], e[5]);
  const V3<double> w = v3(wi[0], wi[1], wi[2]), nn = v3(n[0], n[1], n[2]);
  pol_recombine(er, ei, w, s_axis(w, nn), C2{amps[0], amps[1]}, C2{amps[2], amps[3]}, v3(k[0], k[1], k[2]));
  const double v[6] = {er.x, er.y, er.z, ei.x, ei.y, ei.z};
  for (int i = 0; i < 6; ++i) out[i] = v[i];
}

// The state after a mirror of normal h (the Lambertian and facet paths): out = the new (er, ei)
ORC_API void oracle_pol_mirror(const double h[3], const double e[6], double out[6]) {
  V3<double> er = v3(e[0], e[1], e[2]), ei = v3(e[3], e[4], e[5]);
  pol_mirror(er, ei, v3(h[0], h[1], h[2]));
  const double v[6] = {er.x, er.y, er.z, ei.x, ei.y, ei.z};
  for (int i = 0; i < 6; ++i) out[i] = v[i];
}

// Fresnel amplitudes (re, im of r_s, r_p, t_s, t_p) at incidence cosine c from n1 into the complex index nr + i nk
ORC_API void oracle_fresnel_complex(double c, double n1, double nr, double nk, double out8[8]) {
  const Fres f = fresnel_amps_complex(c, n1, nr, nk);
  const C2 a[4] = {f.rs, f.rp, f.ts, f.tp};
  for (int i = 0; i < 4; ++i) out8[2 * i] = a[i].re, out8[2 * i + 1] = a[i].im;
}

// n GLISUR facet paths off a flat surface (normal +z) at incidence wi, width r = 1 - polish (`dielectric`: from n1 into
// n2, else the metal's), the state er (n*3 per photon, ei = 0); out (n*4): the last direction and the far side
ORC_API void oracle_glisur_paths(int dielectric, double r, const double *wi, double n1, double n2, const double *er,
                                 long long n, uint64_t seed, int g4conv, double *out) {
  for (long long i = 0; i < n; ++i) {
    const Rng g{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, (uint32_t)(i >> 32)};
    V3<double> e_r = v3(er[3 * i], er[3 * i + 1], er[3 * i + 2]), e_i = v3(0.0, 0.0, 0.0);
    const V3<double> w = unit(v3(wi[0], wi[1], wi[2])), z = v3(0.0, 0.0, 1.0);
    const GlisurPath P = dielectric ? glisur_dielectric_path(g, 0, r, r, false, w, z, n1, n2, e_r, e_i, g4conv != 0)
                                    : glisur_metal_path(g, 0, r, r, false, w, z, e_r, e_i);
    out[4 * i] = P.w.x, out[4 * i + 1] = P.w.y, out[4 * i + 2] = P.w.z, out[4 * i + 3] = P.far ? 1.0 : 0.0;
  }
}

// n GLISUR facets (local frame) for V at width r, drawn from the cover mixture with share e when e > 0; out (n*4):
// the facet and its log density log p_m (e = 0) or the mixture's
ORC_API void oracle_glisur_facets(double r, const double *V, double e, long long n, uint64_t seed, double *out) {
  const V3<double> v = unit(v3(V[0], V[1], V[2]));
  for (long long i = 0; i < n; ++i) {
    const Rng g{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)i, (uint32_t)(i >> 32)};
    double u[4];
    g.draw4(0, GLISUR_BLOCK, u);
    const V3<double> m = glisur_facet(g, 0, 0, r, v, e > 0.0 && u[0] < e);
    out[4 * i] = m.x, out[4 * i + 1] = m.y, out[4 * i + 2] = m.z;
    out[4 * i + 3] = e > 0.0 ? glisur_lq(r, v, m, e) : glisur_lp(r, v, m);
  }
}
