// The photon loop on gfx950: one persistent kernel traces every photon of every source and tallies each sensor's
// arrivals in FP64, behind the C ABI of include/lumen_hip.h.
//
// Restates the reference's loop (src/lumen/run.py:293-368 `_loop`: one item per photon stepping until it ends) and
// its step (src/lumen/transport/step.py:516-571 `step`, :318-359 `_step_context`, :362-394 `_surface_event`,
// :397-436 `_bulk`, :464-492 `_after`, :495-513 `_advance`), with the deposit records' readout fused in
// (src/lumen/run.py:443-557 `rebuild`/`_moment_adds`, src/lumen/readout/fill.py:88-213): every deposit adds q w0 to
// `arrival`, q w0 t to `arrival_t`, q w0 hat_k(lam) to `charge` and, when its photon ends, each sensor's square of
// the photon's summed deposits to `arrival2` (the per-photon second moment).
//
// Schedule: each wavefront keeps 64 photons in flight; a lane whose photon ended takes the next one from a chunk its
// wavefront claimed from the device work queue (one atomic per chunk), so lanes refill every step (the reference's
// "step" schedule, src/lumen/run.py:81) and a launch of many sources with uneven photon counts keeps every lane busy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lumen_hip.h"
#include "adjoint.h"
#include "nav.h"
#include "optics.h"
#include "philox.h"
#include "unified.h"

namespace lh_host {
const std::vector<float> &unified_tables();
}

namespace lh {

constexpr float ZERO_STEP_BANDS = 0.125f;
constexpr int ZERO_STEPS_BEFORE_PUSH = 10;
constexpr int ZERO_STEPS_BEFORE_KILL = 25;
constexpr float PUSH_BANDS = 8.0f;
constexpr float FINE_STRUCTURE = 1.0f / 137.035999084f;
constexpr int MOMENT_SLOTS = 4;
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int MAX_FACETS = 256;
constexpr int MAX_FACET_ATTEMPTS = 4096;
// Every photon-loop kernel runs one workgroup of 12 wavefronts per CU (3 per SIMD at a 168-VGPR budget): the BVH
// nodes, the packed tables and the tally or gradient copy are staged once per CU instead of once per 4 wavefronts,
// leaving room for all of them.
#ifndef LH_FWD_BLOCK
#define LH_FWD_BLOCK 768
#endif
constexpr int BLOCK = LH_FWD_BLOCK;
template <unsigned F>
constexpr int block_of() { return BLOCK; }
template <unsigned F>
constexpr int adj_block_of() { return BLOCK; }
constexpr int DEFAULT_CHUNK = 256;
constexpr int MAX_GRAD = 8192;  // gradient slots a block accumulates in LDS (64 KiB)
constexpr long long LDS_TALLY_CELLS = 20480;  // tallies of at most this many cells (160 KiB) may go to LDS
constexpr int OCC_CACHE = 8;
// a kernel's features (template parameter): separate reference values, bulk channels beyond Rayleigh (Mie,
// WLS), the time histogram, BVH nodes read from global memory
constexpr unsigned F_REF = 1, F_CHAN = 2, F_TIME = 4, F_GNODES = 8, F_TORUS = 16, F_REACH = 32;
// the features an adjoint kernel is specialised on
constexpr unsigned F_ADJ_MASK = F_REF | F_CHAN | F_GNODES | F_TORUS | F_REACH;

// The forward kernel's register budget: 3 wavefronts per SIMD (168 VGPRs) hold the step with few spills; at 4
// (128 VGPRs) about 77 VGPRs spill to scratch, several GB of scratch traffic per launch, and the 1,000-PMT tank
// runs 4 % slower; at 2 the walk's load latency shows: the kernel with separate reference values (224 VGPRs at its
// natural 2) takes 9.7 ms for the tank's 2^24 photons against 6.3 at 3.
#ifndef LH_WAVES_PER_EU
#define LH_WAVES_PER_EU 3
#endif
#define LH_OCC_FWD(F) __attribute__((amdgpu_waves_per_eu(LH_WAVES_PER_EU)))

struct Args {
  Geo g;
  const int *__restrict__ L;
  const float *__restrict__ P;
  const float *__restrict__ Pr;
  int has_ref;
  const lh_source *__restrict__ src;
  const long long *__restrict__ src_end;  // inclusive prefix sums of the sources' photon counts
  int nsrc;
  long long total;
  uint32_t k0, k1;
  float roulette, max_dist, reach;
  int max_vert, chunk;
  int keep_zero;  // 1: a photon whose live weight became 0 walks on (the reference's records); 0: it ends
  int reach_on;   // 1: the reach bound where the model has one (the controls' reach_bound)
  double *__restrict__ tally;
  long long tally_stride;  // doubles per event
  int S, K;  // sensors; charge cells per sensor, Kl wavelength (QE) knots x Kt theta (CE) knots
  int Kl, Kt;
  int nt;               // the time histogram's nodes (0: none); its cells start at time_at of each event's tally
  long long time_at;
  float t0, dt;
  double *__restrict__ health;
  unsigned long long *__restrict__ queue;
  const float *__restrict__ ut;  // the unified facet tables (unified_tables.cpp)
  int lds_tally;                 // the block's tallies live in LDS (flushed once at exit), else global atomics
  int lds_values, lds_values_off;  // the forward kernel stages the packed values (lds_values floats) in LDS at
                                   // float offset lds_values_off of lds_cells: its table searches read LDS
  int lds_layout, lds_layout_off;  // ... the layout (words), at that word offset
  int lds_items, lds_ixf_off, lds_ikind_off;  // ... the BVH items' transforms and kinds (float4 / int4 offsets)
  // the diffusion jump (null: off): per medium (lo, h) and (n, offset into jclear, -1 no grid) as float4 pairs
  const float4 *__restrict__ jrec;
  const float *__restrict__ jclear;
  // the sensor-crediting rules (offsets into L, 0: none): a (C x C) table whose non-zero pairs credit a crossing's
  // detected share to the sensor of the volume left (an exit sensor), and per class the row whose detected share a
  // photon emitted inside a sensor volume of that class deposits at its emission (-1: none)
  int left_at, born_at;
};

struct Photon {
  float3 x, d, er, ei;
  float t, lam, w, wr, w0;
  int vol, srf, nz, vertex, cls, event;
  float xs;  // the coordinate scale |x0| + t of the ray whose crossing put the photon on face srf: the start's
             // rounding off that face is within that ray's band, however near the origin the crossing lies
  uint32_t id_lo, id_hi;  // the photon's id: its Philox counter words (the key is the launch's seed)
  // the medium cache: the class it was evaluated for, live and reference values
  int mcls;
  MediumAt ml, mr;
  // the per-photon moment: up to MOMENT_SLOTS sensors and the photon's summed deposits at each
  // (clang vectors: an index known only at run time selects within registers, where an array would go to memory)
  i4v ms;
  f4v mv;
  // the adjoint's walk: the launch index (to replay), the pass (0 totals, 1 accumulates), the photon's total of
  // cotangent-weighted deposits T (from pass 0), their running sum, and the rates' derivative caches
  long long gi;
  int pass;
  float T, Dsum;
  bool hasdep;
  RateD da, dr, dm, dw;
};

struct Acc {  // per-lane health accumulators of the frequent fates, flushed once per launch
  float bulk, surf, rou, rew, start_w;
  int records, vertices, photons;
};

// the rare fates (a photon's end, a stuck push, a moment spill) go to the block's LDS cells by atomics: registers
// held all launch long for them would crowd the step's working set
enum { RF_LOST = 0, RF_ESC, RF_DCUT, RF_ABAND, RF_TRUNC, RF_PUSHED, RF_SPILLS, RF_N };
__shared__ float s_rare[RF_N];
// per wavefront of the block, the volume its last located emission started in: a new photon whose start lies in that
// volume's daughter-free core starts in that volume (no daughter holds a point of the core), with no descent from the
// top (-1: none yet; lanes overwrite it as they like, any value is only a hint that the core test checks)
__shared__ int s_hint[32];
// the block's steps whose query the reach bound limited (a wavefront adds its count once per step)
__shared__ unsigned s_bounded;

__device__ __forceinline__ void rare_add(int cell, float v) {
  if (v != 0.0f) atomicAdd(&s_rare[cell], v);
}

// a[o] by selects (a register array indexed at run time would live in scratch)
__device__ __forceinline__ float pick6(const float a[6], int o) {
  float r = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) r = (k == o) ? a[k] : r;
  return r;
}

__device__ __forceinline__ int class_medium(const Args &a, int cls) {
  return cls < 0 ? -1 : a.L[a.L[L_CLASS_MEDIUM_AT] + cls];
}


__device__ __forceinline__ void tally_add(const Args &a, int event, long long slot, double v) {
  const long long i = (long long)event * a.tally_stride + slot;
  if (a.lds_tally) {
    atomicAdd(lds_cells + i, v);
  } else {
    unsafeAtomicAdd(a.tally + i, v);
  }
}

__device__ inline void moment_flush_slot(const Args &a, Photon &p, int k) {
  if (p.ms[k] >= 0) {
    const double v = (double)p.mv[k];
    tally_add(a, p.event, 2 * a.S + p.ms[k], v * v);
    p.ms[k] = -1;
    p.mv[k] = 0.0f;
  }
}

__device__ inline void moment_add(const Args &a, Photon &p, int s, float v) {
  const i4v hit = p.ms == s;  // lanes of -1 where the slot holds sensor s
  if (hit.x | hit.y | hit.z | hit.w) {
    p.mv += hit ? v : 0.0f;
    return;
  }
  const i4v freev = p.ms < 0;
  int free_k = freev.x ? 0 : freev.y ? 1 : freev.z ? 2 : freev.w ? 3 : -1;
  if (free_k < 0) {  // more sensors than slots: square the oldest now (its later deposits start a new sum)
    moment_flush_slot(a, p, 0);
    p.ms = i4v{p.ms.y, p.ms.z, p.ms.w, -1};
    p.mv = f4v{p.mv.y, p.mv.z, p.mv.w, 0.0f};
    free_k = MOMENT_SLOTS - 1;
    rare_add(RF_SPILLS, 1.0f);
  }
  p.ms[free_k] = s;
  p.mv[free_k] = v;
}

// the cathode position angle of a crossing point x (world) on sensor volume v: the polar angle about the volume's
// +z axis through its origin (the reference's sensors/features.py `theta`, WCSim's collection-efficiency angle)
__device__ inline float cathode_theta(const Geo &g, int v, float3 x) {
  const float4 r0 = g.xf[3 * v], r1 = g.xf[3 * v + 1], r2 = g.xf[3 * v + 2];
  const float lx = fmaf(r0.x, x.x, fmaf(r0.y, x.y, fmaf(r0.z, x.z, r0.w)));
  const float ly = fmaf(r1.x, x.x, fmaf(r1.y, x.y, fmaf(r1.z, x.z, r1.w)));
  const float lz = fmaf(r2.x, x.x, fmaf(r2.y, x.y, fmaf(r2.z, x.z, r2.w)));
  return atan2_as(sqrtf(lx * lx + ly * ly), lz);
}

// a deposit's charge cells: hat weights on the QE wavelength knots and on the CE theta knots (linear between
// knots, held beyond the ends; the reference's readout/grid.py `hats`, `cells`): cell il * Kt + it and its
// neighbours, weights (1 - fl)(1 - ft), fl (1 - ft), (1 - fl) ft, fl ft; and the acceptance QE(lambda) CE(theta)
struct Cells {
  int base;      // the first cell (il * Kt + it)
  float fl, ft;  // the weights of the next wavelength and theta knots
  float acc;     // QE(lambda) CE(theta) (1 without tables)
};

__device__ inline Cells charge_cells(const Args &a, float lam, float theta) {
  Cells c{0, 0.0f, 0.0f, 1.0f};
  if (a.Kl > 1) {
    const TabPos tp = tab_pos(a.P, a.L[L_QE_TABLE], a.Kl, lam);
    c.base = tp.i * a.Kt;
    c.fl = tp.f;
    c.acc = tab_at(a.P, a.L[L_QE_TABLE], a.Kl, tp);
  }
  if (a.Kt > 1) {
    const TabPos tp = tab_pos(a.P, a.L[L_CE_TABLE], a.Kt, theta);
    c.base += tp.i;
    c.ft = tp.f;
    c.acc *= tab_at(a.P, a.L[L_CE_TABLE], a.Kt, tp);
  }
  return c;
}

// one deposit at sensor s: arrival, arrival_t, the charge cells on the QE and CE knots, the photon's moment
template <unsigned F>
__device__ inline void deposit(const Args &a, Photon &p, Acc &acc, int s, float q, float time, float theta) {
  acc.records += 1;
  const float v = q * p.w0;
  if (v == 0.0f) return;
  const double dv = (double)v;
  tally_add(a, p.event, s, dv);
  tally_add(a, p.event, a.S + s, dv * (double)time);
  const Cells c = charge_cells(a, p.lam, theta);
  const long long cb = 3LL * a.S + (long long)s * a.K + c.base;
  tally_add(a, p.event, cb, dv * (double)((1.0f - c.fl) * (1.0f - c.ft)));
  if (c.fl > 0.0f) tally_add(a, p.event, cb + a.Kt, dv * (double)(c.fl * (1.0f - c.ft)));
  if (c.ft > 0.0f) {
    tally_add(a, p.event, cb + 1, dv * (double)((1.0f - c.fl) * c.ft));
    if (c.fl > 0.0f) tally_add(a, p.event, cb + a.Kt + 1, dv * (double)(c.fl * c.ft));
  }
  if ((F & F_TIME) && a.nt > 0) {  // the time histogram (readout/grid.py `time_cells`): cubic Lagrange weights on the four nodes
    // around the time, the shares beyond the last / before the first node kept as (q, q t) per sensor
    const double dq = dv * (double)c.acc;  // weighted by the acceptance: the detected light's time shape
    const float x = fminf(fmaxf((time - a.t0) / a.dt, -4.0f), (float)a.nt + 4.0f);
    const float jf = floorf(x), g = x - jf;
    const int j = (int)jf;
    const float wk[4] = {-g * (g - 1.0f) * (g - 2.0f) / 6.0f, (g + 1.0f) * (g - 1.0f) * (g - 2.0f) / 2.0f,
                         -(g + 1.0f) * g * (g - 2.0f) / 2.0f, (g + 1.0f) * g * (g - 1.0f) / 6.0f};
    const long long tb = a.time_at + (long long)s * a.nt;
    float late = 0.0f, early = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int n = j - 1 + k;
      if (n >= a.nt) {
        late += wk[k];
      } else if (n < 0) {
        early += wk[k];
      } else {
        tally_add(a, p.event, tb + n, dq * (double)wk[k]);
      }
    }
    const long long ob = a.time_at + (long long)a.S * a.nt + 4LL * s;
    if (late != 0.0f) {
      tally_add(a, p.event, ob, dq * (double)late);
      tally_add(a, p.event, ob + 1, dq * (double)late * (double)time);
    }
    if (early != 0.0f) {
      tally_add(a, p.event, ob + 2, dq * (double)early);
      tally_add(a, p.event, ob + 3, dq * (double)early * (double)time);
    }
  }
  moment_add(a, p, s, v);
}

__device__ inline void photon_end(const Args &a, Photon &p) {
#pragma unroll
  for (int k = 0; k < MOMENT_SLOTS; ++k) moment_flush_slot(a, p, k);
}

// the source and local index of launch photon gi
__device__ __forceinline__ int source_of(const Args &a, long long gi) {
  int lo = 0, hi = a.nsrc - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a.src_end[mid] > gi) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// emission: a source row's photon (sources/point.py:73-86, sources/cherenkov.py:33-66), located (item.py:24-36)
template <bool TOR>
__device__ inline bool emit(const Args &a, long long gi, Photon &p, Acc &acc) {
  const int j = source_of(a, gi);
  const lh_source &S = a.src[j];
  const long long local = gi - (j > 0 ? a.src_end[j - 1] : 0);
  const unsigned long long id = (unsigned long long)(S.first_id + local);
  p.id_lo = (uint32_t)id;
  p.id_hi = (uint32_t)(id >> 32);
  const lh_rng rng{a.k0, a.k1, p.id_lo, p.id_hi};
  float u[4], v[4];
  rng.draw4(LH_SOURCE_VERTEX, 0, u);
  rng.draw4(LH_SOURCE_VERTEX, 1, v);
  p.x = mk3(S.pos[0], S.pos[1], S.pos[2]);
  p.t = S.time;
  p.w0 = S.weight;
  p.er = p.ei = mk3(0.f, 0.f, 0.f);
  p.event = S.event;
  if (S.kind == LH_SRC_CHERENKOV) {
    const int *M = a.L + a.L[L_MEDIA_AT] + MEDIUM_WORDS * S.medium;
    const float e_lo = a.P[M[0]], e_hi = a.P[M[0] + M[1] - 1];
    const float lam_min = HC / e_hi, lam_max = HC / e_lo;
    const float zq = 1.0f / lam_min - 1.0f / lam_max;
    p.lam = 1.0f / (1.0f / lam_min - u[0] * zq);
    const float n = tab(a.P, M[0], M[1], HC / p.lam);
    const float bn = S.beta * n;
    const bool emits = bn > 1.0f;
    const float s2 = emits ? 1.0f - 1.0f / (bn * bn) : 0.0f;
    p.w0 = TWO_PI * FINE_STRUCTURE * S.charge * S.charge * S.length * s2 * zq * 1e9f * S.weight;
    const float cos_c = emits ? 1.0f / bn : 1.0f;
    const float sin_c = sqrtf(fmaxf(1.0f - cos_c * cos_c, 0.0f));
    float sp, cp;
    __sincosf(TWO_PI * u[2], &sp, &cp);
    const float3 axis = mk3(S.dir[0], S.dir[1], S.dir[2]);
    p.d = unit3(to_world(mk3(sin_c * cp, sin_c * sp, cos_c), axis));
    const float xs = u[1] * S.length;
    p.x = fma3(xs, axis, p.x);
    p.t = S.time + xs / (S.beta * C_VACUUM);
    p.er = unit3(fma3(dot3(axis, p.d), p.d, mul3(-1.0f, axis)));
  } else {
    p.lam = S.lam_hi > S.lam ? S.lam + u[2] * (S.lam_hi - S.lam) : S.lam;
    if (S.kind == LH_SRC_BEAM) {
      p.d = unit3(mk3(S.dir[0], S.dir[1], S.dir[2]));
    } else {
      const float c = 1.0f - 2.0f * u[0];
      const float s = sqrtf(fmaxf(1.0f - c * c, 0.0f));
      float sp, cp;
      __sincosf(TWO_PI * u[1], &sp, &cp);
      p.d = mk3(s * cp, s * sp, c);
    }
  }
  p.w = 1.0f;
  p.wr = 1.0f;
  int vol = -1;
  const int hint = s_hint[threadIdx.x >> 6];
  if (hint >= 0 && hint < a.g.nvol) {
    float3 xl, dl;
    to_local(a.g, hint, p.x, mk3(0.f, 0.f, 1.f), xl, dl);
    if (core_holds(a.g, hint, xl)) vol = hint;
  }
  if (vol < 0) {
    vol = locate<TOR>(a.g, p.x);
    s_hint[threadIdx.x >> 6] = vol;
  }
  p.vol = vol;
  p.cls = a.g.vinfo[p.vol].z;
  p.srf = -1;
  p.xs = 0.0f;
  p.nz = 0;
  p.vertex = 0;
  p.mcls = -2;
  p.ms = i4v{-1, -1, -1, -1};
  p.mv = f4v{0.0f, 0.0f, 0.0f, 0.0f};
  acc.photons += 1;
  acc.start_w += p.w0;
  if (p.vol >= a.g.nvol || class_medium(a, p.cls) < 0) {  // emitted outside the detector or in an opaque medium
    rare_add(RF_LOST, p.w0);
    return false;
  }
  return true;
}

// Geant4's unified facet path (facets.py:318-397): facets drawn for the current direction and mirrored about
// until the direction leaves the surface; the polarisation mirrored about each facet's half vector.  The path is
// drawn at the reference width `sigma`; with `ratio` its log density ratio live / reference is added to logw
// (facets.py `facet_logpdf` per facet, surface_event.py:92-112), and with `deriv` the live log density's derivative
// in the width to dlogw.
__device__ inline float3 facet_path(const Photon &p, const lh_rng &rng, float sigma, float sigma_live, const float *__restrict__ T,
                                    bool ratio, bool deriv, float3 wi, float3 n, float3 &er, float3 &ei, float &logw,
                                    float &dlogw) {
  float3 V = to_local_frame(mul3(-1.0f, wi), n);
  float3 w = wi;
  const float fmax = fminf(1.0f, 4.0f * sigma);
  for (int k = 0; k < MAX_FACETS; ++k) {
    float a = 0.0f, u2 = 0.5f;
    for (int t = 0; t < MAX_FACET_ATTEMPTS; ++t) {
      float u[4];
      rng.draw4(p.vertex, LH_FACET_BLOCK + ((uint32_t)k << 12) + (uint32_t)t, u);
      const float al = sigma * sqrtf(-2.0f * __logf(u[0]));
      if (al < HALF_PI && u[1] * al * PI_F < fminf(__sinf(al), fmax) * front_arc(V, al)) {
        a = al;
        u2 = u[2];
        break;
      }
    }
    const float psi = (2.0f * u2 - 1.0f) * front_arc(V, a);
    const float ph = atan2_as(V.y, V.x) + psi;
    float sa, ca, sp, cp;
    __sincosf(a, &sa, &ca);
    __sincosf(ph, &sp, &cp);
    const float3 m = mk3(sa * cp, sa * sp, ca);
    if (ratio || deriv) {
      float dl, dr;
      const float lp = unified_lp(T, sigma_live, V, m, dl);
      if (ratio) logw += lp - unified_lp(T, sigma, V, m, dr);
      if (deriv) dlogw += dl;
    }
    const float3 Lo = fma3(2.0f * dot3(V, m), m, mul3(-1.0f, V));
    const float3 wo = to_world(Lo, n);
    pol_mirror(er, ei, unit3(sub3(wo, w)));
    w = wo;
    V = mul3(-1.0f, Lo);
    if (Lo.z >= 0.0f) break;
  }
  return w;
}

// GLISUR's facet (G4OpBoundaryProcess::GetFacetNormal, glisur; facets.py:14-40): M = z + r s with s uniform in the
// unit ball (on the unit sphere when `shell`, the cover's share), redrawn while V . M <= 0; the unit M.  Local frame
__device__ inline float3 glisur_facet(const lh_rng &rng, uint32_t vertex, int k, float r, float3 V, bool shell) {
  float3 M = mk3(0.0f, 0.0f, 1.0f);
  for (int t = 0; t < MAX_FACET_ATTEMPTS; ++t) {
    float u[4];
    rng.draw4(vertex, LH_FACET_BLOCK + ((uint32_t)k << 12) + (uint32_t)t, u);
    const float z = 1.0f - 2.0f * u[1], sxy = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    const float rad = shell ? r : r * cbrtf(u[0]);
    float sp, cp;
    __sincosf(TWO_PI * u[2], &sp, &cp);
    const float3 c = mk3(rad * sxy * cp, rad * sxy * sp, fmaf(rad, z, 1.0f));
    if (dot3(V, c) > 0.0f) {
      M = c;
      break;
    }
  }
  return unit3(M);
}

// the facet's log density per solid angle log p_m, p_m = 2 q (3 c^2 + q^2) / (3 Vol) (c = m_z, q^2 = r^2 - sin^2, Vol
// the ball's volume on V's side), and in dl its derivative in r: 3 r (c^2 + q^2) / (q^2 (3 c^2 + q^2)) - Area / Vol
// (d Vol / dr is the ball's surface on V's side; facets.py `_glisur_log_density`).  -inf (dl 0) off the support.
// With `cover`, lq receives the cover mixture's log density (1 - e) p_m + e p_shell, p_shell = 2 r (c^2 + q^2) /
// (q Area) (facets.py `_glisur_log_proposal`)
__device__ inline float glisur_lp(float r, float3 V, float3 m, float &dl, bool cover, float &lq) {
  const float c = m.z, q2 = r * r - (m.x * m.x + m.y * m.y);
  dl = 0.0f;
  lq = -INFINITY;
  if (!(q2 > 0.0f && c > 0.0f && dot3(V, m) > 0.0f)) return -INFINITY;
  const float h = fmaxf(r - V.z, 0.0f);
  const float vol = (4.0f / 3.0f) * PI_F * r * r * r - PI_F * h * h * (2.0f * r + V.z) * (1.0f / 3.0f);
  const float area = V.z >= r ? 4.0f * PI_F * r * r : TWO_PI * r * (r + fminf(V.z, r));
  const float b = 3.0f * c * c + q2;
  dl = 3.0f * r * (c * c + q2) / (q2 * b) - area / vol;
  const float lp = __logf(2.0f * q2 * b / (3.0f * vol)) - 0.5f * __logf(q2);  // log(2 q b / (3 Vol))
  if (cover) {
    const float ps = 2.0f * r * (c * c + q2) * rsqrtf(q2) / area;
    lq = __logf((1.0f - DEFENSIVE) * __expf(lp) + DEFENSIVE * ps);
  }
  return lp;
}

// one facet's weight terms: the live log density over the density drawn from (the reference width's, or the cover
// mixture's) into logw when `weigh`, the live log density's derivative in r into dlogw when `deriv`.  A facet that
// rounds off the drawn density's support weighs 0 (surface_event.py:124-129)
__device__ inline void glisur_weigh(float r, float r_live, bool cover, bool weigh, bool deriv, float3 V, float3 m,
                                    float &logw, float &dlogw) {
  if (!(weigh || deriv)) return;
  float dl, lq, dr, lqr;
  const float lp = glisur_lp(r_live, V, m, dl, false, lq);
  if (weigh) {
    const float ld = cover ? (glisur_lp(r, V, m, dr, true, lqr), lqr) : glisur_lp(r, V, m, dr, false, lqr);
    logw = ld > -INFINITY ? logw + (lp - ld) : -INFINITY;
  }
  if (deriv) dlogw += dl;
}

// Geant4's DielectricMetal loop with GLISUR facets (facets.py:431-470): a facet for the current direction, the mirror
// reflection about it, repeated while the direction points into the surface; the state mirrored about each facet.
// Drawn at width r (from the cover mixture when `cover`), weighed at r_live
__device__ inline float3 glisur_metal(const lh_rng &rng, uint32_t vertex, float r, float r_live, bool cover,
                                      bool weigh, bool deriv, float3 wi, float3 n, float3 &er, float3 &ei,
                                      float &logw, float &dlogw) {
  float3 w = wi;
  for (int k = 0; k < MAX_FACETS; ++k) {
    float u[4];
    rng.draw4(vertex, LH_GLISUR_BLOCK + (uint32_t)k, u);
    const float3 V = to_local_frame(mul3(-1.0f, w), n);
    const float3 m = glisur_facet(rng, vertex, k, r, V, cover && u[0] < DEFENSIVE);
    glisur_weigh(r, r_live, cover, weigh, deriv, V, m, logw, dlogw);
    const float3 mw = to_world(m, n);
    w = fma3(-2.0f * dot3(w, mw), mw, w);
    pol_mirror(er, ei, mw);
    if (dot3(w, n) >= 0.0f) break;
  }
  return w;
}

// Geant4's DielectricDielectric loop with GLISUR facets (facets.py:471-546): a facet about the current side's normal,
// reflected or refracted by the Fresnel shares of the photon's state there (its own uniform; total reflection always
// reflects), the state recombined into the new direction; a refraction that still points back at the surface swaps
// the side; ends when the direction leaves the surface on either side.  `far`: an odd number of refractions (the
// photon ends beyond the surface).  n1 the index on n's side, n2 the far one; matched indices keep the direction
__device__ inline float3 glisur_dielectric(const lh_rng &rng, uint32_t vertex, float r, float r_live, bool cover,
                                           bool weigh, bool deriv, float3 wi, float3 n, float n1, float n2,
                                           float3 &er, float3 &ei, float &logw, float &dlogw, bool &far) {
  float3 w = wi;
  float side = 1.0f;
  far = false;
  const bool matched = fabsf(n1 - n2) <= INDEX_MATCH * fmaxf(n1, n2);
  for (int k = 0; k < MAX_FACETS; ++k) {
    const float3 N = mul3(side, n);
    const float na = side > 0.0f ? n1 : n2, nb = side > 0.0f ? n2 : n1;
    float u[4];
    rng.draw4(vertex, LH_GLISUR_BLOCK + (uint32_t)k, u);
    const float3 V = to_local_frame(mul3(-1.0f, w), N);
    const float3 ml = glisur_facet(rng, vertex, k, r, V, cover && u[0] < DEFENSIVE);
    glisur_weigh(r, r_live, cover, weigh, deriv, V, ml, logw, dlogw);
    const float3 m = to_world(ml, N);
    const float c1 = clip01(-dot3(w, m));
    const Amps f = fresnel(c1, na, nb);
    const float3 sx = s_axis(w, m), px = cross3(w, sx);
    const float es2 = dot3(er, sx) * dot3(er, sx) + dot3(ei, sx) * dot3(ei, sx);
    const float ep2 = dot3(er, px) * dot3(er, px) + dot3(ei, px) * dot3(ei, px);
    const float Rf = clip01(cabs2(f.rs) * es2 + cabs2(f.rp) * ep2) / fmaxf(es2 + ep2, FLOOR32);
    const float eta = na / nb;
    const bool tir = eta * eta * (1.0f - c1 * c1) >= 1.0f;
    const bool refr = !tir && u[1] >= Rf;
    const float3 w1 = refr ? (matched ? w : refract_dir(w, m, eta)) : fma3(-2.0f * dot3(w, m), m, w);
    if (!(refr && matched)) pol_recombine(er, ei, w, sx, refr ? f.ts : f.rs, refr ? f.tp : f.rp, w1);
    const float d = dot3(w1, N);
    far = far != refr;
    w = w1;
    if (refr ? d <= 0.0f : d >= 0.0f) break;
    if (refr) side = -side;
  }
  return w;
}

template <bool ADJ, unsigned F>
__device__ inline void medium_cache(const Args &a, Photon &p, const AdjArgs *b) {
  constexpr bool REF = F & F_REF, CHAN = F & F_CHAN;
  if (p.mcls == p.cls) return;
  const int m = class_medium(a, p.cls);
  p.ml = medium_at(a.L, a.P, m, p.lam);
  if (REF) p.mr = medium_at(a.L, a.Pr, m, p.lam);
  p.mcls = p.cls;
  if (ADJ) {
    p.da = RateD{-1, -1, 0.0f, 0.0f};
    p.dr = RateD{-1, -1, 0.0f, 0.0f};
    p.dm = RateD{-1, -1, 0.0f, 0.0f};
    p.dw = RateD{-1, -1, 0.0f, 0.0f};
    if (m >= 0) {
      const int *M = a.L + a.L[L_MEDIA_AT] + MEDIUM_WORDS * m;
      if (!(M[8] & MED_OPAQUE)) {
        const float E = HC / p.lam;
        if (M[5] > 0) p.da = rate_d(a.P, b->slots, M[4], M[5], E);
        if (M[7] > 0) p.dr = rate_d(a.P, b->slots, M[6], M[7], E);
        if (CHAN && (M[8] & MED_MIE)) p.dm = rate_d(a.P, b->slots, M[17], M[18], E);
        if (CHAN && (M[8] & MED_WLS)) p.dw = rate_d(a.P, b->slots, M[9], M[10], E);
      }
    }
  }
}

// the cotangent of one deposit at sensor s: dLoss/d(arrival) plus the charge cells' at its wavelength and angle
__device__ inline float deposit_cotangent(const Args &a, const AdjArgs *b, const Photon &p, int s, float theta) {
  const double *c = b->cot + (long long)p.event * a.tally_stride;
  const Cells cl = charge_cells(a, p.lam, theta);
  const long long cb = 3LL * a.S + (long long)s * a.K + cl.base;
  double cw = c[s] + (double)((1.0f - cl.fl) * (1.0f - cl.ft)) * c[cb];
  if (cl.fl > 0.0f) cw += (double)(cl.fl * (1.0f - cl.ft)) * c[cb + a.Kt];
  if (cl.ft > 0.0f) {
    cw += (double)((1.0f - cl.fl) * cl.ft) * c[cb + 1];
    if (cl.fl > 0.0f) cw += (double)(cl.fl * cl.ft) * c[cb + a.Kt + 1];
  }
  return (float)cw;
}

__device__ __forceinline__ float ratio(float live, float ref) { return ref != 0.0f ? live / ref : 0.0f; }

// ------------------------------------------------------------------------------------------- the diffusion jump
// At a Rayleigh event in a medium with a clearance grid, with a linear polarisation (ei = 0): the photon walks its
// own scatterings inside the ball B(x, c) of the proven clearance c at x (the grid's nearest node less the distance
// to it), with no geometry query: free paths Exp(mu_ref), the polarised Rayleigh draw and projection, until a flight
// leaves the ball (the photon stops at the crossing) or the walk has made ceil((c mu_ref)^2) scatterings
// (fastpath.py:166-224 `diffuse`).  Its uniforms are the photon's own, vertex word with the top bit set, block the
// walk's step: no other draw uses such a counter.
struct Jump {
  float3 y, d, e;  // the displacement, the direction and the polarisation after the walk
  float L;         // the path length (0: no walk)
  int n;           // the scatterings
};

__device__ inline Jump diffuse(const Args &a, const Photon &p, const lh_rng &rng, int medium, float mu_r) {
  Jump J{mk3(0.f, 0.f, 0.f), p.d, p.er, 0.0f, 0};
  if (medium < 0 || !(mu_r > 0.0f) || p.ei.x != 0.0f || p.ei.y != 0.0f || p.ei.z != 0.0f) return J;
  const float4 G0 = a.jrec[2 * medium], G1 = a.jrec[2 * medium + 1];
  const int off = __float_as_int(G1.w);
  if (off < 0) return J;
  const int n0 = __float_as_int(G1.x), n1 = __float_as_int(G1.y), n2 = __float_as_int(G1.z);
  const float fx = (p.x.x - G0.x) / G0.w, fy = (p.x.y - G0.y) / G0.w, fz = (p.x.z - G0.z) / G0.w;
  if (!(fx > -0.5f && fx < (float)n0 - 0.5f && fy > -0.5f && fy < (float)n1 - 0.5f && fz > -0.5f &&
        fz < (float)n2 - 0.5f))
    return J;
  const int i0 = min(max((int)floorf(fx + 0.5f), 0), n0 - 1), i1 = min(max((int)floorf(fy + 0.5f), 0), n1 - 1),
            i2 = min(max((int)floorf(fz + 0.5f), 0), n2 - 1);
  const float3 node = mk3(fmaf(G0.w, (float)i0, G0.x), fmaf(G0.w, (float)i1, G0.y), fmaf(G0.w, (float)i2, G0.z));
  const float3 dx = sub3(p.x, node);
  const float c = a.jclear[off + ((long long)i0 * n1 + i1) * n2 + i2] - sqrtf(dot3(dx, dx));
  if (!(c > 0.0f)) return J;
  const float cm = c * mu_r;
  const int budget = (int)ceilf(fminf(cm * cm, 1073741824.0f));
  const float3 zero = mk3(0.f, 0.f, 0.f);
  for (int j = 0; j < budget; ++j) {
    float u[4];
    rng.draw4((uint32_t)p.vertex | 0x80000000u, (uint32_t)j, u);
    const float s = -__logf(u[0]) / mu_r;
    const float by = dot3(J.y, J.d);
    const float t = -by + sqrtf(fmaxf(by * by - (dot3(J.y, J.y) - c * c), 0.0f));  // the flight's exit of the ball
    if (s >= t) {
      J.y = fma3(t, J.d, J.y);
      J.L += t;
      break;
    }
    float3 e2 = J.e, z = zero;
    const float3 k = rayleigh_sample(J.e, zero, u[1], u[2], u[3]);
    pol_project(e2, z, k);
    J.y = fma3(s, J.d, J.y);
    J.L += s;
    J.d = unit3(k);
    J.e = e2;
    J.n += 1;
  }
  return J;
}

// the derivative terms of the rates over one segment: k_c d(mu_c)/d(value) v per rate (absorption, Rayleigh, Mie,
// WLS), through the photon's cached table derivatives
template <bool CHAN>
__device__ __forceinline__ void rates_add(double *sg, const Photon &p, float ka, float kr, float km, float kw,
                                          float v) {
  grad_add(sg, p.da.s0, ka * p.da.d0 * v);
  grad_add(sg, p.da.s1, ka * p.da.d1 * v);
  grad_add(sg, p.dr.s0, kr * p.dr.d0 * v);
  grad_add(sg, p.dr.s1, kr * p.dr.d1 * v);
  if (CHAN) {
    grad_add(sg, p.dm.s0, km * p.dm.d0 * v);
    grad_add(sg, p.dm.s1, km * p.dm.d1 * v);
    grad_add(sg, p.dw.s0, kw * p.dw.d0 * v);
    grad_add(sg, p.dw.s1, kw * p.dw.d1 * v);
  }
}

// one step of one photon (transport/step.py:516-571); returns whether it goes on.  ADJ: the adjoint's walk
// (no tallies; its pass 0 totals the photon's cotangent-weighted deposits, its pass 1 adds the derivative terms
// into the block's accumulators sg)
template <bool ADJ, unsigned F>
__device__ inline bool step(const Args &a, Photon &p, Acc &acc, const AdjArgs *b, double *sg) {
  constexpr bool REF = F & F_REF, CHAN = F & F_CHAN;
  const Geo &g = a.g;
  float b0 = band_at(maxabs3(p.x));
  if (p.nz >= ZERO_STEPS_BEFORE_PUSH) {  // G4Navigator's stuck-track push (step.py:269-282 `_recover`)
    if (p.nz == ZERO_STEPS_BEFORE_PUSH) rare_add(RF_PUSHED, p.w * p.w0);
    p.x = fma3(PUSH_BANDS * b0, p.d, p.x);
    p.vol = locate<(F & F_TORUS) != 0>(g, p.x);
    p.cls = g.vinfo[p.vol].z;
    p.srf = -1;
    b0 = band_at(maxabs3(p.x));
  }
  float u[4], ul[4];
  const lh_rng rng{a.k0, a.k1, p.id_lo, p.id_hi};
  rng.draw4(p.vertex, 0, u);   // path, channel, outcome, roulette
  rng.draw4(p.vertex, 1, ul);  // lobe (3), polarisation
  float3 er0 = p.er, ei0 = p.ei;
  pol_fill(p.d, er0, ei0, ul[3]);
  if (a.born_at && p.vertex == 0) {  // emitted inside a sensor volume that counts by volume (step.py:574-612 `_born`)
    const int br = a.L[a.born_at + p.cls], sensor = g.vinfo[p.vol].w;
    if (br >= 0 && sensor >= 0) {  // the entry's detected share, deposited at the emission; the photon ends there
      const SurfEval se = surface_eval(a.L, a.P, br, p.lam, p.d, mul3(-1.0f, p.d), 1.0f, 1.0f, er0, ei0);
      const float dep_q = p.w * se.detect;
      const float theta = a.Kt > 1 ? cathode_theta(g, p.vol, p.x) : 0.0f;
      if (!ADJ) {
        deposit<F>(a, p, acc, sensor, dep_q, p.t, theta);
        acc.surf += p.w * p.w0;
      } else {
        const float D = dep_q * p.w0 * deposit_cotangent(a, b, p, sensor, theta);
        if (D != 0.0f) {
          p.hasdep = true;
          if (p.pass == 1) {
            float dsh[N_COL], ddet[N_COL];
            const int *Rw = a.L + a.L[L_ROWS_AT] + ROW_WORDS * br;
            surface_dlogs(Rw, a.P, HC / p.lam, -1, dsh, ddet);
            add_columns(sg, Rw, a.P, b->slots, HC / p.lam, ddet, D);
          }
          p.Dsum += D;
        }
      }
      p.w = p.wr = 0.0f;
      p.vertex += 1;
      acc.vertices += 1;
      return false;
    }
  }
  medium_cache<ADJ, F>(a, p, b);
  const MediumAt ml = p.ml, mr = REF ? p.mr : p.ml;
  // the total rates, summed in Geant4's channel order (Rayleigh, Mie, WLS)
  const float mu = CHAN ? (ml.mu_ray + ml.mu_mie) + ml.mu_wls : ml.mu_ray;
  const float mu_r = CHAN ? (mr.mu_ray + mr.mu_mie) + mr.mu_wls : mr.mu_ray;
  const float s = -__logf(u[0]) / fmaxf(mu_r, FLOOR32);  // the free path at the reference rate
  const float max_dist = a.max_dist;
  float q_lim = max_dist;
  if constexpr ((F & F_REACH) != 0) {  // the reach bound: not on a face, in a volume
    if (p.srf < 0 && p.vol < g.nvol) q_lim = reach_limit(g, p.x, p.d, p.vol, s, max_dist);
    const unsigned long long bm = __ballot(q_lim < max_dist);
    if (bm != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)bm) - 1) atomicAdd(&s_bounded, (unsigned)__popcll(bm));
  }
  const Cross cr = query_t<!(F & F_GNODES), (F & F_TORUS) != 0>(g, p.x, p.d, p.vol, p.srf, q_lim, p.xs);
  const bool hit = cr.vol_after >= 0;
  const float dist = cr.t;
  const bool zero = dist <= ZERO_STEP_BANDS * b0;
  const int cb = p.cls;
  const int ca = hit ? g.vinfo[cr.vol_after].z : -1;
  int row = hit ? a.L[a.L[L_ROW_OF_AT] + cb * a.L[L_CLASSES] + ca] : -1;
  // a crossing no surface joins, of zero length or between two volumes of one class (the same medium and surfaces,
  // which the reference's navigator does not stop at: lookup.py:26-31): a straight pass
  const bool cont = hit && (zero || ca == cb) && row < 0;
  if (cont) row = 0;
  const bool scatter = mu_r > 0.0f && s < dist;
  const float reach_live = __expf(-mu * dist), reach_ref = REF ? __expf(-mu_r * dist) : reach_live;
  const float att = __expf(-ml.mu_abs * dist), att_r = REF ? __expf(-mr.mu_abs * dist) : att;
  const float v_g = C_VACUUM / ml.ng;
  SurfEval sl, sr;
  int o = -1;
  float pprop[N_OUT], pdraw[N_OUT];
  const bool at_if = hit && row >= 0;
  if (at_if) {
    const int m2 = class_medium(a, ca);
    const float n2 = index_at(a.L, a.P, m2, p.lam);
    sl = surface_eval(a.L, a.P, row, p.lam, p.d, cr.n, ml.n, n2, er0, ei0);
    if (cont) {
      for (int k = 0; k < N_OUT; ++k) sl.p[k] = 0.0f;
      sl.p[STRAIGHT] = 1.0f;
      sl.detect = 0.0f;
    }
    if (REF) {
      const float n2r = index_at(a.L, a.Pr, m2, p.lam);
      sr = surface_eval(a.L, a.Pr, row, p.lam, p.d, cr.n, mr.n, n2r, er0, ei0);
      if (cont) {
        for (int k = 0; k < N_OUT; ++k) sr.p[k] = 0.0f;
        sr.p[STRAIGHT] = 1.0f;
        sr.detect = 0.0f;
      }
    } else {
      sr = sl;
    }
    // the draw's probabilities: the reference's, or at an edge row the one defensive mixture over the row's
    // support (estimator.py:96-153); the proposal normalises them
    float tot = 0.0f;
    int ncov = 0;
    const int support = (sr.flags >> 8) & 63;
    const bool edge = (sr.flags & FLAG_EDGE) && !cont;
    for (int k = 0; k < N_OUT; ++k) {
      pdraw[k] = sr.p[k];
      if (edge && ((support >> k) & 1) && !(sr.p[k] > 0.0f)) ++ncov;
    }
    if (ncov > 0) {
      const float eps = DEFENSIVE / (1.0f - DEFENSIVE) / (float)ncov;
      for (int k = 0; k < N_OUT; ++k)
        if (((support >> k) & 1) && !(sr.p[k] > 0.0f)) pdraw[k] = eps;
    }
    for (int k = 0; k < N_OUT; ++k) tot += pdraw[k];
    for (int k = 0; k < N_OUT; ++k) pprop[k] = tot > 0.0f ? pdraw[k] / tot : 0.0f;
    o = choose6(pprop, u[2]);
    // the deposit: the expected detection at the crossing (step.py:522-524)
    const float dep_q = p.w * reach_live * att * sl.detect;
    // the sensor entered, or across an exit sensor's border the region left (step.py:544-547 `_sensor_dep`)
    const int vs = (a.left_at && a.L[a.left_at + cb * a.L[L_CLASSES] + ca]) ? p.vol : cr.vol_after;
    const int sensor = g.vinfo[vs].w;
    if (sensor >= 0 && !cont) {
      const float theta = a.Kt > 1 ? cathode_theta(g, vs, fma3(dist, p.d, p.x)) : 0.0f;
      if (!ADJ) {
        deposit<F>(a, p, acc, sensor, dep_q, p.t + dist / v_g, theta);
      } else {
        const float D = dep_q * p.w0 * deposit_cotangent(a, b, p, sensor, theta);
        if (D != 0.0f) {
          p.hasdep = true;
          if (p.pass == 1) {  // the deposit's own factors: reach, attenuation, the detected share
            rates_add<CHAN>(sg, p, -dist, -dist, -dist, -dist, D);
            float dsh[N_COL], ddet[N_COL];
            const int *Rw = a.L + a.L[L_ROWS_AT] + ROW_WORDS * row;
            surface_dlogs(Rw, a.P, HC / p.lam, -1, dsh, ddet);
            add_columns(sg, Rw, a.P, b->slots, HC / p.lam, ddet, D);
          }
          p.Dsum += D;
        }
      }
    }
  }
  // the deposits still ahead of this step's event (the adjoint's pass 1)
  const float ahead = ADJ ? p.T - p.Dsum : 0.0f;
  const bool adj_event = ADJ && p.pass == 1 && ahead != 0.0f;
  bool ok;
  float w_new, wr_new;
  if (scatter) {  // a bulk event at the free path's end (step.py:397-436, `_bulk`)
    // the channel drawn from the reference rates in Geant4's order (Rayleigh, Mie, WLS; estimator.py `choose`:
    // the first whose running sum exceeds u times the total, Rayleigh if rounding leaves none), weighted by the
    // live / reference density of the free path in it, mu_c exp(-mu s), times the absorption over the path
    int ch = 0;
    if (CHAN) {
      const float xc = u[1] * mu_r, c1 = mr.mu_ray + mr.mu_mie;
      ch = xc < mr.mu_ray ? 0 : xc < c1 ? 1 : xc < mu_r ? 2 : 0;
    }
    const float mc = ch == 0 ? ml.mu_ray : ch == 1 ? ml.mu_mie : ml.mu_wls;
    const float mc_r = ch == 0 ? mr.mu_ray : ch == 1 ? mr.mu_mie : mr.mu_wls;
    const float w_path = ratio(mc * __expf(-mu * s), mc_r * __expf(-mu_r * s));
    const float att_s = __expf(-ml.mu_abs * s);
    const float w_at = p.w * w_path;
    const float att_s_r = REF ? __expf(-mr.mu_abs * s) : att_s;
    if (adj_event) {  // the free path's density mu_c exp(-mu s) and the absorption exp(-mu_abs s)
      const float kc = mc > 0.0f ? 1.0f / mc : 0.0f;
      rates_add<CHAN>(sg, p, -s, (ch == 0 ? kc : 0.0f) - s, (ch == 1 ? kc : 0.0f) - s, (ch == 2 ? kc : 0.0f) - s,
                      ahead);
    }
    p.x = fma3(s, p.d, p.x);
    p.t += s / v_g;
    if (ch != 2) {  // a scattering: Rayleigh's polarised dipole lobe (phase.py:45-75) or Mie's two HG lobes
      float3 dn;
      float w_lobe = 1.0f;
      if (ch == 0) {
        dn = rayleigh_sample(er0, ei0, ul[0], ul[1], ul[2]);
      } else {  // G4OpMieHG, drawn at the reference's lobes under the cover of its forward ratio, weighted by the
                // live mixture density over the drawn one (transport/medium.py:96-133 `channel_draw`)
        const int *M = a.L + a.L[L_MEDIA_AT] + MEDIUM_WORDS * class_medium(a, p.cls);
        const MieLobe lv = mie_lobe(a.P, M), lr = REF ? mie_lobe(a.Pr, M) : lv;
        const float rc = mie_cover(lr.r);
        dn = mie_sample(p.d, lr.gf, lr.gb, rc, ul[0], ul[1], ul[2]);
        const float3 dmi = sub3(dn, p.d), dpl = add3(dn, p.d);
        const float cm = 0.5f * dot3(dmi, dmi), cp = 0.5f * dot3(dpl, dpl);
        const float f_l = hg_density(lv.gf, cm, cp), b_l = hg_density(-lv.gb, cm, cp);
        const float live = lv.r * f_l + (1.0f - lv.r) * b_l;
        const float drawn = (REF || rc != lv.r) ? mie_density(lr.gf, lr.gb, rc, cm, cp) : live;
        w_lobe = ratio(live, drawn);
        if (adj_event && live > 0.0f) {  // the live mixture's log density in (g_f, g_b, r)
          const float k = ahead / live;
          grad_add(sg, b->slots[M[19] + M[20]], lv.r * f_l * hg_dlog_g(lv.gf, cm, cp) * k);
          grad_add(sg, b->slots[M[21] + M[22]], -(1.0f - lv.r) * b_l * hg_dlog_g(-lv.gb, cm, cp) * k);
          grad_add(sg, b->slots[M[23] + M[24]], (f_l - b_l) * k);
        }
      }
      w_new = w_at * att_s * w_lobe;
      wr_new = p.wr * att_s_r;
      acc.bulk += w_at * (1.0f - att_s) * p.w0;
      pol_project(er0, ei0, dn);
      p.d = unit3(dn);
      p.er = er0;
      p.ei = ei0;
      if (ch == 0 && a.jrec) {  // the diffusion jump (fastpath.py `diffuse`), in a medium with a clearance grid
        const Jump J = diffuse(a, p, rng, class_medium(a, p.cls), mr.mu_ray);
        if (J.L > 0.0f) {
          // the walk's likelihood ratio and absorption: exp(-mu_abs L) (mu / mu_ref)^n exp(-(mu - mu_ref) L)
          const float wj = __expf(-ml.mu_abs * J.L) *
                           (REF ? __expf((float)J.n * __logf(ml.mu_ray / mr.mu_ray) - (ml.mu_ray - mr.mu_ray) * J.L)
                                : 1.0f);
          acc.bulk += w_new * (1.0f - wj) * p.w0;  // the walk's factor is its path's: absorbed
          w_new *= wj;
          wr_new *= REF ? __expf(-mr.mu_abs * J.L) : __expf(-ml.mu_abs * J.L);
          if (adj_event) rates_add<CHAN>(sg, p, -J.L, (float)J.n / ml.mu_ray - J.L, 0.0f, 0.0f, ahead);
          p.x = add3(p.x, J.y);
          p.d = J.d;
          p.er = J.e;
          p.t += J.L / v_g;
          p.srf = -1;
        }
      }
    } else {  // a wavelength shifter (G4OpWLS): re-emitted isotropic, unpolarised, after the record's delay
      float ur[4];
      rng.draw4(p.vertex, 2, ur);  // (wavelength, delay, -, -)
      const int *M = a.L + a.L[L_MEDIA_AT] + MEDIUM_WORDS * class_medium(a, p.cls);
      const WlsDraw em = wls_emit(a.P, REF ? a.Pr : a.P, M, p.lam, ur[0], ur[1]);
      w_new = w_at * att_s * em.w;
      wr_new = p.wr * att_s_r * em.wr;
      acc.bulk += w_at * (1.0f - att_s * em.w) * p.w0;
      const float c = 1.0f - 2.0f * ul[0];
      const float sn = sqrtf(fmaxf(1.0f - c * c, 0.0f));
      float sp, cp;
      __sincosf(TWO_PI * ul[1], &sp, &cp);
      p.d = unit3(to_world(mk3(sn * cp, sn * sp, c), p.d));
      p.er = p.ei = mk3(0.f, 0.f, 0.f);
      p.lam = em.lam;
      p.t += em.delay;
      p.mcls = -2;  // the medium's values at the new wavelength
    }
    const bool held = s <= ZERO_STEP_BANDS * b0 && p.srf >= 0;
    if (!held) p.srf = -1;
    ok = true;
  } else if (hit) {  // the surface's drawn outcome (step.py:362-394)
    const float w_at = p.w * ratio(reach_live, reach_ref);
    float s_cont = 0.0f;
    if (at_if)
      for (int k = 0; k < N_OUT; ++k) s_cont += sl.p[k];
    acc.bulk += w_at * (1.0f - att) * p.w0;
    if (at_if) acc.surf += w_at * att * (1.0f - s_cont) * p.w0;
    ok = at_if && o >= 0;
    if (!at_if) rare_add(RF_LOST, w_at * att * p.w0);
    if (ok && adj_event) {  // reach, attenuation over the segment, and the drawn outcome's share
      rates_add<CHAN>(sg, p, -dist, -dist, -dist, -dist, ahead);
      float dsh[N_COL], ddet[N_COL];
      const int *Rw = a.L + a.L[L_ROWS_AT] + ROW_WORDS * row;
      if (!cont) {
        surface_dlogs(Rw, a.P, HC / p.lam, o, dsh, ddet);
        add_columns(sg, Rw, a.P, b->slots, HC / p.lam, dsh, ahead);
      }
    }
    if (ok) {
      const float w_if = w_at * att * ratio(pick6(sl.p, o), pick6(pprop, o));
      const float r_o = ratio(pick6(pdraw, o), pick6(pprop, o));
      wr_new = p.wr * (reach_ref != 0.0f ? 1.0f : 0.0f) * att_r * r_o;
      float3 dn = p.d;
      const float3 wi = p.d, n = cr.n;
      bool far = false;
      float w_lobe = 1.0f;  // the drawn lobe's live / reference density (a facet path at another width)
      if (o == STRAIGHT) {
        far = true;
      } else if (o == REFRACT) {
        const float m2n = index_at(a.L, a.P, class_medium(a, ca), p.lam);
        dn = refract_dir(wi, n, ml.n / m2n);
        pol_recombine(er0, ei0, wi, s_axis(wi, n), sl.amps.ts, sl.amps.tp, unit3(dn));
        far = true;
      } else if (o == SPECULAR) {
        dn = fma3(-2.0f * dot3(wi, n), n, wi);
        pol_recombine(er0, ei0, wi, s_axis(wi, n), sl.amps.rs, sl.amps.rp, unit3(dn));
      } else if (o == BACKSCATTER) {
        dn = mul3(-1.0f, wi);
        er0 = mul3(-1.0f, er0);
        ei0 = mul3(-1.0f, ei0);
      } else if (o == GLOSSY) {
        float logw = 0.0f, dlogw = 0.0f;
        const bool want_ratio = REF && sl.sigma != sr.sigma;
        const bool glisur = sr.flags & FLAG_GLISUR;
        if (glisur) {  // GLISUR's facet path at radius r = 1 - polish: a mirror metal's, or an interface's that
          // refracts too and may end beyond the surface (at the live indices, as a polished interface)
          const bool cover = sr.flags & FLAG_COVER, weigh = want_ratio || cover;
          if (sr.flags & FLAG_REFRACTS) {
            const float n2l = index_at(a.L, a.P, class_medium(a, ca), p.lam);
            dn = glisur_dielectric(rng, p.vertex, sr.sigma, sl.sigma, cover, weigh, adj_event, wi, n, ml.n, n2l,
                                   er0, ei0, logw, dlogw, far);
          } else {
            dn = glisur_metal(rng, p.vertex, sr.sigma, sl.sigma, cover, weigh, adj_event, wi, n, er0, ei0, logw,
                              dlogw);
          }
          if (weigh) w_lobe = __expf(logw);
        } else {
          dn = facet_path(p, rng, sr.sigma, sl.sigma, a.ut, want_ratio, adj_event, wi, n, er0, ei0, logw, dlogw);
          if (want_ratio) w_lobe = __expf(logw);
        }
        if (adj_event) {  // the live log density's derivative in the width: sigma_alpha, or -d/dr for polish
          const int col = glisur ? COL_POLISH : COL_SIGMA;
          const float sgn = glisur ? -1.0f : 1.0f;
          const int *Rw = a.L + a.L[L_ROWS_AT] + ROW_WORDS * row;
          const TabD t = tab_d(a.P, b->slots, Rw[2 + 2 * col], Rw[3 + 2 * col], HC / p.lam);
          grad_add(sg, t.s0, sgn * dlogw * t.w0 * ahead);
          grad_add(sg, t.s1, sgn * dlogw * t.w1 * ahead);
        }
      } else {  // DIFFUSE
        dn = lambert_sample(n, ul[0], ul[1]);
        pol_mirror(er0, ei0, unit3(sub3(dn, wi)));
      }
      w_new = w_if * w_lobe;
      acc.rew += (w_at * att * s_cont - w_new) * p.w0;
      p.xs = maxabs3(p.x) + dist;
      p.x = fma3(dist, p.d, p.x);
      p.t += dist / v_g;
      p.d = unit3(dn);
      p.er = er0;
      p.ei = ei0;
      p.srf = cr.srf;
      if (far) {
        p.vol = cr.vol_after;
        p.cls = ca;
      }
    } else {
      w_new = wr_new = 0.0f;
    }
  } else {  // nothing met within max_distance and no bulk event: the photon left (step.py:551-560)
    const float d_end = fminf(dist, max_dist);
    const float w_far = p.w * ratio(__expf(-mu * d_end), __expf(-mu_r * d_end));
    const float att_far = __expf(-ml.mu_abs * d_end);
    rare_add(max_dist < a.reach ? RF_DCUT : RF_ESC, w_far * att_far * p.w0);
    acc.bulk += w_far * (1.0f - att_far) * p.w0;
    ok = false;
    w_new = wr_new = 0.0f;
  }
  // the roulette on the reference weight, and the stuck rule (step.py:464-492)
  bool alive = ok;
  if (ok) {
    if (wr_new < a.roulette) {
      const float ps = wr_new / a.roulette;
      if (u[3] < ps) {
        const float wk = w_new / ps;
        acc.rou += (w_new - wk) * p.w0;
        w_new = wk;
        wr_new = wr_new / ps;
      } else {
        acc.rou += w_new * p.w0;
        w_new = wr_new = 0.0f;
        alive = false;
      }
    }
  }
  // a photon left without live weight (an edge row's floor drew an outcome of live probability 0) deposits nothing
  // more and adds 0 to every fate: it ends here unless the caller keeps it for the records and vertices counts
  if (alive && w_new == 0.0f && !a.keep_zero) alive = false;
  p.nz = (hit && !scatter && dist <= ZERO_STEP_BANDS * b0) ? p.nz + 1 : 0;
  p.w = w_new;
  p.wr = wr_new;
  p.vertex += 1;
  acc.vertices += 1;
  // written as values, not as a branch on which cell to add to: a pointer chosen at run time puts both in memory
  const bool abandon = alive && p.nz >= ZERO_STEPS_BEFORE_KILL;
  const bool truncate = alive && !abandon && p.vertex >= a.max_vert;
  const float wl = p.w * p.w0;
  if (abandon || truncate) rare_add(abandon ? RF_ABAND : RF_TRUNC, wl);
  return alive && !abandon && !truncate;
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// every thread of the block calls it (a barrier inside)
__device__ inline void flush_health(const Args &a, const Acc &acc) {
  const float vals[LH_HEALTH_N] = {acc.bulk, acc.surf, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, acc.rou, acc.rew, 0.0f,
                                   (float)acc.records, 0.0f, (float)acc.vertices, (float)acc.photons, acc.start_w,
                                   0.0f};
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < LH_HEALTH_N; ++k) {
    if (k == LH_H_LOST || k == LH_H_ESCAPED || k == LH_H_DISTANCE_CUT || k == LH_H_ABANDONED ||
        k == LH_H_TRUNCATED || k == LH_H_PUSHED || k == LH_H_MOMENT_SPILLS || k == 15)
      continue;
    const float s = wave_sum(vals[k]);
    if (lane == 0 && s != 0.0f) unsafeAtomicAdd(a.health + k, (double)s);
  }
  __syncthreads();
  if (threadIdx.x < RF_N) {
    const int cell[RF_N] = {LH_H_LOST, LH_H_ESCAPED, LH_H_DISTANCE_CUT, LH_H_ABANDONED, LH_H_TRUNCATED, LH_H_PUSHED,
                            LH_H_MOMENT_SPILLS};
    const float v = s_rare[threadIdx.x];
    if (v != 0.0f) unsafeAtomicAdd(a.health + cell[threadIdx.x], (double)v);
  }
  if (threadIdx.x == RF_N && s_bounded != 0u) unsafeAtomicAdd(a.health + LH_H_BOUNDED, (double)s_bounded);
}

__global__ void reset_queue(unsigned long long *q) { *q = 0ull; }

// the navigator on given rays: the next visible crossing (t, the navigator id entered, the normal) and the volume
// holding each ray's start
__global__ void query_kernel(Geo g, int n, const float *__restrict__ x, const float *__restrict__ d,
                             const int *__restrict__ vol, const int *__restrict__ last, float tmax,
                             float *__restrict__ t_out, int *__restrict__ vol_out, float *__restrict__ n_out,
                             int *__restrict__ srf_out, int *__restrict__ located) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float3 xi = mk3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
  const float3 di = unit3(mk3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
  const int v = vol[i] >= 0 ? vol[i] : locate<true>(g, xi);
  located[i] = locate<true>(g, xi);
  const Cross c = query<true>(g, xi, di, v, last[i], tmax);
  t_out[i] = c.t;
  vol_out[i] = c.vol_after;
  n_out[3 * i] = c.n.x;
  n_out[3 * i + 1] = c.n.y;
  n_out[3 * i + 2] = c.n.z;
  srf_out[i] = c.srf;
}

// the number of lanes below this one set in a wavefront mask
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// dynamic LDS of a launch: [the block's tallies (lds_tally) or gradient accumulators (the adjoint)][BVH nodes]
// the launch's arguments with the tables the launch placed in LDS copied there and pointed at (every thread of the
// block calls it; a barrier follows): the packed values (every table search reads them), the layout (rows, media,
// the class-pair table), the BVH items' transforms and kinds (every item test of the walk)
// A kernel that reads its BVH nodes from LDS (no F_GNODES) has its packed values, layout and BVH items' transforms
// and kinds staged there too (the host launches the F_GNODES variant otherwise), so their pointers are LDS
// addresses the compiler knows as such: LDS reads instead of generic ones that wait on both memory counters (2 %
// of the tank's forward).  The F_GNODES variant stages what fits, behind generic pointers.
template <unsigned F>
__device__ inline Args stage_tables(const Args &a0) {
  Args a = a0;
  constexpr bool ALL = (F & F_GNODES) == 0;
  if (ALL || a0.lds_values) {
    float *v = reinterpret_cast<float *>(lds_cells) + a0.lds_values_off;
    for (int i = threadIdx.x; i < a0.lds_values; i += blockDim.x) v[i] = a0.P[i];
    if constexpr (ALL) {
      if constexpr ((F & F_REF) == 0) a.Pr = v;
    } else if (a0.Pr == a0.P) {
      a.Pr = v;
    }
    a.P = v;
  }
  if (ALL || a0.lds_layout) {
    int *l = reinterpret_cast<int *>(lds_cells) + a0.lds_layout_off;
    for (int i = threadIdx.x; i < a0.lds_layout; i += blockDim.x) l[i] = a0.L[i];
    a.L = l;
  }
  if (ALL || a0.lds_items) {
    float4 *x = reinterpret_cast<float4 *>(lds_cells) + a0.lds_ixf_off;
    int4 *k = reinterpret_cast<int4 *>(lds_cells) + a0.lds_ikind_off;
    for (int i = threadIdx.x; i < 3 * a0.lds_items; i += blockDim.x) x[i] = a0.g.ixf[i];
    for (int i = threadIdx.x; i < a0.lds_items; i += blockDim.x) k[i] = a0.g.ikind[i];
    a.g.ixf = x;
    a.g.ikind = k;
  }
  return a;
}

template <unsigned F>
__global__ void __launch_bounds__(block_of<F>()) LH_OCC_FWD(F) trace_kernel(Args a0, int n_lds) {
  if (a0.lds_tally)
    for (int i = threadIdx.x; i < n_lds; i += blockDim.x) lds_cells[i] = 0.0;
  if (threadIdx.x == 0) s_bounded = 0u;
  if (threadIdx.x < RF_N) s_rare[threadIdx.x] = 0.0f;
  if (threadIdx.x < 32) s_hint[threadIdx.x] = -1;
  stage_nodes(a0.g);
  const Args a = stage_tables<F>(a0);
  __syncthreads();
  Photon p;
  Acc acc{};
  bool alive = false;
  const int lane = threadIdx.x & 63;
  long long cursor = 0, chunk_end = 0;
  bool drained = false;
  for (;;) {
    // refill the lanes whose photon ended from the wavefront's chunk of the work queue
    bool need = !alive;
    unsigned long long mask = __ballot(need);
    while (mask != 0ull && !drained) {
      const long long avail = chunk_end - cursor;
      if (avail <= 0) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(a.queue, (unsigned long long)a.chunk);
        base = __shfl(base, 0);
        if ((long long)base >= a.total) {
          drained = true;
          break;
        }
        cursor = (long long)base;
        chunk_end = min((long long)base + a.chunk, a.total);
        continue;
      }
      const int rank = lanes_below(mask);
      const bool take = need && rank < avail;
      if (take) alive = emit<(F & F_TORUS) != 0>(a, cursor + rank, p, acc);
      cursor += min((long long)__popcll(mask), avail);
      need = need && !take;
      mask = __ballot(need);
    }
    if (!__any(alive) && drained) break;
    if (alive) {
      alive = step<false, F>(a, p, acc, nullptr, nullptr);
      if (!alive) photon_end(a, p);
    }
  }
  flush_health(a, acc);
  if (a.lds_tally) {  // one global atomic per touched cell per block
    __syncthreads();
    for (int i = threadIdx.x; i < n_lds; i += blockDim.x)
      if (lds_cells[i] != 0.0) unsafeAtomicAdd(a.tally + i, lds_cells[i]);
  }
}

// the adjoint: every photon walked twice from its Philox counter (adjoint.h); per-block FP64 accumulators in LDS
// The adjoint runs at 3 wavefronts per SIMD (168 VGPRs, some spilled), one 768-thread block per CU: on the
// 1,000-PMT tank's 2^24 photons 12.1 ms against 14.0 at four 256-thread blocks with the nodes in memory (and 31.8
// at its natural 2 wavefronts); with a Mie or WLS channel (240 VGPRs at its natural 2) the LAr voxels' adjoint
// takes 8 % less time at 3, with separate reference values (256 VGPRs, 1 wavefront) about a third less.
#ifndef LH_ADJ_WAVES_PER_EU
#define LH_ADJ_WAVES_PER_EU 3
#endif
template <unsigned F>
__global__ void __launch_bounds__(adj_block_of<F>())
    __attribute__((amdgpu_waves_per_eu(LH_ADJ_WAVES_PER_EU)))
    adjoint_kernel(Args a0, AdjArgs b) {
  double *sg = lds_cells;
  for (int i = threadIdx.x; i < b.n_grad; i += blockDim.x) sg[i] = 0.0;
  if (threadIdx.x == 0) s_bounded = 0u;
  if (threadIdx.x < RF_N) s_rare[threadIdx.x] = 0.0f;
  if (threadIdx.x < 32) s_hint[threadIdx.x] = -1;
  stage_nodes(a0.g);
  const Args a = stage_tables<F>(a0);
  __syncthreads();
  Photon p;
  Acc acc{};
  bool alive = false;
  const int lane = threadIdx.x & 63;
  long long cursor = 0, chunk_end = 0;
  bool drained = false;
  for (;;) {
    bool need = !alive;
    unsigned long long mask = __ballot(need);
    while (mask != 0ull && !drained) {
      const long long avail = chunk_end - cursor;
      if (avail <= 0) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(a.queue, (unsigned long long)a.chunk);
        base = __shfl(base, 0);
        if ((long long)base >= a.total) {
          drained = true;
          break;
        }
        cursor = (long long)base;
        chunk_end = min((long long)base + a.chunk, a.total);
        continue;
      }
      const int rank = lanes_below(mask);
      const bool take = need && rank < avail;
      if (take) {
        p.gi = cursor + rank;
        alive = emit<(F & F_TORUS) != 0>(a, p.gi, p, acc);
        p.pass = 0;
        p.T = p.Dsum = 0.0f;
        p.hasdep = false;
      }
      cursor += min((long long)__popcll(mask), avail);
      need = need && !take;
      mask = __ballot(need);
    }
    if (!__any(alive) && drained) break;
    if (alive) {
      alive = step<true, F>(a, p, acc, &b, sg);
      if (!alive && p.pass == 0 && p.hasdep) {  // replay the photon from its counter, now adding the terms
        const float T = p.Dsum;
        Acc scratch{};
        alive = emit<(F & F_TORUS) != 0>(a, p.gi, p, scratch);
        p.pass = 1;
        p.T = T;
        p.Dsum = 0.0f;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < b.n_grad; i += blockDim.x)
    if (sg[i] != 0.0) unsafeAtomicAdd(b.grad + i, sg[i]);
  flush_health(a, acc);
}

}  // namespace lh

// ---------------------------------------------------------------------------------------------------- the C ABI
#define LH_API extern "C" __attribute__((visibility("default")))
using namespace lh;

struct SrcBuf {
  lh_source *src = nullptr;
  long long *src_end = nullptr;
  int cap = 0;
};

constexpr int TIME_RING = 256;  // launches whose kernel times a model keeps

struct lh_model_s {
  Geo g;
  void *geo_buf = nullptr;
  int *layout = nullptr;
  int layout_words = 0;
  int S = 0, K = 1, Kl = 1, Kt = 1, n_media = 0;  // K = Kl wavelength knots x Kt theta knots
  // the diffusion jump's grids: per medium two float4 records, the bounds of every grid in one buffer
  std::vector<float4> jrec_h;
  std::vector<float> jclear_h;
  std::vector<int32_t> layout_host;
  float4 *jrec = nullptr;
  float *jclear = nullptr;
  long long tally_stride = 0;
  int NT = 0;  // the time histogram's nodes t0 + j dt (ns), j < NT (0: none)
  float t0 = 0.0f, dt = 1.0f;
  unsigned long long *queue = nullptr;
  lh_source *src = nullptr;
  long long *src_end = nullptr;
  int src_cap = 0;

  long long values_needed = 0;  // the packed values a call must pass (the largest table's end)
  // the photon-loop kernel's start and end of the last TIME_RING launches (launch k in slot k % TIME_RING)
  hipEvent_t ev0[TIME_RING] = {}, ev1[TIME_RING] = {};
  long long launches = 0;
  // pinned host staging for the sources, two buffers used in turn: a call copies from one while the copy from the
  // other may still be queued, so no call waits for the stream (the `done` event marks a buffer's copy finished)
  lh_source *pin_src[2] = {};
  long long *pin_end[2] = {};
  hipEvent_t pin_done[2] = {};
  int pin_cap = 0, pin_next = 0;
  float *ut = nullptr;  // the unified facet tables on the device
  int cus = 0, lds_per_cu = 0, lds_per_block = 0;
  bool has_chan = false;   // a medium has a Mie or WLS channel
  bool has_torus = false;  // a CSG leaf is a torus (the kernels with the torus root isolation)
  // the last launch's shape: grid blocks, blocks per CU, LDS bytes per block, tallies in LDS, nodes in LDS, the
  // LDS budget per block, the CU's LDS, the kernel's feature bits
  int32_t last_launch[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const void *occ_fn[OCC_CACHE] = {};  // the occupancy cache of grid_of
  size_t occ_lds[OCC_CACHE] = {};
  int occ_per[OCC_CACHE] = {};
  int occ_next = 0;
  unsigned char *scratch = nullptr;  // this device's copy of the static scratch array (lh_trace_static)
  int left_at = 0, born_at = 0;      // the layout's sensor-crediting tables (words 16, 17; 0: none)
  int n_items = 0;                   // the BVH items (the forward kernel may stage their transforms in LDS)
};

static thread_local std::string g_err;

// The scratch of the allocation-free entry point lh_trace_static: the work-queue counter, then the staged sources
// and their end offsets.  Each device holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress gives on that device (static_scratch_of), never through the array's name.
constexpr size_t STATIC_SRC_OFF = 256;
constexpr size_t STATIC_END_OFF = STATIC_SRC_OFF + (sizeof(lh_source) * LH_STATIC_MAX_SOURCES + 255) / 256 * 256;
constexpr size_t STATIC_BYTES = STATIC_END_OFF + sizeof(long long) * LH_STATIC_MAX_SOURCES;
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];
constexpr int STATIC_MAX_DEVICES = 64;
static unsigned char *g_scratch_dev[STATIC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail(LH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

LH_API const char *lh_last_error(void) { return g_err.c_str(); }

LH_API int lh_model_create(const lh_geometry *G, const int32_t *layout, int32_t layout_words, lh_model *out) {
  if (!G || !layout || !out) return fail(LH_ERR_ARG, "lh_model_create: null argument");
  if (layout_words < 18 || layout[0] != 0x4C484F31) return fail(LH_ERR_LAYOUT, "lh_model_create: not a packed layout");
  if (layout[14] < 1 || layout[14] > 4096 || layout[15] < 0 || (long long)layout[5] * layout[14] > (1 << 20))
    return fail(LH_ERR_LAYOUT, "lh_model_create: the charge cells (wavelength x theta knots) are out of range");
  {  // the time histogram's grid
    float t0 = 0.0f, dt = 0.0f;
    memcpy(&t0, &layout[12], 4);
    memcpy(&dt, &layout[13], 4);
    if (layout[11] < 0 || layout[11] > (1 << 20) || (layout[11] > 0 && (!(dt > 0.0f) || !std::isfinite(dt) ||
                                                                          !std::isfinite(t0))))
      return fail(LH_ERR_LAYOUT, "lh_model_create: a time histogram needs 0 < nt <= 2^20 nodes, dt > 0, finite t0");
  }
  const int n = G->n_vol;
  if (n <= 0 || n >= (1 << 21)) return fail(LH_ERR_ARG, "lh_model_create: n_vol out of range");
  for (int i = 0; i < G->n_nodes; ++i) {
    if (G->count[i] > 7) return fail(LH_ERR_ARG, "lh_model_create: a BVH leaf holds more than 7 items");
    if (G->count[i] < 0) return fail(LH_ERR_ARG, "lh_model_create: a negative leaf count");
    // every walk moves forward (skip links point past the node's subtree), so it ends
    if (G->skip[i] <= i || G->skip[i] > G->n_nodes)
      return fail(LH_ERR_ARG, "lh_model_create: a skip link does not point forward within the nodes");
    if (G->count[i] > 0 && (G->start[i] < 0 || G->start[i] + G->count[i] > G->n_items))
      return fail(LH_ERR_ARG, "lh_model_create: a leaf's items are out of range");
  }
  for (int i = 0; i < G->n_items; ++i)
    if (G->items[i] < 0 || G->items[i] >= n) return fail(LH_ERR_ARG, "lh_model_create: an item is not a volume");
  for (int i = 0; i < n; ++i)
    if (G->kind[i] < 0 || G->kind[i] >= G->n_kinds) return fail(LH_ERR_ARG, "lh_model_create: a kind is out of range");
  for (int k = 0; k < G->n_kinds; ++k) {
    if (G->kind_l0[k] < 0 || G->kind_l0[k] + G->kind_nl[k] > G->n_leaves || G->kind_nl[k] < 1 ||
        G->kind_nl[k] > MAX_CSG_LEAVES || G->kind_p0[k] < 0 || G->kind_p0[k] + G->kind_pl[k] > G->n_prog)
      return fail(LH_ERR_ARG, "lh_model_create: a kind's leaves or program are out of range (at most 8 leaves)");
    int depth = 0;
    for (int j = 0; j < G->kind_pl[k]; ++j) {
      const int op = G->prog_op[G->kind_p0[k] + j], arg = G->prog_arg[G->kind_p0[k] + j];
      if (op == 0) {
        if (arg < 0 || arg >= G->kind_nl[k]) return fail(LH_ERR_ARG, "lh_model_create: a program pushes a foreign leaf");
        if (++depth > 31) return fail(LH_ERR_ARG, "lh_model_create: a program deeper than 31");
      } else {
        if (depth < 2 || op > 3) return fail(LH_ERR_ARG, "lh_model_create: a malformed program");
        --depth;
      }
    }
    if (depth != 1) return fail(LH_ERR_ARG, "lh_model_create: a program that does not leave one value");
  }
  for (int l = 0; l < G->n_leaves; ++l)
    if (G->leaf_p0[l] < 0 || G->leaf_p0[l] + G->leaf_np[l] > G->n_planes)
      return fail(LH_ERR_ARG, "lh_model_create: a leaf's planes are out of range");
  for (int v = 0; v <= n; ++v) {
    if (G->vg0[v] < 0 || G->vg1[v] > G->n_groups || G->vg0[v] > G->vg1[v])
      return fail(LH_ERR_ARG, "lh_model_create: a frame group range is out of range");
    if (v < n && (G->parent[v] < 0 || G->parent[v] > n)) return fail(LH_ERR_ARG, "lh_model_create: a parent is out of range");
  }
  for (int gi = 0; gi < G->n_groups; ++gi)
    if (G->g0[gi] < 0 || G->g1[gi] > G->n_nodes || G->g0[gi] > G->g1[gi])
      return fail(LH_ERR_ARG, "lh_model_create: a group's nodes are out of range");
  if (G->skin)
    for (int v = 0; v <= n; ++v)
      if (G->skin[v] < -1 || G->skin[v] >= G->n_leaves || (v == n && G->skin[v] != -1))
        return fail(LH_ERR_ARG, "lh_model_create: a daughter-free core's leaf is out of range");
  const int nc = layout[2];
  for (int v = 0; v <= n; ++v) {
    if (G->cls[v] < 0 || G->cls[v] >= nc) return fail(LH_ERR_ARG, "lh_model_create: a class is out of range");
    if (G->sensor[v] < -1 || G->sensor[v] >= layout[4]) return fail(LH_ERR_ARG, "lh_model_create: a sensor is out of range");
  }
  {  // the layout's own ranges: media, classes, the row table
    const int nm = layout[1], nr = layout[3];
    const long long need_media = (long long)layout[6] + (long long)MEDIUM_WORDS * nm;
    const long long need_rows = (long long)layout[9] + (long long)ROW_WORDS * nr;
    if (nm < 0 || nc < 1 || nr < 0 || layout[5] < 1 || layout[4] < 0 || need_media > layout_words ||
        (long long)layout[7] + nc > layout_words || (long long)layout[8] + (long long)nc * nc > layout_words ||
        need_rows > layout_words)
      return fail(LH_ERR_LAYOUT, "lh_model_create: the layout's sections are out of range");
    for (int c = 0; c < nc; ++c)
      if (layout[layout[7] + c] < -1 || layout[layout[7] + c] >= nm)
        return fail(LH_ERR_LAYOUT, "lh_model_create: a class's medium is out of range");
    for (int c = 0; c < nc * nc; ++c)
      if (layout[layout[8] + c] < -1 || layout[layout[8] + c] >= nr)
        return fail(LH_ERR_LAYOUT, "lh_model_create: a surface row index is out of range");
    // the sensor-crediting tables: exit-sensor flags per class pair (0 or 1), a born row per class (-1 or a row)
    if (layout[16] < 0 || layout[17] < 0 || (layout[16] > 0 && (long long)layout[16] + (long long)nc * nc > layout_words) ||
        (layout[17] > 0 && (long long)layout[17] + nc > layout_words))
      return fail(LH_ERR_LAYOUT, "lh_model_create: the sensor-crediting tables are out of range");
    if (layout[16] > 0)
      for (int c = 0; c < nc * nc; ++c)
        if (layout[layout[16] + c] != 0 && layout[layout[16] + c] != 1)
          return fail(LH_ERR_LAYOUT, "lh_model_create: an exit-sensor flag is not 0 or 1");
    if (layout[17] > 0)
      for (int c = 0; c < nc; ++c)
        if (layout[layout[17] + c] < -1 || layout[layout[17] + c] >= nr)
          return fail(LH_ERR_LAYOUT, "lh_model_create: a born row is out of range");
    for (int r = 0; r < nr; ++r) {
      const int32_t *R = layout + layout[9] + ROW_WORDS * r;
      if (R[0] < 0 || R[0] > 2) return fail(LH_ERR_LAYOUT, "lh_model_create: an unknown surface model");
      for (int c = 0; c < N_COL; ++c)
        if (R[2 + 2 * c] < 0 || R[3 + 2 * c] < 1) return fail(LH_ERR_LAYOUT, "lh_model_create: a surface column is empty");
    }
    for (int mi = 0; mi < nm; ++mi) {
      const int32_t *M = layout + layout[6] + MEDIUM_WORDS * mi;
      if (!(M[8] & MED_OPAQUE) && (M[1] < 1 || M[3] < 1))
        return fail(LH_ERR_LAYOUT, "lh_model_create: a medium without RINDEX");
      if (M[8] & ~(MED_OPAQUE | MED_WLS | MED_WLS_EXP | MED_MIE | MED_SAME_ABS | MED_SAME_RAY))
        return fail(LH_ERR_LAYOUT, "lh_model_create: unknown medium flags");
      if (((M[8] & MED_SAME_ABS) && M[5] != M[1]) || ((M[8] & MED_SAME_RAY) && M[7] != M[1]))
        return fail(LH_ERR_LAYOUT, "lh_model_create: a table on RINDEX's knots with another knot count");
      for (int j = 0; j < 8; j += 2)
        if (M[j] < 0 || M[j + 1] < 0) return fail(LH_ERR_LAYOUT, "lh_model_create: a negative table offset");
      if (M[8] & MED_WLS) {  // WLSABSLENGTH, WLSCOMPONENT (two knots or more), the yield and the time constant
        for (int j = 9; j < 17; j += 2)
          if (M[j] < 0 || M[j + 1] < 1) return fail(LH_ERR_LAYOUT, "lh_model_create: a WLS table is empty");
        if (M[12] < 2) return fail(LH_ERR_LAYOUT, "lh_model_create: a WLS emission spectrum needs two knots");
      }
      if (M[8] & MED_MIE) {  // MIEHG (a table), MIEHG_FORWARD, MIEHG_BACKWARD, MIEHG_FORWARD_RATIO (one knot each)
        if (M[17] < 0 || M[18] < 1) return fail(LH_ERR_LAYOUT, "lh_model_create: a MIEHG table is empty");
        for (int j = 19; j < 25; j += 2)
          if (M[j] < 0 || M[j + 1] != 1)
            return fail(LH_ERR_LAYOUT, "lh_model_create: a Mie lobe constant is not one value");
      }
    }
  }

  // the device layout: float4 / int4 arrays in one allocation
  bool has_torus = false;
  std::vector<float4> xf(3 * (size_t)n), gxf(3 * (size_t)std::max(G->n_groups, 1)), node(2 * (size_t)std::max(G->n_nodes, 1)),
      leaf(4 * (size_t)G->n_leaves), plane((size_t)std::max(G->n_planes, 1));
  std::vector<int4> vinfo((size_t)n + 1), kinds((size_t)G->n_kinds), vkind((size_t)n + 1);
  std::vector<int2> vgrp((size_t)n + 1), grange((size_t)std::max(G->n_groups, 1));
  std::vector<int> items((size_t)std::max(G->n_items, 1)), prog((size_t)std::max(G->n_prog, 1));
  for (int v = 0; v < n; ++v) {
    const float *R = G->R + 9 * v, *t = G->t + 3 * v;
    for (int r = 0; r < 3; ++r) {  // row r of R^T is column r of R; offset -(R^T t)_r
      const float m0 = R[0 * 3 + r], m1 = R[1 * 3 + r], m2 = R[2 * 3 + r];
      xf[3 * v + r] = make_float4(m0, m1, m2, -(m0 * t[0] + m1 * t[1] + m2 * t[2]));
    }
  }
  for (int v = 0; v <= n; ++v) {
    vinfo[v] = make_int4(v < n ? (G->touch ? G->touch[v] & 1 : 1) : 0, v < n ? G->parent[v] : n, G->cls[v],
                         G->sensor[v]);
    vgrp[v] = make_int2(G->vg0[v], G->vg1[v]);
  }
  for (int gi = 0; gi < G->n_groups; ++gi) {
    const float *R = G->gR + 9 * gi;
    bool ident = true;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) ident &= R[3 * r + c] == (r == c ? 1.0f : 0.0f);
    for (int r = 0; r < 3; ++r) gxf[3 * gi + r] = make_float4(R[0 * 3 + r], R[1 * 3 + r], R[2 * 3 + r], ident ? 1.0f : 0.0f);
    grange[gi] = make_int2(G->g0[gi], G->g1[gi]);
  }
  for (int i = 0; i < G->n_nodes; ++i) {
    int skip = G->skip[i], sc = G->count[i] > 0 ? G->start[i] * 8 + G->count[i] : 0;
    float4 lo, hi;
    memcpy(&lo.w, &skip, 4);
    memcpy(&hi.w, &sc, 4);
    lo.x = G->lo[3 * i]; lo.y = G->lo[3 * i + 1]; lo.z = G->lo[3 * i + 2];
    hi.x = G->hi[3 * i]; hi.y = G->hi[3 * i + 1]; hi.z = G->hi[3 * i + 2];
    node[2 * i] = lo;
    node[2 * i + 1] = hi;
  }
  for (int i = 0; i < G->n_items; ++i) items[i] = G->items[i];
  for (int l = 0; l < G->n_leaves; ++l) {
    const float *q = G->leaf_q + 3 * l, *o = G->leaf_o + 3 * l, *A = G->leaf_A + 9 * l, *b = G->leaf_b + 3 * l;
    int typ = G->leaf_typ[l], p0 = G->leaf_p0[l], np = G->leaf_np[l];
    if (typ < 0 || typ > 2) return fail(LH_ERR_ARG, "lh_model_create: a CSG leaf type other than 0, 1, 2");
    if (typ == 2 && !(A[0] > 0.0f && A[1] > 0.0f))
      return fail(LH_ERR_ARG, "lh_model_create: a torus leaf needs radii R, r > 0 (its A00, A01)");
    has_torus |= typ == 2;
    float4 L3;
    L3.x = b[2];
    memcpy(&L3.y, &typ, 4);
    memcpy(&L3.z, &p0, 4);
    memcpy(&L3.w, &np, 4);
    leaf[4 * l] = make_float4(q[0], q[1], q[2], G->leaf_c[l]);
    leaf[4 * l + 1] = make_float4(A[0], A[1], A[2], A[4]);
    leaf[4 * l + 2] = make_float4(A[5], A[8], b[0], b[1]);
    leaf[4 * l + 3] = L3;
    for (int k = 0; k < np; ++k) {  // planes relative to q: n.(x - q) <= h - n.(q - o)
      const float *nn = G->plane_n + 3 * (p0 + k);
      const float h = G->plane_h[p0 + k] - (nn[0] * (q[0] - o[0]) + nn[1] * (q[1] - o[1]) + nn[2] * (q[2] - o[2]));
      plane[p0 + k] = make_float4(nn[0], nn[1], nn[2], h);
    }
  }
  for (int k = 0; k < G->n_kinds; ++k) kinds[k] = make_int4(G->kind_l0[k], G->kind_nl[k], G->kind_p0[k], G->kind_pl[k]);
  for (int v = 0; v <= n; ++v) vkind[v] = v < n ? kinds[G->kind[v]] : make_int4(0, 0, 0, 0);
  std::vector<float4> ixf((size_t)3 * std::max(G->n_items, 1));
  std::vector<int4> ikind((size_t)std::max(G->n_items, 1));
  for (int i = 0; i < G->n_items; ++i) {  // the items' volume data in the items' order
    const int u = G->items[i];
    for (int r = 0; r < 3; ++r) ixf[3 * i + r] = xf[3 * u + r];
    ikind[i] = vkind[u];
  }
  for (int j = 0; j < G->n_prog; ++j) prog[j] = (G->prog_op[j] << 24) | (G->prog_arg[j] & 0xFFFFFF);
  auto int_bits = [](int i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
  };
  // the daughter-free cores as the navigator reads them: an axis-aligned box about a centre, cut by a cylinder
  // about z (the shrunk tube or box of a tank), else the leaf itself
  std::vector<float4> vcore(3 * ((size_t)n + 1), make_float4(0.f, 0.f, 0.f, 0.f));
  for (int v = 0; v < n && G->skin; ++v) {
    const int li = G->skin[v];
    if (li < 0) continue;
    const float *A = G->leaf_A + 9 * li, *b = G->leaf_b + 3 * li, *o = G->leaf_o + 3 * li, *q = G->leaf_q + 3 * li;
    const int typ = G->leaf_typ[li], p0 = G->leaf_p0[li], np = G->leaf_np[li];
    bool simple = typ == 0 || (typ == 1 && A[0] == 1.0f && A[4] == 1.0f && A[8] == 0.0f && A[1] == 0.0f &&
                               A[2] == 0.0f && A[3] == 0.0f && A[5] == 0.0f && A[6] == 0.0f && A[7] == 0.0f &&
                               b[0] == 0.0f && b[1] == 0.0f && b[2] == 0.0f);
    const float *c = typ == 1 ? q : o;
    float lo3[3] = {-HUGEF, -HUGEF, -HUGEF}, hi3[3] = {HUGEF, HUGEF, HUGEF};
    for (int k = 0; k < np && simple; ++k) {
      const float *nn = G->plane_n + 3 * (p0 + k);
      int ax = -1;
      for (int j = 0; j < 3; ++j)
        if (std::fabs(nn[j]) == 1.0f && nn[(j + 1) % 3] == 0.0f && nn[(j + 2) % 3] == 0.0f) ax = j;
      if (ax < 0) { simple = false; break; }
      const float h = G->plane_h[p0 + k] - (nn[0] * (c[0] - o[0]) + nn[1] * (c[1] - o[1]) + nn[2] * (c[2] - o[2]));
      if (nn[ax] > 0.0f) hi3[ax] = std::min(hi3[ax], h);
      else lo3[ax] = std::max(lo3[ax], -h);
    }
    if (simple) {
      vcore[3 * v] = make_float4(c[0], c[1], c[2], typ == 1 ? -G->leaf_c[li] : HUGEF);
      vcore[3 * v + 1] = make_float4(lo3[0], lo3[1], lo3[2], int_bits(1));
      vcore[3 * v + 2] = make_float4(hi3[0], hi3[1], hi3[2], 0.0f);
    } else {
      vcore[3 * v + 1] = make_float4(0.f, 0.f, 0.f, int_bits(2));
      vcore[3 * v + 2] = make_float4(0.f, 0.f, 0.f, int_bits(li));
    }
  }

  // the reach bound's boxes: each volume's box and a BVH (one box a leaf) over the sensor volumes' boxes, all padded
  // by twice the relocation step (K_ON bands) at the geometry's scale, the largest coordinate plus the diagonal
  const bool has_reach = G->box_lo && G->box_hi;
  std::vector<float4> rbox(has_reach ? 2 * (size_t)n : 1), rnode(2);
  int n_rnode = -1;
  if (has_reach) {
    double mx = 0.0, ulo[3] = {HUGEF, HUGEF, HUGEF}, uhi[3] = {-HUGEF, -HUGEF, -HUGEF};
    for (int v = 0; v < n; ++v)
      for (int k = 0; k < 3; ++k) {
        const double l = G->box_lo[3 * v + k], h = G->box_hi[3 * v + k];
        if (!(l <= h) || !std::isfinite(l) || !std::isfinite(h))
          return fail(LH_ERR_ARG, "lh_model_create: a volume's box is empty or not finite");
        mx = std::max(mx, std::max(std::fabs(l), std::fabs(h)));
        ulo[k] = std::min(ulo[k], l);
        uhi[k] = std::max(uhi[k], h);
      }
    const double diag = std::sqrt((uhi[0] - ulo[0]) * (uhi[0] - ulo[0]) + (uhi[1] - ulo[1]) * (uhi[1] - ulo[1]) +
                                  (uhi[2] - ulo[2]) * (uhi[2] - ulo[2]));
    const double pad = 2.0 * K_ON * (BAND_ULPS * (double)EPS32 * (mx + diag) + BAND_FLOOR);
    auto down = [](double x) { return std::nextafter((float)x, -HUGEF); };  // float32 bounds that hold the box
    auto up = [](double x) { return std::nextafter((float)x, HUGEF); };
    for (int v = 0; v < n; ++v) {
      rbox[2 * v] = make_float4(down(G->box_lo[3 * v] - pad), down(G->box_lo[3 * v + 1] - pad),
                                down(G->box_lo[3 * v + 2] - pad), 0.0f);
      rbox[2 * v + 1] = make_float4(up(G->box_hi[3 * v] + pad), up(G->box_hi[3 * v + 1] + pad),
                                    up(G->box_hi[3 * v + 2] + pad), 0.0f);
    }
    std::vector<double> slo, shi;
    for (int v = 0; v < n; ++v)
      if (G->sensor[v] >= 0)
        for (int k = 0; k < 3; ++k) {
          slo.push_back(G->box_lo[3 * v + k] - pad);
          shi.push_back(G->box_hi[3 * v + k] + pad);
        }
    const int ns = (int)(slo.size() / 3);
    n_rnode = 0;
    if (ns > 0) {
      std::vector<double> nlo(6 * (size_t)ns), nhi(6 * (size_t)ns);
      std::vector<int32_t> nskip(2 * (size_t)ns), nstart(2 * (size_t)ns), ncount(2 * (size_t)ns), nitems(ns);
      int32_t nn = 0, depth = 0;
      const int rc = lh_bvh_build(slo.data(), shi.data(), ns, 1, 16, nlo.data(), nhi.data(), nskip.data(),
                                  nstart.data(), ncount.data(), nitems.data(), &nn, &depth);
      if (rc != LH_OK) return rc;
      rnode.assign(2 * (size_t)nn, make_float4(0.f, 0.f, 0.f, 0.f));
      for (int i = 0; i < nn; ++i) {
        if (nskip[i] <= i || nskip[i] > nn) return fail(LH_ERR_ARG, "lh_model_create: the sensor boxes' BVH is malformed");
        const int leaf = ncount[i] > 0 ? 1 : 0;
        rnode[2 * i] = make_float4(down(nlo[3 * i]), down(nlo[3 * i + 1]), down(nlo[3 * i + 2]), int_bits(nskip[i]));
        rnode[2 * i + 1] = make_float4(up(nhi[3 * i]), up(nhi[3 * i + 1]), up(nhi[3 * i + 2]), int_bits(leaf));
      }
      n_rnode = nn;
    }
  }

  size_t off = 0;
  auto place = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_xf = place(xf.size() * 16), o_gxf = place(gxf.size() * 16), o_node = place(node.size() * 16),
               o_leaf = place(leaf.size() * 16), o_plane = place(plane.size() * 16), o_vinfo = place(vinfo.size() * 16),
               o_kinds = place(kinds.size() * 16), o_vkind = place(vkind.size() * 16),
               o_ixf = place(ixf.size() * 16), o_ikind = place(ikind.size() * 16), o_vgrp = place(vgrp.size() * 8), o_grange = place(grange.size() * 8),
               o_items = place(items.size() * 4), o_prog = place(prog.size() * 4),
               o_vcore = place(vcore.size() * 16), o_rbox = place(rbox.size() * 16),
               o_rnode = place(rnode.size() * 16);
  lh_model m = new lh_model_s();
  m->has_torus = has_torus;
  char *buf = nullptr;
  hipError_t e = hipMalloc((void **)&buf, off);
  if (e != hipSuccess) {
    delete m;
    return fail(LH_ERR_HIP, std::string("lh_model_create: hipMalloc: ") + hipGetErrorString(e));
  }
  m->geo_buf = buf;
  std::vector<char> host(off);
  memcpy(host.data() + o_xf, xf.data(), xf.size() * 16);
  memcpy(host.data() + o_gxf, gxf.data(), gxf.size() * 16);
  memcpy(host.data() + o_node, node.data(), node.size() * 16);
  memcpy(host.data() + o_leaf, leaf.data(), leaf.size() * 16);
  memcpy(host.data() + o_plane, plane.data(), plane.size() * 16);
  memcpy(host.data() + o_vinfo, vinfo.data(), vinfo.size() * 16);
  memcpy(host.data() + o_kinds, kinds.data(), kinds.size() * 16);
  memcpy(host.data() + o_vkind, vkind.data(), vkind.size() * 16);
  memcpy(host.data() + o_ixf, ixf.data(), ixf.size() * 16);
  memcpy(host.data() + o_ikind, ikind.data(), ikind.size() * 16);
  memcpy(host.data() + o_vgrp, vgrp.data(), vgrp.size() * 8);
  memcpy(host.data() + o_grange, grange.data(), grange.size() * 8);
  memcpy(host.data() + o_items, items.data(), items.size() * 4);
  memcpy(host.data() + o_prog, prog.data(), prog.size() * 4);
  memcpy(host.data() + o_vcore, vcore.data(), vcore.size() * 16);
  memcpy(host.data() + o_rbox, rbox.data(), rbox.size() * 16);
  memcpy(host.data() + o_rnode, rnode.data(), rnode.size() * 16);
  HIP_TRY(hipMemcpy(buf, host.data(), off, hipMemcpyHostToDevice));
  Geo &g = m->g;
  g.nvol = n;
  g.xf = (const float4 *)(buf + o_xf);
  g.gxf = (const float4 *)(buf + o_gxf);
  g.node = (const float4 *)(buf + o_node);
  g.leaf = (const float4 *)(buf + o_leaf);
  g.plane = (const float4 *)(buf + o_plane);
  g.vinfo = (const int4 *)(buf + o_vinfo);
  g.kinds = (const int4 *)(buf + o_kinds);
  g.vkind = (const int4 *)(buf + o_vkind);
  g.ixf = (const float4 *)(buf + o_ixf);
  g.ikind = (const int4 *)(buf + o_ikind);
  m->n_items = G->n_items;
  g.vgrp = (const int2 *)(buf + o_vgrp);
  g.grange = (const int2 *)(buf + o_grange);
  g.items = (const int *)(buf + o_items);
  g.prog = (const int *)(buf + o_prog);
  g.vcore = (const float4 *)(buf + o_vcore);
  g.rbox = (const float4 *)(buf + o_rbox);
  g.rnode = (const float4 *)(buf + o_rnode);
  g.n_rnode = n_rnode;
  g.reach = G->reach;
  g.n_nodes = G->n_nodes;
  g.lds_nodes = 0;
  g.lds_node_off = 0;
  HIP_TRY(hipMalloc((void **)&m->layout, (size_t)layout_words * 4));
  HIP_TRY(hipMemcpy(m->layout, layout, (size_t)layout_words * 4, hipMemcpyHostToDevice));
  m->layout_words = layout_words;
  {
    long long need = std::max((long long)layout[10] + 2LL * layout[5], (long long)layout[15] + 2LL * layout[14]);
    for (int mi = 0; mi < layout[1]; ++mi) {
      const int32_t *M = layout + layout[6] + MEDIUM_WORDS * mi;
      for (int j = 0; j < 8; j += 2) need = std::max(need, (long long)M[j] + 2LL * M[j + 1]);
      if (M[8] & MED_WLS)
        for (int j = 9; j < 17; j += 2) need = std::max(need, (long long)M[j] + 2LL * M[j + 1]);
      if (M[8] & MED_MIE)
        for (int j = 17; j < 25; j += 2) need = std::max(need, (long long)M[j] + 2LL * M[j + 1]);
    }
    for (int r = 0; r < layout[3]; ++r) {
      const int32_t *R = layout + layout[9] + ROW_WORDS * r;
      for (int c = 0; c < N_COL; ++c) need = std::max(need, (long long)R[2 + 2 * c] + 2LL * R[3 + 2 * c]);
    }
    m->values_needed = need;
    for (int mi = 0; mi < layout[1]; ++mi)
      if (layout[layout[6] + MEDIUM_WORDS * mi + 8] & (MED_WLS | MED_MIE)) m->has_chan = true;
  }
  m->n_media = layout[1];
  m->layout_host.assign(layout, layout + layout_words);
  m->S = layout[4];
  m->Kl = layout[5];
  m->Kt = layout[14];
  m->K = m->Kl * m->Kt;
  m->NT = layout[11];
  m->left_at = layout[16];
  m->born_at = layout[17];
  memcpy(&m->t0, &layout[12], 4);
  memcpy(&m->dt, &layout[13], 4);
  m->tally_stride = 3LL * m->S + (long long)m->S * m->K + (m->NT > 0 ? (long long)m->S * (m->NT + 4) : 0);
  HIP_TRY(hipMalloc((void **)&m->queue, 256));
  {
    const std::vector<float> &ut = lh_host::unified_tables();
    if ((int)ut.size() != UT_WORDS) return fail(LH_ERR_ARG, "lh_model_create: the unified tables have the wrong size");
    HIP_TRY(hipMalloc((void **)&m->ut, ut.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(m->ut, ut.data(), ut.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  int dev = 0, cus = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= STATIC_MAX_DEVICES) return fail(LH_ERR_ARG, "lh_model_create: device index beyond 63");
  {  // once per device, with the device current: the address of its copy of the static scratch array
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
      void *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return fail(LH_ERR_HIP, std::string("lh_model_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      g_scratch_dev[dev] = (unsigned char *)ptr;
    }
    m->scratch = g_scratch_dev[dev];
  }
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int lds_cu = 0;
  HIP_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  m->lds_per_cu = lds_cu > 0 ? lds_cu : 64 * 1024;
  int lds_blk = 0;
  HIP_TRY(hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  m->lds_per_block = lds_blk > 0 ? lds_blk : 64 * 1024;
  m->cus = cus;
  for (int k = 0; k < TIME_RING; ++k) {
    HIP_TRY(hipEventCreate(&m->ev0[k]));
    HIP_TRY(hipEventCreate(&m->ev1[k]));
  }
  for (int k = 0; k < 2; ++k) HIP_TRY(hipEventCreateWithFlags(&m->pin_done[k], hipEventDisableTiming));
  *out = m;
  return LH_OK;
}

LH_API int lh_model_destroy(lh_model m) {
  if (!m) return LH_OK;
  (void)hipFree(m->geo_buf);
  (void)hipFree(m->layout);
  (void)hipFree(m->queue);
  (void)hipFree(m->src);
  (void)hipFree(m->src_end);
  (void)hipFree(m->ut);
  (void)hipFree(m->jrec);
  (void)hipFree(m->jclear);
  for (int k = 0; k < 2; ++k) {
    if (m->pin_done[k]) (void)hipEventSynchronize(m->pin_done[k]);
    (void)hipHostFree(m->pin_src[k]);
    (void)hipHostFree(m->pin_end[k]);
    if (m->pin_done[k]) (void)hipEventDestroy(m->pin_done[k]);
  }
  for (int k = 0; k < TIME_RING; ++k) {
    if (m->ev0[k]) (void)hipEventDestroy(m->ev0[k]);
    if (m->ev1[k]) (void)hipEventDestroy(m->ev1[k]);
  }
  delete m;
  return LH_OK;
}

LH_API int lh_model_jump(lh_model m, int32_t medium, const float lo[3], float h, const int32_t n[3],
                         const float *clear) {
  if (!m || !lo || !n || !clear) return fail(LH_ERR_ARG, "lh_model_jump: null argument");
  if (medium < 0 || medium >= m->n_media) return fail(LH_ERR_ARG, "lh_model_jump: medium out of range");
  const int32_t *M = m->layout_host.data() + m->layout_host[L_MEDIA_AT] + MEDIUM_WORDS * medium;
  if ((M[8] & (MED_OPAQUE | MED_MIE | MED_WLS)) || M[7] < 1)
    return fail(LH_ERR_ARG, "lh_model_jump: the medium's only channel must be Rayleigh");
  if (!(h > 0.0f) || !std::isfinite(h) || !std::isfinite(lo[0]) || !std::isfinite(lo[1]) || !std::isfinite(lo[2]))
    return fail(LH_ERR_ARG, "lh_model_jump: a grid needs a spacing h > 0 and a finite origin");
  const long long nn = (long long)n[0] * n[1] * n[2];
  if (n[0] < 1 || n[1] < 1 || n[2] < 1 || nn > (1LL << 26)) return fail(LH_ERR_ARG, "lh_model_jump: grid size out of range");
  for (long long i = 0; i < nn; ++i)
    if (!(clear[i] >= 0.0f) || !std::isfinite(clear[i])) return fail(LH_ERR_ARG, "lh_model_jump: a bound is negative or not finite");
  auto bits = [](int i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
  };
  if (m->jrec_h.empty())  // every medium without a grid: offset -1
    for (int k = 0; k < m->n_media; ++k) {
      m->jrec_h.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
      m->jrec_h.push_back(make_float4(0.f, 0.f, 0.f, bits(-1)));
    }
  if ((long long)m->jclear_h.size() + nn > (1LL << 28)) return fail(LH_ERR_ARG, "lh_model_jump: the grids exceed 1 GiB");
  const float ob = bits((int)m->jclear_h.size());
  float nb[3];
  memcpy(nb, n, 12);
  m->jrec_h[2 * medium] = make_float4(lo[0], lo[1], lo[2], h);
  m->jrec_h[2 * medium + 1] = make_float4(nb[0], nb[1], nb[2], ob);
  m->jclear_h.insert(m->jclear_h.end(), clear, clear + nn);
  (void)hipFree(m->jrec);
  (void)hipFree(m->jclear);
  m->jrec = nullptr;
  m->jclear = nullptr;
  HIP_TRY(hipMalloc((void **)&m->jrec, m->jrec_h.size() * sizeof(float4)));
  HIP_TRY(hipMemcpy(m->jrec, m->jrec_h.data(), m->jrec_h.size() * sizeof(float4), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc((void **)&m->jclear, m->jclear_h.size() * sizeof(float)));
  HIP_TRY(hipMemcpy(m->jclear, m->jclear_h.data(), m->jclear_h.size() * sizeof(float), hipMemcpyHostToDevice));
  return LH_OK;
}

LH_API int lh_model_info(lh_model m, int32_t *n_sensors, int32_t *n_qe, int64_t *tally_per_event) {
  if (!m) return fail(LH_ERR_ARG, "lh_model_info: null model");
  if (n_sensors) *n_sensors = m->S;
  if (n_qe) *n_qe = m->K;
  if (tally_per_event) *tally_per_event = m->tally_stride;
  return LH_OK;
}

// the host side of a staging: the sources and their photon prefix sums, to be copied from (stage_into)
struct Staging {
  const lh_source *src;
  const long long *end;
};

// copy the sources and their photon prefix sums into `buf` (grown as needed) and check them; with `m`, through the
// model's pinned buffers without waiting for the stream, else from pageable memory, waiting for the copy
static int stage_into(SrcBuf &buf, int n_media, const lh_source *sources, int n_sources, long long *total,
                      hipStream_t st, lh_model m = nullptr);

static int stage_sources(lh_model m, const lh_source *sources, int n_sources, long long *total, hipStream_t st) {
  SrcBuf b{m->src, m->src_end, m->src_cap};
  const int rc = stage_into(b, m->n_media, sources, n_sources, total, st, m);
  m->src = b.src;
  m->src_end = b.src_end;
  m->src_cap = b.cap;
  return rc;
}

static int stage_into(SrcBuf &buf, int n_media, const lh_source *sources, int n_sources, long long *total,
                      hipStream_t st, lh_model m) {
  if (!sources || n_sources <= 0) return fail(LH_ERR_ARG, "no sources");
  std::vector<long long> ends(n_sources);
  long long acc = 0;
  for (int i = 0; i < n_sources; ++i) {
    const lh_source &s = sources[i];
    if (s.n < 0) return fail(LH_ERR_ARG, "a source with a negative photon count");
    if (s.kind < 0 || s.kind > 2) return fail(LH_ERR_ARG, "an unknown source kind");
    if (s.event < 0) return fail(LH_ERR_ARG, "a negative event index");
    if (s.kind == LH_SRC_CHERENKOV && (s.medium < 0 || s.medium >= n_media))
      return fail(LH_ERR_ARG, "a Cherenkov source's medium is out of range");
    if (s.kind == LH_SRC_CHERENKOV && !(s.beta > 0.0f)) return fail(LH_ERR_ARG, "a Cherenkov source needs beta > 0");
    acc += s.n;
    ends[i] = acc;
  }
  if (n_sources > buf.cap) {
    (void)hipFree(buf.src);
    (void)hipFree(buf.src_end);
    buf.src = nullptr;
    buf.src_end = nullptr;
    buf.cap = 0;
    HIP_TRY(hipMalloc((void **)&buf.src, sizeof(lh_source) * n_sources));
    HIP_TRY(hipMalloc((void **)&buf.src_end, sizeof(long long) * n_sources));
    buf.cap = n_sources;
  }
  *total = acc;
  if (!m) {
    HIP_TRY(hipMemcpyAsync(buf.src, sources, sizeof(lh_source) * n_sources, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(buf.src_end, ends.data(), sizeof(long long) * n_sources, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));  // the host staging vector goes out of scope
    return LH_OK;
  }
  if (n_sources > m->pin_cap) {  // both buffers regrown, once their queued copies have finished
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(hipEventSynchronize(m->pin_done[k]));
      (void)hipHostFree(m->pin_src[k]);
      (void)hipHostFree(m->pin_end[k]);
      m->pin_src[k] = nullptr;
      m->pin_end[k] = nullptr;
    }
    m->pin_cap = 0;
    const int cap = std::max(n_sources, 64);
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(hipHostMalloc((void **)&m->pin_src[k], sizeof(lh_source) * cap, hipHostMallocDefault));
      HIP_TRY(hipHostMalloc((void **)&m->pin_end[k], sizeof(long long) * cap, hipHostMallocDefault));
    }
    m->pin_cap = cap;
  }
  const int k = m->pin_next;
  m->pin_next ^= 1;
  HIP_TRY(hipEventSynchronize(m->pin_done[k]));  // this buffer's previous copy has left it
  memcpy(m->pin_src[k], sources, sizeof(lh_source) * n_sources);
  memcpy(m->pin_end[k], ends.data(), sizeof(long long) * n_sources);
  HIP_TRY(hipMemcpyAsync(buf.src, m->pin_src[k], sizeof(lh_source) * n_sources, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(buf.src_end, m->pin_end[k], sizeof(long long) * n_sources, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(m->pin_done[k], st));
  return LH_OK;
}

static int fill_args(lh_model m, Args &a, const float *values, const float *ref, const lh_controls *ctl, uint64_t seed,
                     double *tally, double *health) {
  if (!values) return fail(LH_ERR_ARG, "null values");
  if (!tally || !health) return fail(LH_ERR_ARG, "null tally or health");
  a.g = m->g;
  a.L = m->layout;
  a.P = values;
  a.Pr = ref ? ref : values;
  a.has_ref = (ref && ref != values) ? 1 : 0;
  a.k0 = (uint32_t)seed;
  a.k1 = (uint32_t)(seed >> 32);
  lh_controls c{0.01f, 0.0f, 1 << 22, 0, 0, 0, 0, 0, 0};
  if (ctl) c = *ctl;
  if (c.blocks_per_cu < 0 || c.blocks_per_cu > 64) return fail(LH_ERR_ARG, "blocks_per_cu out of range (0 .. 64)");
  if (!(c.roulette > 0.0f)) return fail(LH_ERR_ARG, "the roulette is a number > 0");
  a.roulette = c.roulette;
  a.reach = m->g.reach;
  a.max_dist = c.max_distance > 0.0f ? c.max_distance : m->g.reach;
  a.max_vert = c.max_vertices > 0 ? c.max_vertices : (1 << 22);
  a.chunk = c.chunk > 0 ? c.chunk : DEFAULT_CHUNK;
  a.keep_zero = c.keep_zero_weight ? 1 : 0;
  a.reach_on = c.reach_bound ? 1 : 0;
  a.tally = tally;
  a.tally_stride = m->tally_stride;
  a.nt = m->NT;
  a.time_at = 3LL * m->S + (long long)m->S * m->K;
  a.t0 = m->t0;
  a.dt = m->dt;
  a.S = m->S;
  a.K = m->K;
  a.Kl = m->Kl;
  a.Kt = m->Kt;
  a.health = health;
  a.queue = m->queue;
  a.ut = m->ut;
  a.lds_tally = 0;
  a.lds_values = a.lds_values_off = a.lds_layout = a.lds_layout_off = 0;
  a.lds_items = a.lds_ixf_off = a.lds_ikind_off = 0;
  a.src = m->src;
  a.src_end = m->src_end;
  a.jrec = c.jump && m->jrec ? m->jrec : nullptr;
  a.jclear = m->jclear;
  a.left_at = m->left_at;
  a.born_at = m->born_at;
  return LH_OK;
}

// the kernel of a launch: forward or adjoint, specialised on its features (F_REF: separate reference values;
// F_CHAN: a medium has a Mie or WLS channel), so a run without them carries none of their registers
template <unsigned F>
static const void *kernel_f(bool adj) {
  return adj ? (const void *)adjoint_kernel<F & F_ADJ_MASK> : (const void *)trace_kernel<F>;
}

// the reach bound where the model has its boxes and the controls allow it; not compiled with the torus roots
static unsigned features_of(lh_model m, const Args &a, bool adj) {
  return (a.has_ref ? F_REF : 0u) | (m->has_chan ? F_CHAN : 0u) | (!adj && m->NT > 0 ? F_TIME : 0u) |
         (m->has_torus ? F_TORUS : 0u) | (m->g.n_rnode >= 0 && a.reach_on && !m->has_torus ? F_REACH : 0u);
}

// every feature set a launch can take: any of the first five bits, and the reach bound without the torus roots
#define LH_FEATURES(X)                                                                                         \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20) \
      X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38)  \
      X(39) X(40) X(41) X(42) X(43) X(44) X(45) X(46) X(47)

// the workgroup size of a launch's kernel
static int block_for(bool /* adjoint or forward */, unsigned /* features */) { return BLOCK; }

static const void *kernel_of(bool adj, unsigned f) {
  switch (f) {
#define LH_KCASE(F) \
  case F: return kernel_f<F>(adj);
    LH_FEATURES(LH_KCASE)
#undef LH_KCASE
  }
  return nullptr;
}

template <unsigned F>
static void launch_f(bool adj, dim3 grid, size_t bytes, hipStream_t st, const Args &a, const AdjArgs &b, int n_lds) {
  if (adj)
    hipLaunchKernelGGL(adjoint_kernel<F & F_ADJ_MASK>, grid, dim3(adj_block_of<F & F_ADJ_MASK>()), bytes, st, a, b);
  else
    hipLaunchKernelGGL(trace_kernel<F>, grid, dim3(block_of<F>()), bytes, st, a, n_lds);
}

static void launch_kernel(bool adj, unsigned f, dim3 grid, size_t bytes, hipStream_t st, const Args &a,
                          const AdjArgs &b, int n_lds) {
  switch (f) {
#define LH_LCASE(F) \
  case F: launch_f<F>(adj, grid, bytes, st, a, b, n_lds); break;
    LH_FEATURES(LH_LCASE)
#undef LH_LCASE
  }
}

static int grid_of(lh_model m, const lh_controls *ctl, const void *fn, size_t lds_bytes, int block) {
  // resident blocks per CU of (kernel, LDS bytes), remembered in a small cache
  int per = 0;
  for (int i = 0; i < OCC_CACHE; ++i)
    if (m->occ_fn[i] == fn && m->occ_lds[i] == lds_bytes) per = m->occ_per[i];
  if (per <= 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, lds_bytes) != hipSuccess || per < 1) per = 1;
    const int i = m->occ_next++ % OCC_CACHE;
    m->occ_fn[i] = fn;
    m->occ_lds[i] = lds_bytes;
    m->occ_per[i] = per;
  }
  const int want = (ctl && ctl->blocks_per_cu > 0) ? std::min(ctl->blocks_per_cu, per) : per;
  return m->cus * want;
}

// a launch's dynamic LDS: its cells (the adjoint's gradient accumulators: required; the forward's tallies:
// optional, `cells_optional`), then the BVH nodes.  The budget is the CU's LDS over the blocks its registers let
// reside, so LDS never lowers the occupancy; the nodes (a load on every walk step) come before optional cells.
// Sets a.lds_tally and the nodes' placement; returns the LDS bytes.
static size_t place_lds(lh_model m, Args &a, const void *fn, int block, size_t cells_bytes, bool cells_optional,
                        const lh_controls *ctl) {
  const int per_regs = std::max(1, grid_of(m, nullptr, fn, 0, block) / std::max(m->cus, 1));
  // the kernel's own static LDS (its __shared__ arrays) comes out of the same per-block limit
  hipFuncAttributes attr{};
  const size_t fixed = hipFuncGetAttributes(&attr, fn) == hipSuccess ? (size_t)attr.sharedSizeBytes : 1024;
  const size_t limit = std::min<size_t>((size_t)m->lds_per_block, (size_t)m->lds_per_cu / (size_t)per_regs);
  const size_t budget = (limit > fixed + 16 ? limit - fixed - 16 : 0) / 16 * 16;
  const size_t node_bytes = (size_t)m->g.n_nodes * 32;
  const bool want_nodes = m->g.n_nodes > 0 && !(ctl && ctl->nodes_global == 1);
  const bool want_cells = cells_bytes > 0 && !(cells_optional && ctl && ctl->tally_global == 1);
  const size_t cells_al = (cells_bytes + 15) / 16 * 16;
  // the packed values, the layout and the BVH items' transforms and kinds (3 float4 rows + an int4 kind each): a
  // kernel that reads its nodes from LDS reads all of them there too, so the nodes come with all three or not at all
  const size_t value_bytes = ((size_t)m->values_needed * sizeof(float) + 15) / 16 * 16;
  const size_t layout_bytes = ((size_t)m->layout_words * sizeof(int) + 15) / 16 * 16;
  const size_t item_bytes = (size_t)m->n_items * 64;
  const size_t nodes_need = node_bytes + value_bytes + layout_bytes + item_bytes;
  bool cells = want_cells && (!cells_optional || cells_al <= budget);
  bool nodes = want_nodes && (cells ? cells_al : 0) + nodes_need <= budget;
  if (cells_optional && cells && want_nodes && !nodes && nodes_need <= budget) {  // nodes first
    cells = false;
    nodes = true;
  }
  a.lds_tally = (cells_optional && cells) ? 1 : 0;
  a.g.lds_nodes = nodes ? 1 : 0;
  const size_t off = cells ? cells_al : 0;
  a.g.lds_node_off = (int)(off / 16);
  size_t end = off + (nodes ? node_bytes : 0);
  // then the values (their table searches then read LDS), the layout and the items, in that order: all of them
  // beside LDS nodes, else each where it still fits
  a.lds_values = a.lds_values_off = a.lds_layout = a.lds_layout_off = 0;
  a.lds_items = a.lds_ixf_off = a.lds_ikind_off = 0;
  const bool stage = !(ctl && ctl->nodes_global == 1);
  if (stage && value_bytes > 0 && end + value_bytes <= budget) {
    a.lds_values = (int)m->values_needed;
    a.lds_values_off = (int)(end / sizeof(float));
    end += value_bytes;
  }
  if (stage && layout_bytes > 0 && end + layout_bytes <= budget) {
    a.lds_layout = m->layout_words;
    a.lds_layout_off = (int)(end / sizeof(int));
    end += layout_bytes;
  }
  if (stage && m->n_items > 0 && end + item_bytes <= budget) {
    a.lds_items = m->n_items;
    a.lds_ixf_off = (int)(end / 16);
    a.lds_ikind_off = (int)(end / 16) + 3 * m->n_items;
    end += item_bytes;
  }
  m->last_launch[3] = a.lds_tally;
  m->last_launch[4] = a.g.lds_nodes;
  m->last_launch[5] = (int32_t)budget;
  m->last_launch[6] = m->lds_per_cu;
  return end;
}

static void launch_adjoint(lh_model m, const Args &a, const AdjArgs &b, const lh_controls *ctl, size_t bytes,
                           hipStream_t st) {
  const unsigned f = features_of(m, a, true) | (a.g.lds_nodes ? 0u : F_GNODES);
  const int grid = grid_of(m, ctl, kernel_of(true, f), bytes, block_for(true, f));
  m->last_launch[0] = grid;
  m->last_launch[1] = grid / std::max(m->cus, 1);
  m->last_launch[2] = (int32_t)bytes;
  m->last_launch[7] = (int32_t)f | 0x100;
  launch_kernel(true, f, dim3(grid), bytes, st, a, b, 0);
}

// the forward launch: the BVH nodes and the block's tallies in LDS as they fit
static void launch_trace(lh_model m, Args &a, int n_events, const lh_controls *ctl, hipStream_t st) {
  const long long cells = (long long)n_events * m->tally_stride;
  const unsigned f = features_of(m, a, false);
  const void *fn = kernel_of(false, f);
  const size_t bytes =
      place_lds(m, a, fn, block_for(false, f), cells <= LDS_TALLY_CELLS ? (size_t)cells * sizeof(double) : 0,
                true, ctl);
  const int n_lds = a.lds_tally ? (int)cells : 0;
  hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, st, a.queue);
  const unsigned fk = f | (a.g.lds_nodes ? 0u : F_GNODES);  // the walk reads its nodes from LDS or from memory
  const int grid = grid_of(m, ctl, kernel_of(false, fk), bytes, block_for(false, fk));
  m->last_launch[0] = grid;
  m->last_launch[1] = grid / std::max(m->cus, 1);
  m->last_launch[2] = (int32_t)bytes;
  m->last_launch[7] = (int32_t)fk;
  launch_kernel(false, fk, dim3(grid), bytes, st, a, AdjArgs{}, n_lds);
}

// a launch on `st` after one on another stream: the model's device sources and work-queue counter are rewritten by
// this launch, so `st` first waits for the previous launch's end (a no-op on the same stream)
static int after_previous_launch(lh_model m, hipStream_t st) {
  if (m->launches > 0) HIP_TRY(hipStreamWaitEvent(st, m->ev1[(int)((m->launches - 1) % TIME_RING)], 0));
  return LH_OK;
}

LH_API int lh_trace(lh_model m, const float *values, const float *ref, int64_t n_values, const lh_source *sources,
                    int32_t n_sources, uint64_t seed, const lh_controls *ctl, double *tally, int32_t n_events,
                    double *health, void *stream) {
  if (!m) return fail(LH_ERR_ARG, "lh_trace: null model");
  if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_trace: fewer packed values than the layout reads");
  for (int i = 0; i < n_sources && sources; ++i)
    if (sources[i].event >= n_events) return fail(LH_ERR_ARG, "lh_trace: a source's event is beyond n_events");
  hipStream_t st = (hipStream_t)stream;
  Args a;
  long long total = 0;
  int rc = after_previous_launch(m, st);
  if (rc) return rc;
  rc = stage_sources(m, sources, n_sources, &total, st);
  if (rc) return rc;
  rc = fill_args(m, a, values, ref, ctl, seed, tally, health);
  if (rc) return rc;
  a.nsrc = n_sources;
  a.total = total;
  const int slot = (int)(m->launches % TIME_RING);
  HIP_TRY(hipEventRecord(m->ev0[slot], st));
  launch_trace(m, a, n_events, ctl, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m->ev1[slot], st));
  ++m->launches;
  return LH_OK;
}

// the staged sources' end offsets (a prefix sum of their photon counts) and a cleared work queue, for
// lh_trace_static: one thread, n <= LH_STATIC_MAX_SOURCES
__global__ void static_prepare(const lh_source *src, int n, long long *ends, unsigned long long *queue) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  long long acc = 0;
  for (int i = 0; i < n; ++i) ends[i] = (acc += src[i].n);
  *queue = 0ull;
}

LH_API int lh_trace_static(lh_model m, const float *values, const float *ref, int64_t n_values,
                           const lh_source *sources, int32_t n_sources, uint64_t seed, const lh_controls *ctl,
                           double *tally, int32_t n_events, double *health, void *stream) {
  if (!m) return fail(LH_ERR_ARG, "lh_trace_static: null model");
  if (!m->scratch) return fail(LH_ERR_ARG, "lh_trace_static: the model holds no scratch address");
  if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_trace_static: fewer packed values than the layout reads");
  if (!sources || n_sources <= 0) return fail(LH_ERR_ARG, "lh_trace_static: no sources");
  if (n_sources > LH_STATIC_MAX_SOURCES)
    return fail(LH_ERR_ARG, "lh_trace_static: more sources than its scratch holds (LH_STATIC_MAX_SOURCES = 4096)");
  long long total = 0;
  for (int i = 0; i < n_sources; ++i) {  // the checks of stage_into, without its staging vector
    const lh_source &s = sources[i];
    if (s.n < 0) return fail(LH_ERR_ARG, "lh_trace_static: a source with a negative photon count");
    if (s.kind < 0 || s.kind > 2) return fail(LH_ERR_ARG, "lh_trace_static: an unknown source kind");
    if (s.event < 0 || s.event >= n_events) return fail(LH_ERR_ARG, "lh_trace_static: a source's event is out of range");
    if (s.kind == LH_SRC_CHERENKOV && (s.medium < 0 || s.medium >= m->n_media))
      return fail(LH_ERR_ARG, "lh_trace_static: a Cherenkov source's medium is out of range");
    if (s.kind == LH_SRC_CHERENKOV && !(s.beta > 0.0f))
      return fail(LH_ERR_ARG, "lh_trace_static: a Cherenkov source needs beta > 0");
    total += s.n;
  }
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *queue = (unsigned long long *)m->scratch;
  lh_source *src = (lh_source *)(m->scratch + STATIC_SRC_OFF);
  long long *ends = (long long *)(m->scratch + STATIC_END_OFF);
  Args a;
  int rc = fill_args(m, a, values, ref, ctl, seed, tally, health);
  if (rc) return rc;
  a.queue = queue;
  a.src = src;
  a.src_end = ends;
  a.nsrc = n_sources;
  a.total = total;
  HIP_TRY(hipMemcpyAsync(src, sources, sizeof(lh_source) * (size_t)n_sources, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(static_prepare, dim3(1), dim3(64), 0, st, (const lh_source *)src, (int)n_sources, ends, queue);
  HIP_TRY(hipGetLastError());
  const int slot = (int)(m->launches % TIME_RING);
  HIP_TRY(hipEventRecord(m->ev0[slot], st));
  launch_trace(m, a, n_events, ctl, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m->ev1[slot], st));
  ++m->launches;
  return LH_OK;
}

LH_API int lh_adjoint(lh_model m, const float *values, const float *ref, int64_t n_values, const int32_t *slots,
                      const lh_source *sources, int32_t n_sources, uint64_t seed, const lh_controls *ctl,
                      const double *cot, int32_t n_events, double *grad, int32_t n_grad, double *health,
                      void *stream) {
  if (!m) return fail(LH_ERR_ARG, "lh_adjoint: null model");
  if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_adjoint: fewer packed values than the layout reads");
  if (!slots || !cot || !grad) return fail(LH_ERR_ARG, "lh_adjoint: null slots, cotangent or gradient");
  if (n_grad < 1 || n_grad > MAX_GRAD) return fail(LH_ERR_ARG, "lh_adjoint: n_grad out of range (1 .. 8192)");
  for (int i = 0; i < n_sources && sources; ++i)
    if (sources[i].event >= n_events) return fail(LH_ERR_ARG, "lh_adjoint: a source's event is beyond n_events");
  hipStream_t st = (hipStream_t)stream;
  {  // every slot indexes the block's accumulators: checked on the host before the launch
    std::vector<int32_t> hs((size_t)n_values);
    HIP_TRY(hipMemcpyAsync(hs.data(), slots, sizeof(int32_t) * (size_t)n_values, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (long long i = 0; i < n_values; ++i)
      if (hs[i] < -1 || hs[i] >= n_grad) return fail(LH_ERR_ARG, "lh_adjoint: a gradient slot is out of range");
  }
  Args a;
  long long total = 0;
  int rc = after_previous_launch(m, st);
  if (rc) return rc;
  rc = stage_sources(m, sources, n_sources, &total, st);
  if (rc) return rc;
  rc = fill_args(m, a, values, ref, ctl, seed, const_cast<double *>(cot), health);
  if (rc) return rc;
  a.nsrc = n_sources;
  a.total = total;
  AdjArgs b{cot, slots, grad, n_grad};
  hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, st, m->queue);
  HIP_TRY(hipGetLastError());
  const int slot = (int)(m->launches % TIME_RING);
  HIP_TRY(hipEventRecord(m->ev0[slot], st));
  const size_t bytes = place_lds(m, a, kernel_of(true, features_of(m, a, true)), block_for(true, features_of(m, a, true)),
                                 (size_t)n_grad * sizeof(double), false, ctl);
  launch_adjoint(m, a, b, ctl, bytes, st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m->ev1[slot], st));
  ++m->launches;
  return LH_OK;
}

// a staged batch (lh_batch_create): its own sources, work queue and checked gradient slots
struct lh_batch_s {
  lh_model m = nullptr;
  SrcBuf src;
  unsigned long long *queue = nullptr;
  int nsrc = 0, events = 0;
  long long total = 0;
  const int32_t *slots = nullptr;
  int ngrad = 0;
};

LH_API int lh_batch_create(lh_model m, const lh_source *sources, int32_t n_sources, const int32_t *slots_dev,
                           int64_t n_values, int32_t n_grad, lh_batch *out, int64_t *total) {
  if (!m || !out) return fail(LH_ERR_ARG, "lh_batch_create: null model or output");
  *out = nullptr;
  std::vector<int32_t> hs;
  if (slots_dev) {
    if (n_grad < 1 || n_grad > MAX_GRAD) return fail(LH_ERR_ARG, "lh_batch_create: n_grad out of range (1 .. 8192)");
    if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_batch_create: fewer packed values than the layout reads");
    hs.resize((size_t)n_values);
    HIP_TRY(hipMemcpy(hs.data(), slots_dev, sizeof(int32_t) * (size_t)n_values, hipMemcpyDeviceToHost));
    for (long long i = 0; i < n_values; ++i)
      if (hs[i] < -1 || hs[i] >= n_grad) return fail(LH_ERR_ARG, "lh_batch_create: a gradient slot is out of range");
  }
  std::unique_ptr<lh_batch_s> b(new lh_batch_s());
  b->m = m;
  long long tot = 0;
  int rc = stage_into(b->src, m->n_media, sources, n_sources, &tot, nullptr);
  if (rc) {
    (void)hipFree(b->src.src);
    (void)hipFree(b->src.src_end);
    return rc;
  }
  if (hipMalloc((void **)&b->queue, 256) != hipSuccess) {
    (void)hipFree(b->src.src);
    (void)hipFree(b->src.src_end);
    return fail(LH_ERR_HIP, "lh_batch_create: hipMalloc of the work queue failed");
  }
  int ev = 0;
  for (int i = 0; i < n_sources; ++i) ev = std::max(ev, sources[i].event + 1);
  b->nsrc = n_sources;
  b->total = tot;
  b->events = ev;
  b->slots = slots_dev;
  b->ngrad = slots_dev ? n_grad : 0;
  if (total) *total = tot;
  *out = b.release();
  return LH_OK;
}

LH_API int lh_batch_destroy(lh_batch b) {
  if (!b) return LH_OK;
  (void)hipFree(b->src.src);
  (void)hipFree(b->src.src_end);
  (void)hipFree(b->queue);
  delete b;
  return LH_OK;
}

LH_API int lh_batch_trace(lh_batch b, const float *values, const float *ref, int64_t n_values, uint64_t seed,
                          const lh_controls *ctl, double *tally, int32_t n_events, double *health, void *stream) {
  if (!b) return fail(LH_ERR_ARG, "lh_batch_trace: null batch");
  lh_model m = b->m;
  if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_batch_trace: fewer packed values than the layout reads");
  if (n_events < b->events) return fail(LH_ERR_ARG, "lh_batch_trace: a staged source's event is beyond n_events");
  hipStream_t st = (hipStream_t)stream;
  Args a;
  int rc = fill_args(m, a, values, ref, ctl, seed, tally, health);
  if (rc) return rc;
  a.src = b->src.src;
  a.src_end = b->src.src_end;
  a.nsrc = b->nsrc;
  a.total = b->total;
  a.queue = b->queue;
  launch_trace(m, a, n_events, ctl, st);
  HIP_TRY(hipGetLastError());
  return LH_OK;
}

LH_API int lh_batch_adjoint(lh_batch b, const float *values, const float *ref, int64_t n_values, uint64_t seed,
                            const lh_controls *ctl, const double *cot, int32_t n_events, double *grad, int32_t n_grad,
                            double *health, void *stream) {
  if (!b) return fail(LH_ERR_ARG, "lh_batch_adjoint: null batch");
  lh_model m = b->m;
  if (!b->slots) return fail(LH_ERR_ARG, "lh_batch_adjoint: the batch was created without gradient slots");
  if (n_grad != b->ngrad) return fail(LH_ERR_ARG, "lh_batch_adjoint: n_grad differs from the batch's");
  if (n_values < m->values_needed) return fail(LH_ERR_ARG, "lh_batch_adjoint: fewer packed values than the layout reads");
  if (n_events < b->events) return fail(LH_ERR_ARG, "lh_batch_adjoint: a staged source's event is beyond n_events");
  if (!cot || !grad) return fail(LH_ERR_ARG, "lh_batch_adjoint: null cotangent or gradient");
  hipStream_t st = (hipStream_t)stream;
  Args a;
  int rc = fill_args(m, a, values, ref, ctl, seed, const_cast<double *>(cot), health);
  if (rc) return rc;
  a.src = b->src.src;
  a.src_end = b->src.src_end;
  a.nsrc = b->nsrc;
  a.total = b->total;
  a.queue = b->queue;
  AdjArgs g{cot, b->slots, grad, n_grad};
  hipLaunchKernelGGL(reset_queue, dim3(1), dim3(1), 0, st, a.queue);
  const size_t bytes = place_lds(m, a, kernel_of(true, features_of(m, a, true)), block_for(true, features_of(m, a, true)),
                                 (size_t)n_grad * sizeof(double), false, ctl);
  launch_adjoint(m, a, g, ctl, bytes, st);
  HIP_TRY(hipGetLastError());
  return LH_OK;
}

LH_API int lh_query(lh_model m, int32_t n, const float *x_dev, const float *d_dev, const int32_t *vol_dev,
                    const int32_t *last_dev, float tmax, float *t_dev, int32_t *vol_after_dev, float *normal_dev,
                    int32_t *srf_dev, int32_t *located_dev, void *stream) {
  if (!m) return fail(LH_ERR_ARG, "lh_query: null model");
  if (n < 0) return fail(LH_ERR_ARG, "lh_query: n < 0");
  if (n == 0) return LH_OK;
  if (!x_dev || !d_dev || !vol_dev || !last_dev || !t_dev || !vol_after_dev || !normal_dev || !srf_dev || !located_dev)
    return fail(LH_ERR_ARG, "lh_query: null buffer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(query_kernel, dim3((n + 255) / 256), dim3(256), 0, st, m->g, n, x_dev, d_dev, vol_dev, last_dev,
                     tmax > 0.0f ? tmax : m->g.reach, t_dev, vol_after_dev, normal_dev, srf_dev, located_dev);
  HIP_TRY(hipGetLastError());
  return LH_OK;
}

LH_API int lh_last_launch(lh_model m, int32_t *info) {
  if (!m || !info) return fail(LH_ERR_ARG, "lh_last_launch: null argument");
  for (int i = 0; i < 8; ++i) info[i] = m->last_launch[i];
  return LH_OK;
}

LH_API int lh_last_kernel_ms(lh_model m, float *ms) {
  if (!m || !ms) return fail(LH_ERR_ARG, "lh_last_kernel_ms: null argument");
  if (m->launches == 0) return fail(LH_ERR_ARG, "lh_last_kernel_ms: nothing launched yet");
  return lh_kernel_ms_of(m, m->launches - 1, ms);
}

LH_API int lh_launches(lh_model m, int64_t *count) {
  if (!m || !count) return fail(LH_ERR_ARG, "lh_launches: null argument");
  *count = m->launches;
  return LH_OK;
}

LH_API int lh_kernel_ms_of(lh_model m, int64_t launch, float *ms) {
  if (!m || !ms) return fail(LH_ERR_ARG, "lh_kernel_ms_of: null argument");
  if (launch < 0 || launch >= m->launches || launch < m->launches - TIME_RING)
    return fail(LH_ERR_ARG, "lh_kernel_ms_of: the launch is not among the last 256 of this model");
  const int slot = (int)(launch % TIME_RING);
  HIP_TRY(hipEventSynchronize(m->ev1[slot]));
  HIP_TRY(hipEventElapsedTime(ms, m->ev0[slot], m->ev1[slot]));
  return LH_OK;
}
