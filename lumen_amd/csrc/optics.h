// Optics on the device: tables, the bulk medium at a wavelength, the surface kernel and the direction samplers.
//
// Restates the reference's per-step physics (src/lumen/transport/step.py:318-359 `_step_context`,
// src/lumen/media/bulk.py:76-110 `rate`/`evaluate`, src/lumen/media/laws.py `_table`, `geant4_group_index`,
// src/lumen/surfaces/types.py:211-296 `chained_cuts`/`_none`/`_real`/`_geant4_chain`, src/lumen/surfaces/fresnel.py
// `amplitudes`, src/lumen/surfaces/polar.py, src/lumen/surfaces/outcomes.py, src/lumen/media/phase.py:45-75
// Rayleigh, src/lumen/lobes.py:60-78 Lambert, src/lumen/surfaces/facets.py:318-397 the unified facet path).
#pragma once
#include <hip/hip_runtime.h>

#include "nav.h"

namespace lh {

constexpr float HC = 1239.8419843320026f;  // eV nm
constexpr float C_VACUUM = 0.299792458f;   // m / ns
constexpr float TWO_PI = 6.283185307179586f;
constexpr float PI_F = 3.141592653589793f;
constexpr float HALF_PI = 1.5707963267948966f;
constexpr float FLOOR32 = 1e-30f;
constexpr float TINY_NORMAL = 1.17549435e-38f;
constexpr float DEFENSIVE = 0.01f;
constexpr float INDEX_MATCH = 128.0f * 1.1920929e-7f;  // |n_a - n_b| / max at or below: a facet refraction keeps w
constexpr float WLS_TRIES = 100.0f;  // G4OpWLS redraws an emission energy above the absorbed one up to 100 times
constexpr int N_OUT = 6;
enum { STRAIGHT = 0, REFRACT = 1, SPECULAR = 2, BACKSCATTER = 3, GLOSSY = 4, DIFFUSE = 5 };
enum { MODEL_REFLECT = 0, MODEL_FRESNEL = 1, MODEL_COMPLEX = 2 };
// FLAG_GLISUR: GLOSSY is GLISUR's facet path (width 1 - polish); FLAG_REFRACTS: that path refracts at each facet
// (GLISUR_DIELECTRIC, which may end beyond the surface); FLAG_COVER: its facets are drawn from the cover mixture
// (polish is a parameter: it moves the support's edge)
enum { FLAG_NORMAL = 1, FLAG_FAR_OPAQUE = 2, FLAG_EDGE = 4, FLAG_PER_NORMAL = 8, FLAG_GLISUR = 16, FLAG_REFRACTS = 32,
       FLAG_COVER = 64 };
enum { COL_R = 0, COL_T, COL_EFF, COL_SPIKE, COL_LOBE, COL_BACK, COL_SIGMA, COL_POLISH, N_COL };
// the packed layout's header words (lumen_amd/pack.py)
enum { L_MAGIC = 0, L_MEDIA, L_CLASSES, L_ROWS, L_SENSORS, L_QE, L_MEDIA_AT, L_CLASS_MEDIUM_AT, L_ROW_OF_AT,
       L_ROWS_AT, L_QE_TABLE, L_NT, L_T0, L_DT, L_CE_KNOTS, L_CE_TABLE, L_LEFT_AT, L_BORN_AT };
constexpr int MEDIUM_WORDS = 25;
// a medium's flag word (its word 8); MED_SAME_ABS / _RAY: ABSLENGTH / RAYLEIGH tabulated on RINDEX's knots
enum { MED_OPAQUE = 1, MED_WLS = 2, MED_WLS_EXP = 4, MED_MIE = 8, MED_SAME_ABS = 16, MED_SAME_RAY = 32 };
constexpr float MIE_G_MAX = 0.999f;  // the largest |g| of a Mie lobe
constexpr int ROW_WORDS = 2 + 2 * N_COL;

struct TabPos {
  int i;    // the knot pair (i, i + 1); for one knot or a clamped end, i with f 0
  float f;  // the weight of knot i + 1
};

// a table linear in its knot coordinate between knots, held beyond them (G4PhysicsVector::Value)
__device__ __forceinline__ TabPos tab_pos(const float *__restrict__ P, int off, int n, float x) {
  if (n <= 1 || x <= P[off]) return TabPos{0, 0.0f};
  if (x >= P[off + n - 1]) return TabPos{n - 1, 0.0f};
  // the last knot i <= n - 2 with E_i <= x, by bisection: ceil(log2(n - 1)) loads whatever x, so the lanes of a
  // wavefront stay in step (a scan from the first knot runs as long as its farthest lane)
  int i = 0, cnt = n - 1;
  while (cnt > 1) {
    const int half = cnt >> 1;
    const bool up = P[off + i + half] <= x;
    i = up ? i + half : i;
    cnt = up ? cnt - half : half;
  }
  const float e0 = P[off + i], e1 = P[off + i + 1];
  return TabPos{i, (x - e0) / (e1 - e0)};
}

__device__ __forceinline__ float tab_at(const float *__restrict__ P, int off, int n, TabPos p) {
  const float v0 = P[off + n + p.i];
  return p.f == 0.0f ? v0 : fmaf(p.f, P[off + n + p.i + 1] - v0, v0);
}

__device__ __forceinline__ float tab(const float *__restrict__ P, int off, int n, float x) {
  return tab_at(P, off, n, tab_pos(P, off, n, x));
}

// a mean free path's rate; a length of 0 (or below the smallest normal float) is 1 / FLT_MIN (media/bulk.py:61-73)
__device__ __forceinline__ float rate_of(float L) { return L >= TINY_NORMAL ? 1.0f / L : 1.0f / TINY_NORMAL; }

struct MediumAt {
  float n, ng, mu_abs, mu_ray, mu_mie, mu_wls;  // the event channels' rates in Geant4's order: Rayleigh, Mie, WLS
};

__device__ inline MediumAt medium_at(const int *__restrict__ L, const float *__restrict__ P, int m, float lam) {
  MediumAt r{1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (m < 0) return r;
  const int *M = L + L[L_MEDIA_AT] + MEDIUM_WORDS * m;
  if (M[8] & MED_OPAQUE) return r;
  const float E = HC / lam;
  const TabPos pn = tab_pos(P, M[0], M[1], E);
  r.n = tab_at(P, M[0], M[1], pn);
  const TabPos pg = tab_pos(P, M[2], M[3], E);  // Geant4 interpolates the group velocity (1 / n_g) in energy
  const float g0 = 1.0f / P[M[2] + M[3] + pg.i];
  const float inv = pg.f == 0.0f ? g0 : fmaf(pg.f, 1.0f / P[M[2] + M[3] + pg.i + 1] - g0, g0);
  r.ng = 1.0f / inv;
  if (M[5] > 0) r.mu_abs = rate_of((M[8] & MED_SAME_ABS) ? tab_at(P, M[4], M[5], pn) : tab(P, M[4], M[5], E));
  if (M[7] > 0) r.mu_ray = rate_of((M[8] & MED_SAME_RAY) ? tab_at(P, M[6], M[7], pn) : tab(P, M[6], M[7], E));
  if (M[8] & MED_MIE) r.mu_mie = rate_of(tab(P, M[17], M[18], E));
  if (M[8] & MED_WLS) r.mu_wls = rate_of(tab(P, M[9], M[10], E));
  return r;
}

__device__ __forceinline__ float clip01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// ------------------------------------------------------------------------------------------- WLS emission
// Geant4's sampler of an emission spectrum (media/spectra.py `g4_arrays`): each knot interval drawn with its
// trapezoid mass, uniform in energy within it; optionally under the one defensive mixture over the intervals
// without mass (`covered`: share `cover` uniform over them).  Walked interval by interval: spectra are short.
struct Spectrum {
  const float *E, *S;  // knots (eV) and intensities
  int n;
  float inv_tot, cover, inv_zero;
  __device__ float dens(int k) const {  // the density (1/eV) on interval k
    const float m = S[k] + S[k + 1];
    float d = 0.5f * m * inv_tot;
    if (cover > 0.0f) d = (1.0f - cover) * d + ((m <= 0.0f && E[k + 1] > E[k]) ? cover * inv_zero : 0.0f);
    return d;
  }
};

__device__ inline Spectrum spectrum_of(const float *E, const float *S, int n, float cover) {
  float tot = 0.0f, zero = 0.0f;
  for (int k = 0; k + 1 < n; ++k) {
    const float h = E[k + 1] - E[k];
    tot += 0.5f * (S[k] + S[k + 1]) * h;
    if (S[k] + S[k + 1] <= 0.0f && h > 0.0f) zero += h;
  }
  return Spectrum{E, S, n, tot > 0.0f ? 1.0f / tot : 0.0f, zero > 0.0f ? cover : 0.0f,
                  zero > 0.0f ? 1.0f / zero : 0.0f};
}

// the mass at energies up to e (spectra.py `cumulative`)
__device__ inline float spectrum_cdf(const Spectrum &q, float e) {
  if (e < q.E[0]) return 0.0f;
  float c = 0.0f, u = 0.0f;
  for (int k = 0; k + 1 < q.n; ++k) {
    const float h = q.E[k + 1] - q.E[k], d = q.dens(k);
    if (q.E[k] <= e || k == 0) u = c + d * fminf(fmaxf(e - q.E[k], 0.0f), h);  // the last interval starting <= e
    c += d * h;
  }
  return clip01(u);
}

// the energy at a uniform r (spectra.py `invert`): the exact inverse of the cumulative mass
__device__ inline float spectrum_invert(const Spectrum &q, float r) {
  float c = 0.0f, ck = 0.0f;
  int k = 0;
  for (int j = 0; j + 1 < q.n; ++j) {
    if (c <= r) {
      k = j;
      ck = c;
    }
    c += q.dens(j) * (q.E[j + 1] - q.E[j]);
  }
  const float d = q.dens(k), h = q.E[k + 1] - q.E[k];
  const float x = d > 0.0f ? fmaxf(r - ck, 0.0f) / d : 0.0f;
  return q.E[k] + fminf(fmaxf(x, 0.0f), h);
}

// the density of the interval holding e, 0 outside the knots (spectra.py `on_intervals`)
__device__ inline float spectrum_density(const Spectrum &q, float e) {
  if (!(e >= q.E[0] && e <= q.E[q.n - 1])) return 0.0f;
  int k = 0;
  for (int j = 0; j + 1 < q.n; ++j)
    if (q.E[j] <= e) k = j;
  return q.dens(k);
}

// G4OpWLS's chance that one of its 100 draws lands at or below the absorbed energy, f that of one draw
// (spectra.py `kept`)
__device__ __forceinline__ float wls_kept(float f) {
  return -expm1f(WLS_TRIES * log1pf(-fminf(fmaxf(f, 0.0f), 1.0f - EPS32)));
}

struct WlsDraw {
  float lam, delay, w, wr;  // the emitted wavelength (nm), the delay (ns), the weight and the reference weight
};

__device__ __forceinline__ float safe_ratio(float a, float b) { return b != 0.0f ? a / b : 0.0f; }

// a re-emission of a photon absorbed at lam_in (media/emission.py `EmissionRecord.sample`): the wavelength drawn
// from the reference spectrum (under the defensive mixture) truncated at the absorbed energy by its exact inverse
// CDF; the weight the yield times Geant4's chance of an energy below the absorbed one times the live / drawn
// density; the reference weight the same at the reference values; the delay from the live time constant
__device__ inline WlsDraw wls_emit(const float *__restrict__ P, const float *__restrict__ Pr, const int *M, float lam_in,
                                   float u_lam, float u_time) {
  const int off = M[11], n = M[12];
  const Spectrum q = spectrum_of(Pr + off, Pr + off + n, n, DEFENSIVE);
  const Spectrum l = spectrum_of(P + off, P + off + n, n, 0.0f);
  const float e_in = HC / lam_in;
  const float f_q = spectrum_cdf(q, e_in);
  const float e = spectrum_invert(q, u_lam * f_q);
  const float qd = spectrum_density(q, e);
  const float dens_q = safe_ratio(qd, f_q);
  const float f_l = spectrum_cdf(l, e_in);
  const float y = P[M[13] + M[14]], y_r = Pr[M[13] + M[14]], tau = P[M[15] + M[16]];
  WlsDraw r;
  r.lam = HC / e;
  r.w = safe_ratio(y * wls_kept(f_l) * safe_ratio(spectrum_density(l, e), f_l), dens_q);
  r.wr = safe_ratio(y_r * wls_kept(f_q) * safe_ratio(qd, f_q), dens_q);
  r.delay = (M[8] & MED_WLS_EXP) ? -tau * __logf(u_time) : tau;
  return r;
}

__device__ __forceinline__ float index_at(const int *__restrict__ L, const float *__restrict__ P, int m, float lam) {
  if (m < 0) return 1.0f;
  const int *M = L + L[L_MEDIA_AT] + MEDIUM_WORDS * m;
  if (M[8] & MED_OPAQUE) return 1.0f;
  return tab(P, M[0], M[1], HC / lam);
}


// ------------------------------------------------------------------------------------------- polarisation
__device__ __forceinline__ void basis(float3 n, float3 &t1, float3 &t2) {  // Duff et al., JCGT 2017
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  t1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  t2 = mk3(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ float3 to_world(float3 l, float3 n) {
  float3 t1, t2;
  basis(n, t1, t2);
  return mk3(l.x * t1.x + l.y * t2.x + l.z * n.x, l.x * t1.y + l.y * t2.y + l.z * n.y,
             l.x * t1.z + l.y * t2.z + l.z * n.z);
}

__device__ __forceinline__ float3 to_local_frame(float3 v, float3 n) {
  float3 t1, t2;
  basis(n, t1, t2);
  return mk3(dot3(v, t1), dot3(v, t2), dot3(v, n));
}

// an unpolarised state (|E|^2 < 1/2) becomes a random linear one perpendicular to d (polar.py `fill`)
__device__ __forceinline__ void pol_fill(float3 d, float3 &er, float3 &ei, float u) {
  if (dot3(er, er) + dot3(ei, ei) < 0.5f) {
    float3 t1, t2;
    basis(d, t1, t2);
    float s, c;
    __sincosf(TWO_PI * u, &s, &c);
    er = add3(mul3(c, t1), mul3(s, t2));
    ei = mk3(0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ void pol_normalise(float3 &er, float3 &ei) {
  const float s = dot3(er, er) + dot3(ei, ei);
  if (s > FLOOR32) {
    const float r = rsqrtf(s);
    er = mul3(r, er);
    ei = mul3(r, ei);
  } else {
    er = ei = mk3(0.f, 0.f, 0.f);
  }
}

// the part of E perpendicular to k, normalised (polar.py `project`)
__device__ __forceinline__ void pol_project(float3 &er, float3 &ei, float3 k) {
  er = fma3(-dot3(k, er), k, er);
  ei = fma3(-dot3(k, ei), k, ei);
  pol_normalise(er, ei);
}

// -E + 2 (E.h) h (polar.py `mirror`)
__device__ __forceinline__ void pol_mirror(float3 &er, float3 &ei, float3 h) {
  er = fma3(2.0f * dot3(er, h), h, mul3(-1.0f, er));
  ei = fma3(2.0f * dot3(ei, h), h, mul3(-1.0f, ei));
}

struct cf {
  float re, im;
};
__device__ __forceinline__ cf cmul(cf a, cf b) { return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cf cdiv(cf a, cf b) {
  const float d = b.re * b.re + b.im * b.im;
  return cf{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ float cabs2(cf a) { return a.re * a.re + a.im * a.im; }

// sqrt with Im >= 0 (the decaying wave); 0 near 0 (fresnel.py `decaying_sqrt`)
__device__ __forceinline__ cf decaying_sqrt(cf z) {
  const float m = sqrtf(z.re * z.re + z.im * z.im);
  if (m < 1e-12f) return cf{0.f, 0.f};
  float re = sqrtf(fmaxf(0.5f * (m + z.re), 0.0f));
  float im = sqrtf(fmaxf(0.5f * (m - z.re), 0.0f));
  if (z.im < 0.0f) im = -im;
  if (im < 0.0f) { re = -re; im = -im; }
  return cf{re, im};
}

struct Amps {
  cf rs, rp, ts, tp;
};

// Fresnel amplitudes from real n1 into real n2 at incidence cos_i (fresnel.py `amplitudes`)
__device__ inline Amps fresnel(float cos_i, float n1, float n2) {
  const float c = clip01(cos_i);
  const cf q = decaying_sqrt(cf{n2 * n2 - n1 * n1 * (1.0f - c * c), 0.0f});
  const cf ds{n1 * c + q.re, q.im}, dp{n2 * n2 * c + n1 * q.re, n1 * q.im};
  Amps a;
  const bool oks = sqrtf(cabs2(ds)) > 1e-12f, okp = sqrtf(cabs2(dp)) > 1e-12f;
  a.rs = oks ? cdiv(cf{n1 * c - q.re, -q.im}, ds) : cf{-1.f, 0.f};
  a.rp = okp ? cdiv(cf{n2 * n2 * c - n1 * q.re, -n1 * q.im}, dp) : cf{1.f, 0.f};
  a.ts = oks ? cdiv(cf{2.0f * n1 * c, 0.f}, ds) : cf{0.f, 0.f};
  a.tp = okp ? cdiv(cf{2.0f * n1 * n2 * c, 0.f}, dp) : cf{0.f, 0.f};
  return a;
}

// Fresnel amplitudes from real n1 into a complex index nr + i nk (an absorbing medium, nk >= 0) at incidence cos_i:
// q = sqrt(n2^2 - n1^2 sin^2) with Im q >= 0, r_s = (n1 c - q) / (n1 c + q), r_p = (n2^2 c - n1 q) / (n2^2 c + n1 q)
// (fresnel.py `amplitudes` with a complex n2); t_s, t_p likewise (unused: the photon never enters the medium)
__device__ inline Amps fresnel_complex(float cos_i, float n1, float nr, float nk) {
  const float c = clip01(cos_i);
  const cf n2sq{nr * nr - nk * nk, 2.0f * nr * nk};
  const cf q = decaying_sqrt(cf{n2sq.re - n1 * n1 * (1.0f - c * c), n2sq.im});
  const cf ds{n1 * c + q.re, q.im};
  const cf n2sqc{n2sq.re * c, n2sq.im * c};
  const cf dp{n2sqc.re + n1 * q.re, n2sqc.im + n1 * q.im};
  Amps a;
  const bool oks = sqrtf(cabs2(ds)) > 1e-12f, okp = sqrtf(cabs2(dp)) > 1e-12f;
  a.rs = oks ? cdiv(cf{n1 * c - q.re, -q.im}, ds) : cf{-1.f, 0.f};
  a.rp = okp ? cdiv(cf{n2sqc.re - n1 * q.re, n2sqc.im - n1 * q.im}, dp) : cf{1.f, 0.f};
  a.ts = oks ? cdiv(cf{2.0f * n1 * c, 0.f}, ds) : cf{0.f, 0.f};
  a.tp = okp ? cdiv(cmul(cf{2.0f * n1 * c, 0.f}, cf{nr, nk}), dp) : cf{0.f, 0.f};
  return a;
}

// the s axis unit(wi x n); at normal incidence a fixed perpendicular of wi (polar.py `s_axis`)
__device__ __forceinline__ float3 s_axis(float3 wi, float3 n) {
  const float3 c = cross3(wi, n);
  if (dot3(c, c) > EPS32 * EPS32) return unit3(c);
  float3 t1, t2;
  basis(wi, t1, t2);
  return t1;
}

// the state a_s E_s s + a_p E_p (k x s), normalised (polar.py `recombine`)
__device__ inline void pol_recombine(float3 &er, float3 &ei, float3 wi, float3 s, cf as, cf ap, float3 k) {
  const float3 p_in = cross3(wi, s);
  const cf es{dot3(er, s), dot3(ei, s)}, ep{dot3(er, p_in), dot3(ei, p_in)};
  const cf cs = cmul(as, es), cp = cmul(ap, ep);
  const float3 p = cross3(k, s);
  er = add3(mul3(cs.re, s), mul3(cp.re, p));
  ei = add3(mul3(cs.im, s), mul3(cp.im, p));
  pol_normalise(er, ei);
}

// ------------------------------------------------------------------------------------------- the surface kernel
struct SurfEval {
  float p[N_OUT];  // the continuing shares
  float detect;    // the detected share
  float sigma;     // the glossy lobe's width (GLISUR: its radius 1 - polish)
  Amps amps;       // (r_s, r_p, t_s, t_p) of the deterministic outcomes
  int flags;
};

__device__ __forceinline__ float col_at(const int *__restrict__ R, const float *__restrict__ P, int c, float E) {
  return tab(P, R[2 + 2 * c], R[3 + 2 * c], E);
}

// one surface row at a crossing: Geant4's chain (types.py:211-296): r, t, a = chained_cuts(Rc, T), r split by the
// reflection constants (or the Fresnel interface), t straight on, a detected with eff
__device__ inline SurfEval surface_eval(const int *__restrict__ L, const float *__restrict__ P, int row, float lam,
                                        float3 wi, float3 n, float n1, float n2, float3 er, float3 ei) {
  SurfEval s;
  const int *R = L + L[L_ROWS_AT] + ROW_WORDS * row;
  const int model = R[0], flags = R[1];
  s.flags = flags;
  const float E = HC / lam;
  const float Rv = col_at(R, P, COL_R, E), Tv = col_at(R, P, COL_T, E), eff = col_at(R, P, COL_EFF, E);
  float c1 = clip01(Rv), c2 = clip01(Rv + Tv);
  float r = c1, t = c2 - c1, a = 1.0f - c2;
  s.sigma = (flags & FLAG_GLISUR) ? 1.0f - col_at(R, P, COL_POLISH, E) : col_at(R, P, COL_SIGMA, E);
  for (int k = 0; k < N_OUT; ++k) s.p[k] = 0.0f;
  s.amps = Amps{cf{-1.f, 0.f}, cf{1.f, 0.f}, cf{1.f, 0.f}, cf{1.f, 0.f}};  // the perfect mirror's
  if (model == MODEL_REFLECT) {
    const float ps = col_at(R, P, COL_SPIKE, E), pl = col_at(R, P, COL_LOBE, E), pb = col_at(R, P, COL_BACK, E);
    const float k1 = clip01(ps), k2 = clip01(ps + pl), k3 = clip01(ps + pl + pb);
    const float ss = k1, sl = k2 - k1, bs = k3 - k2, lm = 1.0f - k3;
    const bool normal = flags & FLAG_NORMAL;
    s.p[SPECULAR] = r * (normal ? ss + sl : ss);
    s.p[GLOSSY] = normal ? 0.0f : r * sl;
    s.p[BACKSCATTER] = r * bs;
    s.p[DIFFUSE] = r * lm;
  } else {
    const float cos_i = clip01(-dot3(wi, n));
    const bool cx = model == MODEL_COMPLEX;  // a complex index (nr, nk in the p_lobe, p_back slots)
    s.amps = cx ? fresnel_complex(cos_i, n1, col_at(R, P, COL_LOBE, E), col_at(R, P, COL_BACK, E))
                : fresnel(cos_i, n1, n2);
    const float3 sx = s_axis(wi, n);
    const float3 px = cross3(wi, sx);
    const float es2 = dot3(er, sx) * dot3(er, sx) + dot3(ei, sx) * dot3(ei, sx);
    const float ep2 = dot3(er, px) * dot3(er, px) + dot3(ei, px) * dot3(ei, px);
    const float F = clip01(cabs2(s.amps.rs) * es2 + cabs2(s.amps.rp) * ep2);
    if (cx) {  // the reflectance is the cut (Geant4's CalculateReflectivity): the mirror's share, then T, the rest
      c1 = F;  // absorbed and detected with eff (types.py `_complex` in `_geant4_chain`)
      c2 = clip01(F + Tv);
      r = c1;
      t = c2 - c1;
      a = 1.0f - c2;
      s.p[SPECULAR] = r;
    } else {
      s.p[SPECULAR] = r * F;
      s.p[REFRACT] = r * (1.0f - F);
    }
  }
  s.p[STRAIGHT] = t;
  if (flags & FLAG_FAR_OPAQUE) s.p[STRAIGHT] = s.p[REFRACT] = 0.0f;
  s.detect = a * eff;
  if (flags & FLAG_PER_NORMAL) {  // read for a QE applied after tracking: the detected share over A_n, the type's
    // absorbed share at normal incidence (types.py `per_normal`, `normal_absorptance`): the cuts' a for a reflector
    // or an interface, 1 - clip(R_n + T) for a complex index
    float an = a;
    if (model == MODEL_COMPLEX) {
      const Amps n0 = fresnel_complex(1.0f, n1, col_at(R, P, COL_LOBE, E), col_at(R, P, COL_BACK, E));
      an = 1.0f - clip01(cabs2(n0.rs) + Tv);
    }
    s.detect = an > 0.0f ? s.detect / an : 0.0f;
  }
  return s;
}

// the outcome drawn from q with one uniform: the first k with cumsum_k > u * total; -1 when q is all zero
// (estimator.py `choose`)
__device__ __forceinline__ int choose6(const float q[N_OUT], float u) {
  float tot = 0.0f;
  for (int k = 0; k < N_OUT; ++k) tot += q[k];
  if (!(tot > 0.0f)) return -1;
  const float th = u * tot;
  float c = 0.0f;
  int pick = -1;
  for (int k = 0; k < N_OUT; ++k) {
    c += q[k];
    if (pick < 0 && c > th && q[k] > 0.0f) pick = k;
  }
  if (pick < 0)
    for (int k = N_OUT - 1; k >= 0; --k)
      if (q[k] > 0.0f) { pick = k; break; }
  return pick;
}

// ------------------------------------------------------------------------------------------- directions
// arccos by Abramowitz & Stegun 4.4.46 (Hastings' minimax form sqrt(1 - x) P7(x)): |error| <= 2.2e-8 on [-1, 1]
__device__ __forceinline__ float acos_as(float x) {
  const float ax = fminf(fabsf(x), 1.0f);
  float p = -0.0012624911f;
  p = fmaf(p, ax, 0.0066700901f);
  p = fmaf(p, ax, -0.0170881256f);
  p = fmaf(p, ax, 0.0308918810f);
  p = fmaf(p, ax, -0.0501743046f);
  p = fmaf(p, ax, 0.0889789874f);
  p = fmaf(p, ax, -0.2145988016f);
  p = fmaf(p, ax, 1.5707963050f);
  const float r = sqrtf(1.0f - ax) * p;
  return x >= 0.0f ? r : PI_F - r;
}

// arctan2 by Abramowitz & Stegun 4.4.49 (arctan z = z P8(z^2) on [0, 1], |error| <= 1.4e-8) on the octants
__device__ __forceinline__ float atan2_as(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float z = mx > 0.0f ? mn / mx : 0.0f, z2 = z * z;
  float p = 0.0028662257f;
  p = fmaf(p, z2, -0.0161657367f);
  p = fmaf(p, z2, 0.0429096138f);
  p = fmaf(p, z2, -0.0752896400f);
  p = fmaf(p, z2, 0.1065626393f);
  p = fmaf(p, z2, -0.1420889944f);
  p = fmaf(p, z2, 0.1999355085f);
  p = fmaf(p, z2, -0.3333314528f);
  p = fmaf(p, z2, 1.0f);
  float r = z * p;
  if (ay > ax) r = HALF_PI - r;
  if (x < 0.0f) r = PI_F - r;
  return copysignf(r, y);
}

// Rayleigh, polarised: the dipole axis er with probability |er|^2, else ei; c from p(c) = 3/4 (1 - c^2)
// (phase.py `rayleigh_sample`, `dipole_cos`)
__device__ inline float3 rayleigh_sample(float3 er, float3 ei, float u0, float u1, float u2) {
  const float a2 = dot3(er, er), norm2 = a2 + dot3(ei, ei);
  const float3 axis = unit3(u2 * norm2 < a2 ? er : ei);
  const float c = fminf(fmaxf(2.0f * __cosf((acos_as(fminf(fmaxf(1.0f - 2.0f * u0, -1.0f), 1.0f)) - TWO_PI) / 3.0f),
                              -1.0f), 1.0f);
  const float s = sqrtf(fmaxf(1.0f - c * c, 0.0f));
  float sp, cp;
  __sincosf(TWO_PI * u1, &sp, &cp);
  return to_world(mk3(s * cp, s * sp, c), axis);
}

// ------------------------------------------------------------------------------------------------- Mie
// G4OpMieHG (media/phase.py:98-190): a forward Henyey-Greenstein lobe HG(g_f) with probability r, else the
// backward one HG(-g_b); the constants MIEHG_FORWARD, MIEHG_BACKWARD, MIEHG_FORWARD_RATIO are the medium's words
// 19-24 (tables of one knot).  Cosines travel as (1 - mu, 1 + mu), both sums of positive terms, so neither cancels.
struct MieLobe {
  float gf, gb, r;
};

__device__ __forceinline__ MieLobe mie_lobe(const float *__restrict__ P, const int *M) {
  return MieLobe{P[M[19] + M[20]], P[M[21] + M[22]], P[M[23] + M[24]]};
}

// Geant4's inverse CDF mu = 2u(1+g)^2(1-g+gu)/(1-g+2gu)^2 - 1 as (1 - mu, 1 + mu) through the identity
// D^2 - u(1+g)^2(1-g+gu) = (1-g)^2(1-u)(1+gu), D = 1 - g + 2gu, each factor written without cancellation for
// either sign of g (phase.py `mie_hg_cos`)
__device__ __forceinline__ void hg_cos(float g, float u, float &cm, float &cp) {
  const bool up = g >= 0.0f;
  const float d = up ? (1.0f - g) + 2.0f * g * u : (1.0f - g) * (1.0f - u) + (1.0f + g) * u;
  const float one_gu = up ? 1.0f + g * u : (1.0f - u) + (1.0f + g) * u;
  const float inv_d2 = 1.0f / (d * d);
  cm = 2.0f * (1.0f - g) * (1.0f - g) * (1.0f - u) * one_gu * inv_d2;
  cp = 2.0f * u * (1.0f + g) * (1.0f + g) * ((1.0f - g) + g * u) * inv_d2;
}

// HG(mu; g) per steradian from cm = 1 - mu, cp = 1 + mu: 1 + g^2 - 2 g mu as (1-g)^2 + 2 g cm (g >= 0) or
// (1+g)^2 - 2 g cp (g < 0) (phase.py `hg_density`)
__device__ __forceinline__ float hg_density(float g, float cm, float cp) {
  const float q = g >= 0.0f ? (1.0f - g) * (1.0f - g) + 2.0f * g * cm : (1.0f + g) * (1.0f + g) - 2.0f * g * cp;
  return (1.0f - g) * (1.0f + g) / (2.0f * TWO_PI * q * sqrtf(q));
}

// d HG(mu; g) / dg over HG: -2g / (1 - g^2) - 3 (g - mu) / (1 + g^2 - 2 g mu)
__device__ __forceinline__ float hg_dlog_g(float g, float cm, float cp) {
  const float q = g >= 0.0f ? (1.0f - g) * (1.0f - g) + 2.0f * g * cm : (1.0f + g) * (1.0f + g) - 2.0f * g * cp;
  const float mu = cm < cp ? 1.0f - cm : cp - 1.0f;
  return -2.0f * g / ((1.0f - g) * (1.0f + g)) - 3.0f * (g - mu) / q;
}

// the lobe choice's forward ratio the draw is made at: the one defensive mixture over (r, 1 - r) (phase.py
// `_mie_cover`): e at r = 0, 1 - e at r = 1, r itself between
__device__ __forceinline__ float mie_cover(float r) { return r <= 0.0f ? DEFENSIVE : r >= 1.0f ? 1.0f - DEFENSIVE : r; }

// a direction about wi from the Mie lobes at (g_f, g_b, forward ratio rc): uniforms lobe, cosine, azimuth
// (phase.py `mie_hg_sample`, `hg_direction`)
__device__ inline float3 mie_sample(float3 wi, float gf, float gb, float rc, float u0, float u1, float u2) {
  const bool fwd = u0 <= rc;
  float cm, cp;
  hg_cos(fwd ? gf : gb, u1, cm, cp);
  if (!fwd) {
    const float t = cm;
    cm = cp;
    cp = t;
  }
  const float mu = cm < cp ? 1.0f - cm : cp - 1.0f;
  const float s = sqrtf(fmaxf(cm * cp, 0.0f));
  float sp, cph;
  __sincosf(TWO_PI * u2, &sp, &cph);
  return to_world(mk3(s * cph, s * sp, mu), wi);
}

// the two lobes' mixture density per steradian at (cm, cp) of the outgoing direction (phase.py `mie_hg_logpdf`)
__device__ __forceinline__ float mie_density(float gf, float gb, float r, float cm, float cp) {
  return r * hg_density(gf, cm, cp) + (1.0f - r) * hg_density(-gb, cm, cp);
}

// Lambertian about n (lobes.py `_cosine`)
__device__ __forceinline__ float3 lambert_sample(float3 n, float u0, float u1) {
  const float r = sqrtf(u0);
  float sp, cp;
  __sincosf(TWO_PI * u1, &sp, &cp);
  return to_world(mk3(r * cp, r * sp, sqrtf(fmaxf(1.0f - u0, 0.0f))), n);
}

// Snell's refraction at facing normal n with eta = n1 / n2; the mirror under total internal reflection
// (outcomes.py `refract`)
__device__ __forceinline__ float3 refract_dir(float3 wi, float3 n, float eta) {
  const float cos_i = -dot3(wi, n);
  const float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return fma3(2.0f * cos_i, n, wi);
  const float cos_t = sqrtf(fmaxf(1.0f - sin2_t, 0.0f));
  return unit3(fma3(eta * cos_i - cos_t, n, mul3(eta, wi)));
}

// the half-width of the azimuth arc of facets at tilt a that face V (facets.py `_front_arc`)
__device__ __forceinline__ float front_arc(float3 V, float a) {
  const float sv = sqrtf(V.x * V.x + V.y * V.y);
  float sa, ca;
  __sincosf(a, &sa, &ca);
  const float ratio = V.z * ca / fmaxf(sv * sa, FLOOR32);
  return acos_as(-fminf(ratio, 1.0f));
}

}  // namespace lh
