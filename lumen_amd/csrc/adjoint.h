// The adjoint's derivative terms: each step's log-derivatives of its weight factors with respect to the packed
// table values, added into per-block FP64 accumulators.
//
// The reference differentiates the photon loop through tallymap (a jvp by default, a reverse kernel keeping 2C+1
// checkpoints per photon: src/lumen/run.py:267-269, src/lumen/api.py:748-787).  Every random choice is drawn at the
// reference values and weighted live / reference (src/lumen/transport/estimator.py:1-40), so at the reference a
// deposit q_k = w0 prod_{j<k} f_j g_k has the derivative q_k (sum_{j<k} dlog f_j + dlog g_k): f_j the event factor of
// step j (attenuation, scattering density, the drawn outcome's share p_live / q), g_k the deposit's own factors
// (reach, attenuation, the detected share).  With a cotangent c per tally cell, dL/dtheta = sum_k D_k dlog g_k +
// sum_j dlog f_j R_{j+1}, D_k = c q_k and R_{j+1} = sum_{k>j} D_k the deposits still ahead: the adjoint replays each
// photon from its Philox counter (no stored path) -- once to total T = sum_k D_k, once to add the terms with
// R_{j+1} = T - sum_{k<=j} D_k.
//
// Pathwise terms (src/lumen/transport/step.py:318-359, :397-436): reach exp(-mu d) and attenuation exp(-mu_abs d)
// give -d dmu, -d dmu_abs; a bulk event's density mu exp(-mu s) gives dmu / mu - s dmu.  Score terms: the surface
// share of the drawn outcome (src/lumen/surfaces/types.py:211-296: chained_cuts with clip01's derivative, 1 on the
// closed interval [0, 1]) and the detected share a eff.  A table value v_i enters its column linearly (weights
// 1 - f, f), a rate mu = 1 / L with dmu = -dL / L^2 (src/lumen/media/bulk.py:61-73; 0 below FLT_MIN).
#pragma once
#include <hip/hip_runtime.h>

#include "optics.h"

namespace lh {

struct AdjArgs {
  const double *__restrict__ cot;  // dLoss / d(tally), the tally layout
  const int *__restrict__ slots;   // the gradient slot of each packed value (-1 none)
  double *__restrict__ grad;       // n_grad global accumulators
  int n_grad;
};

// a table's (up to two) value slots at x and the weights of its knots
struct TabD {
  int s0, s1;
  float w0, w1;
  float v;  // the table's value at x
};

__device__ __forceinline__ TabD tab_d(const float *__restrict__ P, const int *__restrict__ slots, int off, int n,
                                      float x) {
  const TabPos tp = tab_pos(P, off, n, x);
  TabD t;
  t.v = tab_at(P, off, n, tp);
  t.s0 = slots[off + n + tp.i];
  t.w0 = 1.0f - tp.f;
  t.s1 = tp.f > 0.0f ? slots[off + n + tp.i + 1] : -1;
  t.w1 = tp.f;
  return t;
}

__device__ __forceinline__ void grad_add(double *sg, int slot, float v) {
  if (slot >= 0 && v != 0.0f) atomicAdd(sg + slot, (double)v);
}

// d(rate)/d(table value) of a length table: dmu = -dL / L^2
struct RateD {
  int s0, s1;
  float d0, d1;
};

__device__ __forceinline__ RateD rate_d(const float *__restrict__ P, const int *__restrict__ slots, int off, int n,
                                        float E) {
  RateD r{-1, -1, 0.0f, 0.0f};
  if (n <= 0) return r;
  const TabD t = tab_d(P, slots, off, n, E);
  if (!(t.v >= TINY_NORMAL)) return r;
  const float k = -1.0f / (t.v * t.v);
  r.s0 = t.s0;
  r.d0 = k * t.w0;
  r.s1 = t.s1;
  r.d1 = k * t.w1;
  return r;
}

// the log-derivatives of the drawn outcome's share (dshare) and of the detected share (ddet) with respect to the
// row's columns (R, T, eff, p_spike, p_lobe, p_back); o < 0: the detected share only
__device__ inline void surface_dlogs(const int *__restrict__ R, const float *__restrict__ P, float E, int o,
                                     float dshare[N_COL], float ddet[N_COL]) {
  for (int c = 0; c < N_COL; ++c) dshare[c] = ddet[c] = 0.0f;
  const float Rv = col_at(R, P, COL_R, E), Tv = col_at(R, P, COL_T, E), eff = col_at(R, P, COL_EFF, E);
  const float iR = (Rv >= 0.0f && Rv <= 1.0f) ? 1.0f : 0.0f;
  const float iRT = (Rv + Tv >= 0.0f && Rv + Tv <= 1.0f) ? 1.0f : 0.0f;
  const float r = clip01(Rv), t = clip01(Rv + Tv) - r, a = 1.0f - clip01(Rv + Tv);
  if (a > 0.0f && !(R[1] & FLAG_PER_NORMAL)) ddet[COL_R] = ddet[COL_T] = -iRT / a;  // per normal: a / A_n = 1
  if (eff != 0.0f) ddet[COL_EFF] = 1.0f / eff;
  if (o < 0) return;
  if (o == STRAIGHT) {
    if (t > 0.0f) {
      dshare[COL_R] = (iRT - iR) / t;
      dshare[COL_T] = iRT / t;
    }
    return;
  }
  if (r > 0.0f) dshare[COL_R] = iR / r;
  if (R[0] != MODEL_REFLECT) return;  // the Fresnel shares: F has no parameter
  const float ps = col_at(R, P, COL_SPIKE, E), pl = col_at(R, P, COL_LOBE, E), pb = col_at(R, P, COL_BACK, E);
  const float k1 = clip01(ps), k2 = clip01(ps + pl), k3 = clip01(ps + pl + pb);
  const float I1 = (ps >= 0.0f && ps <= 1.0f) ? 1.0f : 0.0f;
  const float I2 = (ps + pl >= 0.0f && ps + pl <= 1.0f) ? 1.0f : 0.0f;
  const float I3 = (ps + pl + pb >= 0.0f && ps + pl + pb <= 1.0f) ? 1.0f : 0.0f;
  const bool normal = R[1] & FLAG_NORMAL;
  float x = 0.0f, dps = 0.0f, dpl = 0.0f, dpb = 0.0f;
  if (o == SPECULAR) {
    if (normal) { x = k2; dps = I2; dpl = I2; } else { x = k1; dps = I1; }
  } else if (o == GLOSSY) {
    x = k2 - k1; dps = I2 - I1; dpl = I2;
  } else if (o == BACKSCATTER) {
    x = k3 - k2; dps = I3 - I2; dpl = I3 - I2; dpb = I3;
  } else if (o == DIFFUSE) {
    x = 1.0f - k3; dps = dpl = dpb = -I3;
  }
  if (x > 0.0f) {
    dshare[COL_SPIKE] = dps / x;
    dshare[COL_LOBE] = dpl / x;
    dshare[COL_BACK] = dpb / x;
  }
}

// add scale * sum_c dlog[c] dcol_c / dv into the slots of the row's column tables
__device__ inline void add_columns(double *sg, const int *__restrict__ R, const float *__restrict__ P,
                                   const int *__restrict__ slots, float E, const float dlog[N_COL], float scale) {
  for (int c = 0; c < N_COL; ++c) {
    if (dlog[c] == 0.0f) continue;
    const TabD t = tab_d(P, slots, R[2 + 2 * c], R[3 + 2 * c], E);
    grad_add(sg, t.s0, scale * dlog[c] * t.w0);
    grad_add(sg, t.s1, scale * dlog[c] * t.w1);
  }
}

}  // namespace lh
