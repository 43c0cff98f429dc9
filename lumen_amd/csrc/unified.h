// The unified facet density on the device: one facet reflection's log density in sigma and its derivative, from the
// tables of unified_tables.cpp (bicubic Hermite in (x, log sigma) for F, cubic Hermite in log sigma for lz).
//
// log p(m | V; sigma) = -alpha^2 / (2 sigma^2) + log min(1, f_max / sin alpha) - log(2 pi) - log Z(sigma)
//                       - log F(theta, sigma) (- log(4 V.h), which no ratio or derivative in sigma sees),
// f_max = min(1, 4 sigma), the reference's `_unified_log_density` (src/lumen/surfaces/facets.py:211-229), whose
// tables and grid are `unified_tables` (:127-151) and `table_front` / `table_lz` (:170-208).
#pragma once
#include <hip/hip_runtime.h>

#include "optics.h"

namespace lh {

constexpr int UT_NX = 257, UT_NY = 69, UT_J_LO = -52;
constexpr float UT_Y0 = -1.3862943611198906f, UT_DY = 0.08664339756999316f;
constexpr int UT_WORDS = 4 * UT_NX * UT_NY + 2 * UT_NY;

__device__ __forceinline__ void hermite4(float t, float h[4], float dh[4]) {
  const float t2 = t * t, t3 = t2 * t;
  h[0] = 2 * t3 - 3 * t2 + 1;
  h[1] = t3 - 2 * t2 + t;
  h[2] = -2 * t3 + 3 * t2;
  h[3] = t3 - t2;
  dh[0] = 6 * t2 - 6 * t;
  dh[1] = 3 * t2 - 4 * t + 1;
  dh[2] = -6 * t2 + 6 * t;
  dh[3] = 3 * t2 - 2 * t;
}

__device__ __forceinline__ void ut_cell(float v, int n, int &i, float &t) {
  const float f = v < 0.0f ? 0.0f : (v > n - 1.0f ? n - 1.0f : v);
  i = min(max((int)floorf(f), 0), n - 2);
  t = f - (float)i;
}

// F and its partial derivatives in x and y at (x, y = log sigma)
__device__ inline void ut_front(const float *__restrict__ T, float x, float y, float &F, float &Fx, float &Fy) {
  const float *F0 = T, *FX = T + UT_NX * UT_NY, *FY = T + 2 * UT_NX * UT_NY, *FXY = T + 3 * UT_NX * UT_NY;
  int i, j;
  float tx, ty, hx[4], dhx[4], hy[4], dhy[4];
  ut_cell(x * (UT_NX - 1), UT_NX, i, tx);
  ut_cell((y - UT_Y0) / UT_DY - UT_J_LO, UT_NY, j, ty);
  hermite4(tx, hx, dhx);
  hermite4(ty, hy, dhy);
  const float sx = 1.0f / (UT_NX - 1);
  F = Fx = Fy = 0.0f;
#pragma unroll
  for (int di = 0; di < 2; ++di)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int k = (i + di) * UT_NY + (j + dj);
      const float ax = hx[2 * di], bx = hx[2 * di + 1] * sx, ay = hy[2 * dj], by = hy[2 * dj + 1] * UT_DY;
      const float dax = dhx[2 * di], dbx = dhx[2 * di + 1] * sx, day = dhy[2 * dj], dby = dhy[2 * dj + 1] * UT_DY;
      F += ax * ay * F0[k] + bx * ay * FX[k] + ax * by * FY[k] + bx * by * FXY[k];
      Fx += dax * ay * F0[k] + dbx * ay * FX[k] + dax * by * FY[k] + dbx * by * FXY[k];
      Fy += ax * day * F0[k] + bx * day * FX[k] + ax * dby * FY[k] + bx * dby * FXY[k];
    }
  Fx *= (float)(UT_NX - 1);
  Fy *= 1.0f / UT_DY;
}

__device__ inline void ut_lz(const float *__restrict__ T, float y, float &lz, float &dlz) {
  const float *V = T + 4 * UT_NX * UT_NY, *D = V + UT_NY;
  int j;
  float t, h[4], dh[4];
  ut_cell((y - UT_Y0) / UT_DY - UT_J_LO, UT_NY, j, t);
  hermite4(t, h, dh);
  lz = h[0] * V[j] + h[2] * V[j + 1] + UT_DY * (h[1] * D[j] + h[3] * D[j + 1]);
  dlz = (dh[0] * V[j] + dh[2] * V[j + 1] + UT_DY * (dh[1] * D[j] + dh[3] * D[j + 1])) / UT_DY;
}

// log p of facet m (local frame, unit) for the incoming V (local, V.z > 0) at sigma, without the sigma-free terms;
// dlp: its derivative in sigma
__device__ inline float unified_lp(const float *__restrict__ T, float sigma, float3 V, float3 m, float &dlp) {
  const float s = sqrtf(m.x * m.x + m.y * m.y), c = m.z;
  const float al = atan2_as(s, c);
  const float fmax = fminf(1.0f, 4.0f * sigma);
  const float y = __logf(sigma);
  float lz, dlz, F, Fx, Fy;
  ut_lz(T, y, lz, dlz);
  const float vs = sqrtf(V.x * V.x + V.y * V.y);
  const float x = atan2_as(V.z, sigma * vs) / HALF_PI;
  ut_front(T, x, y, F, Fx, Fy);
  F = fmaxf(F, FLOOR32);
  const float lf = fminf(0.0f, __logf(4.0f * sigma)), ls = __logf(fmaxf(s, fmax));
  const float lp = -al * al / (2.0f * sigma * sigma) + (lf - ls) - 2.0f * y - lz - __logf(F);
  const float dx = (1.0f / HALF_PI) * (-V.z * vs) / (V.z * V.z + sigma * sigma * vs * vs);
  dlp = al * al / (sigma * sigma * sigma) + ((4.0f * sigma < 1.0f && s >= fmax) ? 1.0f / sigma : 0.0f) -
        (2.0f + dlz) / sigma - (Fx * dx + Fy / sigma) / F;
  return lp;
}

}  // namespace lh
