// The unified facet model's tables, computed once per process on the host in FP64 and stored in float32:
// F(x, y), the share of facets facing the photon, on x = (2/pi) atan2(cos theta, sigma sin theta) in [0, 1] (257
// nodes) and y = log sigma (69 nodes, log(1/4) + k log(2)/8 for k in [-52, 16]), with its derivatives Fx, Fy, Fxy for
// a bicubic Hermite interpolant; and lz(y) = log(Z(sigma) / sigma^2) with d lz / dy, Z the normaliser of Geant4's
// tilt density.  The definitions and grids are the reference's (src/lumen/surfaces/facets.py:65-151,
// src/lumen/rules/choices.py:101-104): composite Gauss-Legendre quadrature of order 16 (64 panels for Z split at
// the f_max kink, 16 panels for F in tau with alpha = alpha_k + span tau^2), derivatives by central differences
// (step 1e-5; 1e-6 one-sided at the x ends), Fxy by second-order differences of Fx along y.
#include <cmath>
#include <mutex>
#include <vector>

namespace lh_host {

constexpr int NX = 257, NY = 69, J_LO = -52, ORDER = 16, PANELS_Z = 64, PANELS_F = 16;
constexpr double Y0 = -1.3862943611198906;       // log(1/4)
constexpr double DY = 0.6931471805599453 / 8.0;  // log(2) / 8
constexpr double CUT_SIGMAS = 14.0, FD = 1e-5, FD_END = 1e-6;
constexpr double HALF_PI = 1.5707963267948966, PI = 3.141592653589793;

// Gauss-Legendre nodes and weights of order n on [-1, 1] (Newton on the Legendre recurrence)
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-15) break;
    }
    x[i] = -z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

struct Quad {
  double x[ORDER], w[ORDER];
  Quad() { gauss_legendre(ORDER, x, w); }
};
static const Quad &quad() {
  static Quad q;
  return q;
}

// Geant4's unified tilt density, unnormalised: exp(-a^2 / 2 s^2) min(sin a, f_max), f_max = min(1, 4 s)
static inline double tilt(double a, double s) {
  return std::exp(-a * a / (2.0 * s * s)) * std::min(std::sin(a), std::min(1.0, 4.0 * s));
}

template <class F> static double integrate(double a, double b, int panels, F f) {
  const Quad &q = quad();
  double tot = 0.0;
  for (int p = 0; p < panels; ++p) {
    const double lo = a + (b - a) * p / panels, hi = a + (b - a) * (p + 1) / panels;
    for (int k = 0; k < ORDER; ++k) tot += 0.5 * (hi - lo) * q.w[k] * f(0.5 * (lo + hi) + 0.5 * (hi - lo) * q.x[k]);
  }
  return tot;
}

static double tilt_norm(double s) {
  const double top = std::min(HALF_PI, CUT_SIGMAS * s);
  const double af = std::min(std::asin(std::min(1.0, 4.0 * s)), top);
  double tot = 0.0;
  if (af > 0.0) tot += integrate(0.0, af, PANELS_Z, [&](double a) { return tilt(a, s); });
  if (top > af) tot += integrate(af, top, PANELS_Z, [&](double a) { return tilt(a, s); });
  return tot;
}

// the share of facets facing the photon at x (1 - the tilt-weighted share of azimuths facing away)
static double front(double x, double s, double z) {
  const bool inf = x >= 1.0;
  const double c = inf ? INFINITY : s * std::tan(HALF_PI * std::min(x, 1.0));
  const double ak = std::atan(c);
  const double top = std::min(HALF_PI, ak + CUT_SIGMAS * s);
  const double span = top - ak;
  const double tf =
      std::sqrt(std::min(std::max((std::asin(std::min(1.0, 4.0 * s)) - ak) / (span > 0 ? span : 1.0), 0.0), 1.0));
  auto f = [&](double t) {
    const double al = ak + span * t * t;
    const double arg = inf ? 1.0 : std::min(std::max(c * std::cos(al) / std::sin(al), -1.0), 1.0);
    return tilt(al, s) * std::acos(arg) / PI * 2.0 * span * t;
  };
  double acc = 0.0;
  if (tf > 0.0) acc += integrate(0.0, tf, PANELS_F, f);
  if (tf < 1.0) acc += integrate(tf, 1.0, PANELS_F, f);
  return 1.0 - acc / z;
}

// tables as float32: F, Fx, Fy, Fxy (NX x NY, row-major in x), then lz, lzy (NY)
const std::vector<float> &unified_tables() {
  static std::vector<float> out;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<double> F(NX * NY), Fx(NX * NY), Fy(NX * NY), Fxy(NX * NY), lz(NY), lzy(NY);
    for (int j = 0; j < NY; ++j) {
      const double y = Y0 + DY * (J_LO + j), s = std::exp(y);
      const double z = tilt_norm(s), zp = tilt_norm(std::exp(y + FD)), zm = tilt_norm(std::exp(y - FD));
      lz[j] = std::log(z / (s * s));
      lzy[j] = (std::log(zp / std::exp(2 * (y + FD))) - std::log(zm / std::exp(2 * (y - FD)))) / (2 * FD);
      for (int i = 0; i < NX; ++i) {
        const double x = (double)i / (NX - 1);
        double xp = std::min(x + FD, 1.0), xm = std::max(x - FD, 0.0);
        if (i == 0) xp = FD_END;
        if (i == NX - 1) xm = 1.0 - FD_END;
        F[i * NY + j] = front(x, s, z);
        Fx[i * NY + j] = (front(xp, s, z) - front(xm, s, z)) / (xp - xm);
        Fy[i * NY + j] = (front(x, std::exp(y + FD), zp) - front(x, std::exp(y - FD), zm)) / (2 * FD);
      }
    }
    for (int i = 0; i < NX; ++i)  // numpy.gradient along y, second order at the ends
      for (int j = 0; j < NY; ++j) {
        const double *c = &Fx[i * NY];
        double d;
        if (j == 0) d = (-3 * c[0] + 4 * c[1] - c[2]) / (2 * DY);
        else if (j == NY - 1) d = (3 * c[NY - 1] - 4 * c[NY - 2] + c[NY - 3]) / (2 * DY);
        else d = (c[j + 1] - c[j - 1]) / (2 * DY);
        Fxy[i * NY + j] = d;
      }
    out.reserve(4 * NX * NY + 2 * NY);
    for (const auto *v : {&F, &Fx, &Fy, &Fxy})
      for (double x : *v) out.push_back((float)x);
    for (double x : lz) out.push_back((float)x);
    for (double x : lzy) out.push_back((float)x);
  });
  return out;
}

}  // namespace lh_host
