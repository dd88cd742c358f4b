// Oracle: cosmology, CAMB transfer-function power spectrum and the natural
// cubic spline it interpolates with.
//   growthFactor / zeldovichVelocityToOffsetRatio: genetIC/src/cosmology/parameters.hpp:55-106
//   CAMB reader, sigma8 normalisation, P(k):       genetIC/src/cosmology/camb.hpp:101-332
//   LogInterpolator over gsl_interp_cspline:       genetIC/src/tools/numerics/interpolation.hpp:20-110
// GSL's natural cubic spline (unpinned distro GSL) is restated from its
// published construction: tridiagonal solve for the second-derivative
// coefficients with c_0 = c_{n-1} = 0, evaluation y_i + d(b + d(c + d e)).
// Pinned by the theory column of the reference's .ps fixtures.
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"
#include <cmath>
#include <stdexcept>
#include <limits>
#include <algorithm>

namespace orc {

double Grid::kmin() const { return 2. * M_PI / size; }
int Grid::simEquivalentSize() const { return (int)llround(periodic / dx); }

double growthFactor(const Cosmology &c) {
  const double a = c.scalefactor, Om = c.OmegaM0, Ol = c.OmegaLambda0;
  // the reference computes a^3 with single-precision powf here
  double Hsq = c.OmegaM0 / (double)powf((float)a, 3.0f) + (1. - Om - Ol) / a / a + Ol;
  double d = 2.5 * a * Om / pow(a, 3.0) / Hsq /
             (pow(Om / Hsq / a / a / a, (double)(4.f / 7.f)) - Ol / Hsq +
              (1. + 0.5 * Om / pow(a, 3.0) / Hsq) * (1. + 1. / 70. * Ol / Hsq));
  return d;
}

double zeldovichVelocityToOffsetRatio(const Cosmology &c) {
  double f = 1.0, Om = c.OmegaM0, Ol = c.OmegaLambda0, a = c.scalefactor;
  return f * 100. * sqrt(Om / (a * a * a) + Ol) * sqrt(a);
}

// --- natural cubic spline -----------------------------------------------------
void CubicSpline::init(const std::vector<double> &xs, const std::vector<double> &ys) {
  x = xs;
  y = ys;
  size_t n = x.size();
  c.assign(n, 0.0);
  if (n < 3) return;
  size_t sys = n - 2;
  std::vector<double> diag(sys), off(sys), g(sys);
  for (size_t i = 0; i < sys; ++i) {
    double hi = x[i + 1] - x[i], hip = x[i + 2] - x[i + 1];
    double gi = hi != 0 ? 1.0 / hi : 0.0, gip = hip != 0 ? 1.0 / hip : 0.0;
    off[i] = hip;
    diag[i] = 2.0 * (hip + hi);
    g[i] = 3.0 * ((y[i + 2] - y[i + 1]) * gip - (y[i + 1] - y[i]) * gi);
  }
  if (sys == 1) {
    c[1] = g[0] / diag[0];
    return;
  }
  // symmetric tridiagonal LDL^T solve
  std::vector<double> gam(sys), alp(sys), z(sys), cc(sys);
  alp[0] = diag[0];
  gam[0] = off[0] / alp[0];
  for (size_t i = 1; i < sys - 1; ++i) {
    alp[i] = diag[i] - off[i - 1] * gam[i - 1];
    gam[i] = off[i] / alp[i];
  }
  alp[sys - 1] = diag[sys - 1] - off[sys - 2] * gam[sys - 2];
  z[0] = g[0];
  for (size_t i = 1; i < sys; ++i) z[i] = g[i] - gam[i - 1] * z[i - 1];
  for (size_t i = 0; i < sys; ++i) cc[i] = z[i] / alp[i];
  c[sys] = cc[sys - 1];
  for (long i = long(sys) - 2; i >= 0; --i) c[i + 1] = cc[i] - gam[i] * c[i + 2];
}

double CubicSpline::eval(double xv) const {
  size_t n = x.size();
  if (xv < x[0] || xv > x[n - 1]) return 0;
  // gsl_interp_bsearch
  size_t lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    size_t i = (hi + lo) / 2;
    if (x[i] > xv) hi = i; else lo = i;
  }
  double xl = x[lo], xh = x[lo + 1], dx = xh - xl;
  if (!(dx > 0)) return y[lo];
  double dy = y[lo + 1] - y[lo], delx = xv - xl;
  double ci = c[lo], cip = c[lo + 1];
  double bi = dy / dx - dx * (cip + 2.0 * ci) / 3.0;
  double di = (cip - ci) / (3.0 * dx);
  return y[lo] + delx * (bi + delx * (ci + delx * di));
}

void LogInterpolator::init(const std::vector<double> &xv, const std::vector<double> &yv) {
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> lx, ly;
  for (size_t i = 0; i < xv.size(); ++i) {
    double a = std::log(xv[i]), b = std::log(yv[i]);
    if (a != inf && a != -inf && !std::isnan(a) && b != inf && b != -inf && !std::isnan(b)) {
      lx.push_back(a);
      ly.push_back(b);
    }
  }
  if (lx.size() < 2) throw std::range_error("Could not find enough valid values in the transfer function.");
  s.init(lx, ly);
  minx = lx.front();
  maxx = lx.back();
}

double LogInterpolator::operator()(double xv) const {
  double lx = std::log(xv);
  if (lx < minx || lx > maxx) return 0;
  return std::exp(s.eval(lx));
}

// --- CAMB --------------------------------------------------------------------
void CAMB::init(const Cosmology &c, const std::vector<double> &input, int numCols) {
  if (numCols != 7 && numCols != 13) throw std::runtime_error("CAMB transfer file doesn't have a sensible number of columns");
  double tnorm = input[1];
  std::vector<double> tdm, tall;
  for (size_t j = 0; j < input.size() / numCols; j++) {
    if (input[numCols * j] > 0) {
      kpts.push_back(input[numCols * j]);
      tdm.push_back(input[numCols * j + 1] / tnorm);
      tall.push_back(input[numCols * j + 6] / tnorm);
    }
  }
  tf_dm.init(kpts, tdm);
  tf_all.init(kpts, tall);
  ns = c.ns;
  // calculateOverallNormalization (camb.hpp:289-297)
  Cosmology c0 = c;
  c0.scalefactor = 1. / (1. + 0.);
  double gf = growthFactor(c) / growthFactor(c0);
  double s8pre = varianceInSphere(8.);
  double lr = (c.sigma8 / s8pre) * gf;
  amplitude = lr * lr;
}

double CAMB::varianceInSphere(double radius) const {
  // camb.hpp:300-328: rectangle rule with the reference's accumulated k
  double s = 0., k, t;
  double amp = 9. / 2. / M_PI / M_PI;
  double kmax = std::min(kpts.back(), 200.0 / radius) * 0.999999;
  double kmin = kpts[0] * 1.000001;
  double dk = (kmax - kmin) / 50000.;
  for (k = kmin; k < kmax; k += dk) {
    t = tf_all(k);
    double kr = k * radius;
    double w = (sin(kr) - kr * cos(kr)) / (kr * kr * kr);
    s += pow(k, ns + 2.) * w * w * t * t;
  }
  return sqrt(s * amp * dk);
}

double CAMB::P(double k, bool all) const {
  double lt = (k != 0) ? (all ? tf_all(k) : tf_dm(k)) : 0.0;
  return amplitude * pow(k, ns) * lt * lt;
}

double CAMB::normalisationForGrid(const Grid &g) {
  double kw = 2. * M_PI / g.size;
  double norm = kw * kw * kw / pow(2.f * M_PI, 3.f);
  return norm * double(g.n3());
}

void powerSpectrumField(const CAMB &camb, const Grid &g, Field &out, bool all) {
  // camb.hpp:101-119 via forEachFourierCell (k from kx*kMin etc.)
  out = Field(g, true);
  double norm = CAMB::normalisationForGrid(g);
  double kMin = g.kmin();
  out.forEachK([&](cplx, int ix, int iy, int iz) {
    double kx = ix * kMin, ky = iy * kMin, kz = iz * kMin;
    double k = sqrt(kx * kx + ky * ky + kz * kz);
    return cplx(camb.P(k, all) * norm, 0);
  });
}

double lowPassFermi(double k, double kcut) {
  double temperature = kcut * 0.1;
  return 1. / (1. + exp((k - kcut) / temperature));
}
double highPassCov(double k, double kcut) {
  double f = lowPassFermi(k, kcut);
  return sqrt(1. - f * f);
}

}  // namespace orc
