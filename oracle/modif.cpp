// Oracle: constrained realisations (Hoffman-Ribak modifications), restated
// serially from the reference.  TEST INFRASTRUCTURE ONLY.
//
//   filters per level        filters/filterfamily.hpp:79-143
//   inner product            field/multilevelfield.hpp:248-294 (weights of fourier.hpp:110-151)
//   metric                   field/multilevelfield.hpp:318-384, field.hpp:752-800
//   cross-level transfer     field.hpp:804-849 (FILTER_ON_COARSE_GRID), evaluator.hpp:166-206,
//                            grid/virtualgrid.hpp:411-515,576-612
//   deinterpolation          field.hpp:420-480, multilevelgrid.hpp:246-321
//   linear covectors         modifications/linearmodification.hpp:22-412
//   quadratic (variance)     modifications/quadraticmodification.hpp:50-221
//   manager                  modifications/modificationmanager.hpp:80-331
//   chi^2                    multilevelfield.hpp:461-487
//   potential, velocity      cosmology/parameters.hpp:83-106,208-225
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <strings.h>

namespace orc {

namespace {

// Window<T> of doubles (simulation/window.hpp): corners wrapped, the upper one
// through the point just inside it.
struct FWindow {
  double L, lo[3], hi[3];
  double wrap(double v) const {
    if (v < 0) v += L;
    if (v >= L) v -= L;
    return v;
  }
  FWindow(double periodic, const double *a, const double *b) : L(periodic) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = wrap(a[d]);
      hi[d] = std::nextafter(wrap(std::nextafter(b[d], -INFINITY)), INFINITY);
    }
  }
  bool contains(const double *p) const {
    for (int d = 0; d < 3; ++d) {
      bool in = hi[d] >= lo[d] ? (p[d] < hi[d] && p[d] >= lo[d]) : (p[d] < hi[d] || p[d] >= lo[d]);
      if (!in) return false;
    }
    return true;
  }
};

FWindow gridWindow(const Grid &g) {   // Grid::getWindow (grid.hpp:502-504)
  double a[3] = {g.ox, g.oy, g.oz};
  double b[3] = {g.ox + g.size, g.oy + g.size, g.oz + g.size};
  return FWindow(g.periodic, a, b);
}

void centroid(const Grid &g, size_t i, double *p) {
  size_t n = g.n;
  size_t x = i / (n * n), y = (i / n) % n, z = i % n;
  p[0] = x * g.dx + g.ox + g.dx / 2;
  p[1] = y * g.dx + g.oy + g.dx / 2;
  p[2] = z * g.dx + g.oz + g.dx / 2;
}

// fine -> coarse: each coarse cell whose section corner cells lie in the fine
// grid gets the mean of its contained fine subcells (SubSampleEvaluator)
void addFineToCoarse(Field &coarse, const Field &fine) {
  const Grid &gc = *coarse.g, &gf = *fine.g;
  int r = (int)llround(gc.dx / gf.dx);
  int cellOffset[3] = {(int)llround((gc.ox - gf.ox) / gf.dx), (int)llround((gc.oy - gf.oy) / gf.dx),
                       (int)llround((gc.oz - gf.oz) / gf.dx)};
  int simf = gf.simEquivalentSize();
  int nsec = gc.n * r;
  auto wrapc = [simf](int v) {
    if (v > simf - 1) v -= simf;
    if (v < 0) v += simf;
    return v;
  };
  auto underlying = [&](int sx, int sy, int sz, int *u) {
    if (sx < 0 || sy < 0 || sz < 0 || sx >= nsec || sy >= nsec || sz >= nsec) return false;
    u[0] = wrapc(sx + cellOffset[0]);
    u[1] = wrapc(sy + cellOffset[1]);
    u[2] = wrapc(sz + cellOffset[2]);
    return u[0] >= 0 && u[1] >= 0 && u[2] >= 0 && u[0] < gf.n && u[1] < gf.n && u[2] < gf.n;
  };
  int n = gc.n;
  for (int X = 0; X < n; ++X)
    for (int Y = 0; Y < n; ++Y)
      for (int Z = 0; Z < n; ++Z) {
        int u[3];
        bool c0 = underlying(X * r, Y * r, Z * r, u);
        bool c1 = underlying(X * r + r - 1, Y * r + r - 1, Z * r + r - 1, u);
        if (!c0 && !c1) continue;
        double sum = 0;
        int cnt = 0;
        for (int i = 0; i < r; ++i)
          for (int j = 0; j < r; ++j)
            for (int k = 0; k < r; ++k)
              if (underlying(X * r + i, Y * r + j, Z * r + k, u)) {
                sum += fine.d[(size_t(u[0]) * gf.n + u[1]) * gf.n + u[2]];
                ++cnt;
              }
        coarse.d[(size_t(X) * n + Y) * n + Z] += sum / cnt;
      }
}

// addFieldFromDifferentGridWithFilter, FILTER_ON_COARSE_GRID branch
void addFromOtherGridWithFilter(Field &target, const Field &source, const std::function<double(double)> &filt) {
  Field temp = source;
  temp.applyFilterFn(filt);
  temp.toReal();
  target.toReal();
  if (source.g->dx > target.g->dx) addFieldFromDifferentGrid(target, temp);
  else addFineToCoarse(target, temp);
}

// applyFilterInWindow (field.hpp:752-775)
void applyFilterInWindow(Field &f, const std::function<double(double)> &filt, const FWindow &w, bool windowFirst) {
  Field in = f;
  const Grid &g = *f.g;
  auto zeroOutside = [&](Field &x) {
    x.toReal();
    double p[3];
    for (size_t i = 0; i < g.n3(); ++i) {
      centroid(g, i, p);
      if (!w.contains(p)) x.d[i] = 0;
    }
  };
  if (windowFirst) {
    zeroOutside(in);
    in.toFourier();
    in.applyFilterFn(filt);
  } else {
    in.toFourier();
    in.applyFilterFn(filt);
    zeroOutside(in);
  }
  f.toReal();
  double p[3];
  for (size_t i = 0; i < g.n3(); ++i) {
    centroid(g, i, p);
    if (w.contains(p)) f.d[i] = 0;
  }
  if (in.fourier) f.toFourier();
  else f.toReal();
  for (size_t i = 0; i < f.d.size(); ++i) f.d[i] += in.d[i];
}

// LocalUnitTricubicApproximation weights: the interpolation is a tensor
// product of Catmull-Rom weights, so its transpose elements are the same
// products (tricubic.hpp:450-553)
double crw(int i, double t) {
  double t2 = t * t, t3 = t2 * t;
  switch (i) {
    case 0: return -0.5 * t3 + t2 - 0.5 * t;
    case 1: return 1.5 * t3 - 2.5 * t2 + 1.0;
    case 2: return -1.5 * t3 + 2.0 * t2 + 0.5 * t;
    default: return 0.5 * t3 - 0.5 * t2;
  }
}

// Field::deInterpolate (field.hpp:420-480), with its index arithmetic:
// coordinates are wrapped by the grid size only when < 0 or > size, and the
// linear index is formed on the contiguous real-space layout.
void deInterpolate(Field &lores, const double *loc0, double value) {
  const Grid &g = *lores.g;
  double loc[3] = {loc0[0] - g.ox, loc0[1] - g.oy, loc0[2] - g.oz};
  int p0[3];
  double t[3];
  for (int a = 0; a < 3; ++a) {
    double v = fmod(loc[a], g.periodic);
    if (v < 0) v += g.periodic;
    double u = v / g.dx - 0.5;
    p0[a] = (int)floor(u);
    t[a] = u - p0[a];
  }
  int ix[3][4];
  long n = g.n;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 4; ++i) {
      int c = p0[a] + i - 1;
      if (c < 0) c += (int)n;
      if (c > n) c -= (int)n;
      ix[a][i] = c;
    }
  // getIndexFromCoordinate of the key cell wraps with the simulation size
  int sim = g.simEquivalentSize();
  long key[3];
  for (int a = 0; a < 3; ++a) {
    long c = ix[a][0];
    if (c > sim - 1) c -= sim;
    if (c < 0) c += sim;
    key[a] = c;
  }
  long keyIndex = (key[0] * n + key[1]) * n + key[2];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        long idx = keyIndex + (ix[0][i] - ix[0][0]) * n * n + (ix[1][j] - ix[1][0]) * n + (ix[2][k] - ix[2][0]);
        double w = crw(i, t[0]) * crw(j, t[1]) * crw(k, t[2]);
        if (idx >= 0 && size_t(idx) < lores.d.size()) lores.d[idx] += value * w;
      }
}

}  // namespace

// ---------------------------------------------------------------- filters -
double IC::levelFilter(int level, double k) const {
  int nl = (int)grids.size();
  if (combined) return level == nl - 1 ? 1.0 : 0.0;
  double f = 1.0;
  for (int j = 0; j < level; ++j) f *= highPassCov(k, kcutForLevel(j));
  if (level < nl - 1) f *= lowPassFermi(k, kcutForLevel(level));
  return f;
}

double IC::innerProduct(const std::vector<Field> &a, const std::vector<Field> &b) const {
  double r = 0;
  for (size_t l = 0; l < a.size(); ++l) {
    Field fa = a[l], fb = b[l];
    fa.toFourier();
    fb.toFourier();
    r += fa.accumulateK([&](cplx va, int kx, int ky, int kz) {
            return cplx(std::real(std::conj(va) * fb.getK(kx, ky, kz)), 0);
          }).real();
  }
  return r;
}

// applyMetric (multilevelfield.hpp:335-384)
void IC::applyMetric(std::vector<Field> &f, bool toCovector) const {
  size_t nl = f.size();
  for (auto &x : f) x.toFourier();
  std::vector<Field> out;
  for (size_t l = 0; l < nl; ++l) {
    Field res = f[l];
    auto Fl = [this, l](double k) { return levelFilter(int(l), k); };
    if (toCovector && l + 1 < nl) {
      FWindow w = gridWindow(*grids[l + 1]);
      applyFilterInWindow(res, Fl, w, true);
      applyFilterInWindow(res, Fl, w, false);
    } else {
      res.applyFilterFn([&](double k) { return Fl(k) * Fl(k); });
    }
    for (size_t s = 0; s < nl; ++s) {
      if (s == l) continue;
      double ratio = weights[l] / weights[s];
      double sq = sqrt(ratio);
      addFromOtherGridWithFilter(res, f[s], [&](double k) {
        return levelFilter(int(l), k) * levelFilter(int(s), k) * sq;
      });
    }
    res.toFourier();
    out.push_back(std::move(res));
  }
  f = std::move(out);
}

// applyTransferRatio from the field's species to dm (multilevelfield.hpp:421-454)
void IC::transferToDm(std::vector<Field> &f, bool fromWhiteNoise) const {
  for (size_t l = 0; l < f.size(); ++l) {
    Field Pdm, Pall;
    powerSpectrumField(*camb, *grids[l], Pdm, false);
    f[l].toFourier();
    if (!fromWhiteNoise) {
      powerSpectrumField(*camb, *grids[l], Pall, true);
      f[l].forEachK([&](cplx v, int kx, int ky, int kz) {
        double s = Pall.getK(kx, ky, kz).real();
        if (s != 0.0) s = pow(s, -0.5);
        return v * s;
      });
    }
    f[l].forEachK([&](cplx v, int kx, int ky, int kz) {
      double s = Pdm.getK(kx, ky, kz).real();
      if (s != 0.0) s = pow(s, 0.5);
      return v * s;
    });
  }
}

int IC::finestFlaggedCheck(const std::string &name, const std::vector<std::vector<size_t>> &fl) const {
  (void)name;
  int L = (int)grids.size() - 1;
  if (fl[L].empty())
    throw std::runtime_error(
        "Linear modifications use exclusively information from the highest resolution grid but no cells have been "
        "flagged on this grid.\nChange flagged cells selection to use linear modifications.");
  return L;
}

// LinearModification::getCovector (linearmodification.hpp:47-77) and the
// subclasses' localisation / modification covectors
std::vector<Field> IC::linearCovector(const std::string &name, const std::vector<std::vector<size_t>> &fl) const {
  int L = finestFlaggedCheck(name, fl);
  const Grid &g = *grids[L];
  const std::vector<size_t> &cells = fl[L];
  int kind;   // 0 overdensity, 1 potential, 2-4 velocity, 5-7 angular momentum
  if (!strcasecmp(name.c_str(), "overdensity")) kind = 0;
  else if (!strcasecmp(name.c_str(), "potential")) kind = 1;
  else if (!strcasecmp(name.c_str(), "vx")) kind = 2;
  else if (!strcasecmp(name.c_str(), "vy")) kind = 3;
  else if (!strcasecmp(name.c_str(), "vz")) kind = 4;
  else if (!strcasecmp(name.c_str(), "Lx")) kind = 5;
  else if (!strcasecmp(name.c_str(), "Ly")) kind = 6;
  else if (!strcasecmp(name.c_str(), "Lz")) kind = 7;
  else throw std::runtime_error(name + " is an unknown modification name");
  if (kind >= 2 && kind <= 4 && combined)
    throw std::runtime_error("Cannot create velocity covector for recombined fields (i.e. after output generation is complete)");
  Field hr(g, false);
  int n = g.n, sim = g.simEquivalentSize();
  auto step = [&](size_t index, int dir, int s) -> long {
    long c[3] = {long(index / (size_t(n) * n)), long((index / n) % n), long(index % n)};
    c[dir] += s;
    for (int a = 0; a < 3; ++a) {
      if (c[a] > sim - 1) c[a] -= sim;
      if (c[a] < 0) c[a] += sim;
    }
    return (c[0] * n + c[1]) * n + c[2];
  };
  auto add = [&](long i, double v) {
    if (i >= 0 && size_t(i) < g.n3()) hr.d[i] += v;
  };
  double w = 1.0 / cells.size();
  if (kind <= 1) {
    for (size_t c : cells) hr.d[c] += w;
  } else if (kind <= 4) {
    int dir = kind - 2;
    double a = -w / 12. / g.dx, b = w * 2. / 3. / g.dx;
    for (size_t c : cells) {
      long m1 = step(c, dir, -1), p1 = step(c, dir, +1);
      long m2 = step(size_t(m1), dir, -1), p2 = step(size_t(p1), dir, +1);
      add(m2, a);
      add(m1, b);
      add(p1, -b);
      add(p2, -a);
    }
  } else {
    int dir = kind - 5;
    // getCentreWrapped of the flagged cells (grid.hpp:466-494)
    double p0[3], s[3] = {0, 0, 0}, q[3];
    centroid(g, cells[0], p0);
    auto woff = [&](double x0, double x1) {
      double r = x0 - x1;
      if (r > g.periodic / 2) r -= g.periodic;
      if (r < -g.periodic / 2) r += g.periodic;
      return r;
    };
    for (size_t i = 1; i < cells.size(); ++i) {
      centroid(g, cells[i], q);
      for (int d = 0; d < 3; ++d) s[d] += woff(q[d], p0[d]);
    }
    double qc[3];
    for (int d = 0; d < 3; ++d) {
      double v = s[d] / cells.size() + p0[d];
      v = fmod(v, g.periodic);
      if (v < 0) v += g.periodic;
      qc[d] = v;
    }
    double norm = cells.size() * g.dx;
    double a = 1. / 12. / norm, b = -2. / 3. / norm;
    int dp1 = (dir + 1) % 3, dp2 = (dir + 2) % 3;
    for (size_t c : cells) {
      centroid(g, c, q);
      double dq[3];
      for (int d = 0; d < 3; ++d) dq[d] = woff(qc[d], q[d]);
      double cross[3];
      cross[dir] = 0;
      cross[dp2] = dq[dp1];
      cross[dp1] = -dq[dp2];
      for (int d = 0; d < 3; ++d) {
        if (d == dir) continue;
        long m1 = step(c, d, -1), m2 = step(size_t(m1), d, -1);
        long p1 = step(c, d, +1), p2 = step(size_t(p1), d, +1);
        add(m2, cross[d] * a);
        add(m1, cross[d] * b);
        add(p1, -cross[d] * b);
        add(p2, -cross[d] * a);
      }
    }
  }
  hr.toFourier();
  // generateMultilevelCovectorFromHiresCovector (multilevelgrid.hpp:246-321)
  size_t nl = grids.size();
  std::vector<Field> f;
  for (size_t l = 0; l < nl; ++l) f.push_back(l == size_t(L) ? hr : Field(*grids[l], false));
  if (L > 0) {
    f[L].toReal();
    for (int l = L - 1; l >= 0; --l) {
      Field &hi = f[l + 1], &lo = f[l];
      const Grid &gh = *hi.g;
      int r = (int)llround(lo.g->dx / gh.dx);
      double vr = double(r) * r * r;
      double p[3];
      for (size_t i = 0; i < gh.n3(); ++i) {
        centroid(gh, i, p);
        deInterpolate(lo, p, hi.d[i] / vr);
      }
    }
  }
  for (size_t l = 0; l < nl; ++l) {
    double s = weights[l] / weights[nl - 1];
    for (auto &v : f[l].d) v *= s;
  }
  for (size_t l = 0; l < nl; ++l) {
    f[l].applyFilterFn([this, l](double k) { return levelFilter(int(l), k); });
    f[l].toReal();
  }
  // turnLocalisationCovectorIntoModificationCovector
  transferToDm(f, !powerApplied);
  if (kind == 1) {
    const double prefac = 3. / 2. * cosmo.OmegaM0 / cosmo.scalefactor * 100. * 100. / (299792.) / (299792.);
    for (auto &x : f) {
      double kMin = x.g->kmin();
      x.forEachK([&](cplx v, int ix, int iy, int iz) {
        double k1 = ix * kMin, k2 = iy * kMin, k3 = iz * kMin;
        double ksq = k1 * k1 + k2 * k2 + k3 * k3;
        double inv = 1.0 / ksq;
        if (k2 == 0.0) inv = 0.0;
        return -prefac * v * inv;
      });
    }
  } else if (kind >= 2) {
    double scale = zeldovichVelocityToOffsetRatio(cosmo);
    for (auto &x : f) {
      double kMin = x.g->kmin();
      x.forEachK([&](cplx v, int ix, int iy, int iz) {
        double k1 = ix * kMin, k2 = iy * kMin, k3 = iz * kMin;
        double k2s = k1 * k1 + k2 * k2 + k3 * k3;
        if (k2s == 0) return cplx(0);
        return -scale * v / k2s;
      });
    }
  }
  for (auto &x : f) x.toFourier();
  return f;
}

// FilteredVarianceModification::pushMultiLevelFieldThroughMatrix
// (quadraticmodification.hpp:58-92,152-221)
std::vector<Field> IC::pushVariance(const std::vector<Field> &field, const std::vector<std::vector<size_t>> &fl,
                                    double scale) const {
  std::vector<Field> f = field;
  transferToDm(f, !powerApplied);
  double kcut = 2 * M_PI / scale;
  auto hp = [kcut](double k) { return 1. - lowPassFermi(k, kcut); };
  for (size_t l = 0; l < f.size(); ++l) {
    Field &x = f[l];
    const std::vector<size_t> &cells = fl[l];
    auto window = [&]() {
      x.toReal();
      std::vector<double> keep(x.d.size(), 0.0);
      for (size_t c : cells) keep[c] = x.d[c];
      x.d.swap(keep);
    };
    window();
    x.applyFilterFn(hp);
    x.toReal();
    window();
    double sum = 0;
    for (size_t c : cells) sum += x.d[c];
    size_t m = cells.size();
    for (size_t c : cells) {
      x.d[c] -= sum / m;
      x.d[c] /= m;
    }
    x.applyFilterFn(hp);
    x.toReal();
    window();
  }
  applyMetric(f, true);
  transferToDm(f, !powerApplied);
  return f;
}

double IC::varianceValue(const std::vector<Field> &field, const std::vector<std::vector<size_t>> &fl,
                         double scale) const {
  auto pushed = pushVariance(field, fl, scale);
  return innerProduct(pushed, field);
}

void IC::checkVarianceScale(double scale) const {
  if (scale < 0.0)
    throw std::runtime_error(
        "Trying to calculate filtered variance without initialising variance filtering scale. Use filtering_scale "
        "command to do this.");
  if (scale < grids.back()->dx)
    throw std::runtime_error("Variance high-pass filtering scale is smaller than the smallest pixel size.");
}

double IC::calculate(const std::string &name) {
  if (!drawn) draw();
  auto fl = flags;
  for (auto &x : noise) x.toFourier();
  if (!strcasecmp(name.c_str(), "variance")) {
    checkVarianceScale(varianceScale);
    return varianceValue(noise, fl, varianceScale);
  }
  auto cov = linearCovector(name, fl);
  return innerProduct(cov, noise);
}

void IC::addModification(const std::string &name, bool relative, double target) {
  if (!drawn) draw();
  Mod m;
  m.name = name;
  m.flags = flags;
  m.scale = varianceScale;
  m.quadratic = !strcasecmp(name.c_str(), "variance");
  if (m.quadratic) checkVarianceScale(varianceScale);
  else linearCovector(name, flags);   // validates the name and the flags
  if (relative) target *= calculate(name);
  m.target = target;
  (m.quadratic ? quadratic : linear).push_back(m);
}

double IC::chi2() {
  if (!drawn) draw();
  if (powerApplied) throw std::runtime_error("Cannot calculate the chi^2 after the output field has been generated");
  std::vector<Field> c = noise;
  applyMetric(c, true);
  return innerProduct(c, noise);
}

size_t IC::numDof() const {
  size_t dof = 0;
  for (size_t i = 0; i < grids.size(); ++i) {
    dof += grids[i]->n3();
    if (i > 0) {
      size_t f = (size_t)llround(grids[i - 1]->dx / grids[i]->dx);
      dof -= grids[i]->n3() / (f * f * f);
    }
  }
  return dof;
}

// ModificationManager::applyModifications (modificationmanager.hpp:80-115)
void IC::applyModifications(double *preChi2, double *postChi2) {
  if (!drawn) draw();
  double pre = chi2();
  std::vector<std::vector<Field>> alphas;
  std::vector<double> targets;
  for (auto &m : linear) {
    alphas.push_back(linearCovector(m.name, m.flags));
    targets.push_back(m.target);
  }
  // orthonormaliseModifications: covector . covector goes through the metric
  auto covDot = [&](const std::vector<Field> &a, const std::vector<Field> &b) {
    std::vector<Field> bv = b;
    applyMetric(bv, false);
    return innerProduct(a, bv);
  };
  for (size_t i = 0; i < alphas.size(); ++i) {
    for (size_t j = 0; j < i; ++j) {
      double r = covDot(alphas[i], alphas[j]);
      for (size_t l = 0; l < alphas[i].size(); ++l)
        for (size_t q = 0; q < alphas[i][l].d.size(); ++q) alphas[i][l].d[q] -= r * alphas[j][l].d[q];
      targets[i] -= r * targets[j];
    }
    double norm = sqrt(covDot(alphas[i], alphas[i]));
    for (auto &x : alphas[i])
      for (auto &v : x.d) v /= norm;
    targets[i] /= norm;
  }
  // applyLinearModif: the alphas become vectors
  std::vector<double> existing;
  for (auto &a : alphas) existing.push_back(innerProduct(a, noise));
  for (size_t i = 0; i < alphas.size(); ++i) {
    double dval = targets[i] - existing[i];
    applyMetric(alphas[i], false);
    for (size_t l = 0; l < noise.size(); ++l) {
      noise[l].toFourier();
      for (size_t q = 0; q < noise[l].d.size(); ++q) noise[l].d[q] += dval * alphas[i][l].d[q];
    }
  }
  // applyLinQuadModif (modificationmanager.hpp:174-246)
  for (auto &m : quadratic) {
    auto iterate = [&](std::vector<Field> &field, int nsteps) {
      double start = varianceValue(field, m.flags, m.scale);
      std::vector<double> tg(nsteps + 1);
      double stepv = (m.target - start) / nsteps;   // linspace (tools/util_functions.hpp:78-91)
      for (int i = 0; i <= nsteps; ++i) tg[i] = start + i * stepv;
      for (int i = 0; i < nsteps; ++i) {
        double cur = varianceValue(field, m.flags, m.scale);
        auto pushed = pushVariance(field, m.flags, m.scale);
        double norm = sqrt(covDot(pushed, pushed));
        for (auto &a : alphas) {   // vectors now: plain products
          double r = innerProduct(pushed, a);
          for (size_t l = 0; l < pushed.size(); ++l)
            for (size_t q = 0; q < pushed[l].d.size(); ++q) pushed[l].d[q] -= r * a[l].d[q];
        }
        double n2 = sqrt(covDot(pushed, pushed));
        for (auto &x : pushed)
          for (auto &v : x.d) v /= n2;
        double mult = 0.5 * (tg[i + 1] - cur) / norm;
        applyMetric(pushed, false);
        for (size_t l = 0; l < field.size(); ++l) {
          field[l].toFourier();
          for (size_t q = 0; q < field[l].d.size(); ++q) field[l].d[q] += mult * pushed[l].d[q];
        }
      }
    };
    int init = m.initSteps;
    std::vector<Field> test = noise;
    iterate(test, init);
    double achieved = std::fabs(varianceValue(test, m.flags, m.scale) - m.target);
    double want = m.target * m.precision;
    int nsteps = init * (int)std::ceil(std::sqrt(achieved / want));
    iterate(noise, nsteps > init ? nsteps : init);
  }
  double post = chi2();
  if (preChi2) *preChi2 = pre;
  if (postChi2) *postChi2 = post;
}

void IC::clearModifications() {
  linear.clear();
  quadratic.clear();
}

}  // namespace orc
