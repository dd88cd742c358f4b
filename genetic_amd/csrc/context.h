// Internal state of a gic_context and the pipeline steps shared by the
// C-ABI (api.hip) and the constrained-realisation engine (modif.hip).
#pragma once
#include "../../include/gic.h"
#include "gic_common.h"
#include "fft.h"
#include "rng.h"
#include "kspace.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace gic {

// ---------------------------------------------------------------- host maths --
struct Cosmo {
  double Om = 0.3, Ol = 0, Ob = 0, h = 0.701, s8 = 0, ns = 0.96, z = 0, a = 1;
};

// growthFactor, cosmology/parameters.hpp:55-70 (a^3 through single-precision powf there)
inline double growth_factor(double a, double Om, double Ol) {
  double Hsq = Om / (double)powf((float)a, 3.0f) + (1. - Om - Ol) / a / a + Ol;
  return 2.5 * a * Om / pow(a, 3.0) / Hsq /
         (pow(Om / Hsq / a / a / a, (double)(4.f / 7.f)) - Ol / Hsq + (1. + 0.5 * Om / pow(a, 3.0) / Hsq) * (1. + 1. / 70. * Ol / Hsq));
}

// Natural cubic spline (the gsl_interp_cspline construction): second-derivative
// coefficients from a symmetric tridiagonal system with zero end conditions.
struct Spline {
  std::vector<double> x, y, c;
  void init(const std::vector<double> &xs, const std::vector<double> &ys) {
    x = xs;
    y = ys;
    size_t n = x.size();
    c.assign(n, 0.0);
    if (n < 3) return;
    size_t m = n - 2;
    std::vector<double> dg(m), of(m), g(m);
    for (size_t i = 0; i < m; ++i) {
      double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
      of[i] = h1;
      dg[i] = 2.0 * (h0 + h1);
      g[i] = 3.0 * ((y[i + 2] - y[i + 1]) * (h1 != 0 ? 1.0 / h1 : 0.0) - (y[i + 1] - y[i]) * (h0 != 0 ? 1.0 / h0 : 0.0));
    }
    if (m == 1) {
      c[1] = g[0] / dg[0];
      return;
    }
    std::vector<double> gam(m), al(m), zz(m);
    al[0] = dg[0];
    gam[0] = of[0] / al[0];
    for (size_t i = 1; i + 1 < m; ++i) {
      al[i] = dg[i] - of[i - 1] * gam[i - 1];
      gam[i] = of[i] / al[i];
    }
    al[m - 1] = dg[m - 1] - of[m - 2] * gam[m - 2];
    zz[0] = g[0];
    for (size_t i = 1; i < m; ++i) zz[i] = g[i] - gam[i - 1] * zz[i - 1];
    c[m] = zz[m - 1] / al[m - 1];
    for (long i = long(m) - 2; i >= 0; --i) c[i + 1] = zz[i] / al[i] - gam[i] * c[i + 2];
  }
  double eval(double xv) const {
    size_t n = x.size();
    if (xv < x[0] || xv > x[n - 1]) return 0;
    size_t lo = 0, hi = n - 1;
    while (hi > lo + 1) {
      size_t i = (hi + lo) / 2;
      if (x[i] > xv) hi = i;
      else lo = i;
    }
    double dx = x[lo + 1] - x[lo];
    if (!(dx > 0)) return y[lo];
    double dy = y[lo + 1] - y[lo], dl = xv - x[lo];
    double bi = dy / dx - dx * (c[lo + 1] + 2.0 * c[lo]) / 3.0, di = (c[lo + 1] - c[lo]) / (3.0 * dx);
    return y[lo] + dl * (bi + dl * (c[lo] + dl * di));
  }
};

struct LogSpline {
  Spline s;
  double minx = 0, maxx = 0;
  void init(const std::vector<double> &xv, const std::vector<double> &yv) {
    std::vector<double> lx, ly;
    for (size_t i = 0; i < xv.size(); ++i) {
      double a = std::log(xv[i]), b = std::log(yv[i]);
      if (std::isfinite(a) && std::isfinite(b)) {
        lx.push_back(a);
        ly.push_back(b);
      }
    }
    if (lx.size() < 2)
      throw std::range_error(
          "Could not find enough valid values in the transfer function. This may mean that one of the transfer "
          "function columns is full of zeros.");
    s.init(lx, ly);
    minx = lx.front();
    maxx = lx.back();
  }
  double operator()(double k) const {
    double l = std::log(k);
    if (l < minx || l > maxx) return 0;
    return std::exp(s.eval(l));
  }
};

// ---------------------------------------------------------------- timing -----
struct EventTimer : TimingHook {
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> totals;
  std::map<std::string, double> algo_bytes;   // summed over launches
  void bytes(const char *name, double b) override { algo_bytes[name] += b; }
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    GIC_HIP(hipEventCreate(&e));
    return e;
  }
  void begin(const char *name, hipStream_t st) override {
    Rec r{name, get(), nullptr};
    GIC_HIP(hipEventRecord(r.a, st));
    pending.push_back(r);
  }
  void end(const char *name, hipStream_t st) override {
    for (auto it = pending.rbegin(); it != pending.rend(); ++it)
      if (it->name == name && it->b == nullptr) {
        it->b = get();
        GIC_HIP(hipEventRecord(it->b, st));
        break;
      }
  }
  void collect() {
    for (auto &r : pending) {
      if (!r.b) continue;
      GIC_HIP(hipEventSynchronize(r.b));
      float ms = 0;
      GIC_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto &t = totals[r.name];
      t.first += ms;
      t.second += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  ~EventTimer() override {
    for (auto &r : pending) {
      (void)hipEventDestroy(r.a);
      if (r.b) (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    GIC_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;   // one owner per allocation
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
};

struct Level {
  LevelGeom g;
  DevBuf noise;            // white noise, then the overdensity
  DevBuf psi[3];
  DevBuf psi2[3];          // second-order displacements (2LPT)
  DevBuf sqrtp;            // sqrt P by |k|^2, element type T
  DevBuf sqrtp_dm;         // the same for the dark-matter transfer function (covectors)
  bool noise_fourier = false;
  bool dc_pending = false;   // Re(w_0) still to be zeroed after the first r2c
  bool psi_fourier = false;
  bool have_psi = false;
  bool have_psi2 = false;
  DevBuf adj_acc;          // adjoint: the noise gradient's spectrum
  DevBuf adj_u;            // adjoint: gradient arriving from the finer level
  DevBuf adj_tab, adj_lp;  // adjoint: |k|^2 tables of the level's operators
};

}  // namespace gic

struct ModState;   // modif.hip
namespace gic {
struct BatchState;   // batch.hip
}

struct gic_context {
  using Cosmo = gic::Cosmo;
  using LogSpline = gic::LogSpline;
  using SplineHost = gic::SplineHost;
  using DevBuf = gic::DevBuf;
  using Level = gic::Level;
  using EventTimer = gic::EventTimer;
  using TimingHook = gic::TimingHook;
  gic_context();
  ~gic_context();
  int device = 0;
  int prec = 64;
  hipStream_t st = nullptr;
  std::string err;
  Cosmo cosmo;
  // frozen at the camb command
  bool have_camb = false;
  LogSpline tf_all, tf_dm;
  std::vector<double> kpts;
  double amplitude = 0, ns_frozen = 0;
  SplineHost spline_dev, spline_dm_dev;
  DevBuf spl_x, spl_y, spl_c, spl_dm_x, spl_dm_y, spl_dm_c;
  // modifications: flagged cells per level, variance filter scale, the
  // modification lists and their device scratch
  std::vector<std::vector<uint64_t>> flags;
  double variance_scale = -1;
  std::unique_ptr<ModState> mod;
  std::shared_ptr<gic::BatchState> batch;   // streams and buffers of batched realisations
  std::vector<std::unique_ptr<Level>> levels;
  unsigned long seed = 0;
  int seed_mode = 0;
  bool seeded = false, drawn = false, power_applied = false, zeldovich_done = false;
  int lpt_order = 1;   // 1: Zel'dovich, 2: second-order LPT on top
  bool fix_power = false;   // fix_power: unit-modulus white noise after the draw
  DevBuf work, stage, tab_a, tab_b, tab_c, lpt_win, adj_scratch;
  // three padded fields of data transformed along x only (fft_c2r_xpass), so
  // that one x pass serves several displacement or second-derivative fields;
  // after the Zel'dovich pass it holds the base level's X0, X1 for 2LPT
  DevBuf xhalf;
  bool xhalf_base = false;
  std::unique_ptr<EventTimer> timer;
  TimingHook *hook() { return timer.get(); }
  size_t tsize() const { return prec == 32 ? 4 : 8; }
};


namespace gic {

// A named region on the context's stream for the event timer (no-op when
// timing is off); regions nest and overlap the per-kernel records.
// One kernel launch or device copy timed under `name`, with its
// algorithmic HBM bytes (when event timing is on).
struct Timed {
  gic_context *c;
  const char *name;
  Timed(gic_context *cc, const char *n, double bytes) : c(cc), name(n) {
    if (c->hook()) {
      c->hook()->begin(name, c->st);
      c->hook()->bytes(name, bytes);
    }
  }
  ~Timed() {
    if (c->hook()) c->hook()->end(name, c->st);
  }
};

struct Phase {
  gic_context *c;
  const char *name;
  Phase(gic_context *cc, const char *n) : c(cc), name(n) {
    if (c->hook()) c->hook()->begin(name, c->st);
  }
  ~Phase() {
    if (c->hook()) c->hook()->end(name, c->st);
  }
};

// ---- pipeline steps (api.hip) ------------------------------------------------------
Level &level_at(gic_context *c, int level);
void alloc_field(gic_context *c, DevBuf &b, const LevelGeom &g);
double kcut_for_level(gic_context *c, int level);
InterpGeom interp_geom(const LevelGeom &gf, const LevelGeom &gc);
// coarse cells [w0, w0 + wl) (mod n per axis) that the tricubic stencils of a fine level read
void coarse_window(const LevelGeom &gf, const LevelGeom &gc, int *w0, int *wl);
template <typename T> void to_fourier(gic_context *c, DevBuf &b, int n, bool &flag) {
  if (flag) return;
  fft3d<T>((T *)b.p, n, -1, c->st, c->hook());
  flag = true;
}
template <typename T> void to_real(gic_context *c, DevBuf &b, int n, bool &flag) {
  if (!flag) return;
  fft3d<T>((T *)b.p, n, +1, c->st, c->hook());
  flag = false;
}
template <typename T> void build_sqrtp(gic_context *c, Level &L);
template <typename T> void build_sqrtp_dm(gic_context *c, Level &L);
template <typename T> void do_draw(gic_context *c);
template <typename T> void ensure_noise_fourier(gic_context *c, Level &L);
template <typename T> void do_zeldovich(gic_context *c);

// ---- displacement map and its transpose (adjoint.hip) ----------------------------------
// noise[l]: dense n_l^3 white noise; psi[3l+a]: dense displacement components
template <typename T> void do_displacement_forward(gic_context *c, const T *const *noise, T *const *psi);
template <typename T> void do_displacement_adjoint(gic_context *c, const T *const *psi_bar, T *const *noise_bar);

// ---- constrained realisations (modif.hip) ------------------------------------------
template <typename T> double mod_calculate(gic_context *c, const std::string &name);
template <typename T> void mod_add(gic_context *c, const std::string &name, bool relative, double target, int init_steps,
                                   double precision);
template <typename T> void mod_apply(gic_context *c, double *pre, double *post);
void mod_clear(gic_context *c);
template <typename T> double mod_chi2(gic_context *c);
size_t num_dof(gic_context *c);
// filtered variance of an arbitrary white-noise field and its gradient 2 Q w
template <typename T> double mod_variance_gradient(gic_context *c, const T *const *noise, T *const *grad);
// quadratic modification steps: the reference's linearised steps (off) or an
// exact line search along the same direction (on)
void mod_set_line_search(gic_context *c, bool on);
template <typename T> void do_splice(gic_context *c, unsigned long new_seed, int *cg_iterations);
template <typename T>
void do_batch(gic_context *c, int nreal, const uint64_t *seeds, int ncons, const double *cons, T *out, int nlanes,
              bool graph, double *ms);

}  // namespace gic
