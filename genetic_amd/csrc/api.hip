// C-ABI implementation: the ICGenerator field path on one GPU.
//
// The orchestration follows the reference's ICGenerator (genetIC/src/ic.hpp)
// and MultiLevelParticleGenerator (particles/multilevelgenerator.hpp:78-128);
// the arithmetic runs in the kernels of fft.hip, rng.hip and kspace.hip.
#include "context.h"

using namespace gic;

namespace gic {

size_t total_bytes(gic_context *c) {
  size_t b = c->work.bytes + c->stage.bytes + c->tab_a.bytes + c->tab_b.bytes + c->tab_c.bytes + c->spl_x.bytes * 3;
  for (auto &l : c->levels) {
    b += l->noise.bytes + l->sqrtp.bytes;
    for (auto &p : l->psi) b += p.bytes;
  }
  return b;
}

Level &level_at(gic_context *c, int level) {
  if (level < 0 || level >= (int)c->levels.size()) throw std::runtime_error("Specified level does not exist.");
  return *c->levels[level];
}


void alloc_field(gic_context *c, DevBuf &b, const LevelGeom &g) {
  b.ensure(g.padded_reals() * c->tsize());
}

// camb.hpp:300-328: sigma(R) by a 50000-step rectangle rule with accumulated k
double variance_in_sphere(gic_context *c, double radius) {
  double s = 0., k, t;
  double amp = 9. / 2. / M_PI / M_PI;
  double kmax = std::min(c->kpts.back(), 200.0 / radius) * 0.999999;
  double kmin = c->kpts[0] * 1.000001;
  double dk = (kmax - kmin) / 50000.;
  for (k = kmin; k < kmax; k += dk) {
    t = c->tf_all(k);
    double kr = k * radius;
    double w = (sin(kr) - kr * cos(kr)) / (kr * kr * kr);
    s += pow(k, c->ns_frozen + 2.) * w * w * t * t;
  }
  return sqrt(s * amp * dk);
}

template <typename T> void build_sqrtp(gic_context *c, Level &L) {
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
  long h = L.g.n / 2;
  long nent = 3 * h * h + 1;
  L.sqrtp.ensure(nent * sizeof(T));
  double kw = 2. * M_PI / L.g.box;
  double norm = kw * kw * kw / pow(2.f * M_PI, 3.f) * double(L.g.n3());
  sqrtp_table<T>(c->spline_dev, c->amplitude, c->ns_frozen, L.g.kmin(), norm, nent, (T *)L.sqrtp.p, c->st);
}

// the dark-matter column of the transfer table (camb.hpp:181-185): the
// covectors of the modifications carry this spectrum
template <typename T> void build_sqrtp_dm(gic_context *c, Level &L) {
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
  long h = L.g.n / 2;
  long nent = 3 * h * h + 1;
  L.sqrtp_dm.ensure(nent * sizeof(T));
  double kw = 2. * M_PI / L.g.box;
  double norm = kw * kw * kw / pow(2.f * M_PI, 3.f) * double(L.g.n3());
  sqrtp_table<T>(c->spline_dm_dev, c->amplitude, c->ns_frozen, L.g.kmin(), norm, nent, (T *)L.sqrtp_dm.p, c->st);
}

// The white noise stays in real space after the draw; the first transform
// to Fourier space zeroes Re(w_0) (randomfieldgenerator.hpp:194-195).  When
// the power spectrum is applied to noise that is still real, the r2c and the
// sqrt P multiply share one pass.
template <typename T> void noise_to_fourier(gic_context *c, Level &L, const T *post) {
  if (L.noise_fourier) {
    if (post) mul_table<T>((T *)L.noise.p, L.g.n, post, c->st);
    return;
  }
  fft_r2c_post<T>((T *)L.noise.p, L.g.n, post, c->st, c->hook());
  L.noise_fourier = true;
  if (L.dc_pending) {
    if (!post) fill<T>((T *)L.noise.p, 1, T(0), c->st);   // with sqrt P applied the k=0 mode is already 0
    L.dc_pending = false;
  }
}

template <typename T> void do_draw(gic_context *c) {
  if (!c->seeded) throw std::runtime_error("The random number generator has not been seeded");
  if (c->drawn) throw std::runtime_error("Trying to re-draw the random field after it was already initialised");
  if (c->levels.empty()) throw std::runtime_error("Cannot draw a random field before the base grid is initialised");
  Phase ph(c, "phase:draw");
  for (auto &L : c->levels) alloc_field(c, L->noise, L->g);
  if (c->seed_mode == GIC_SEED_REAL_SPACE) {
    std::vector<RngLevelOut> outs;
    size_t total = 0;
    for (auto &L : c->levels) {
      outs.push_back({L->noise.p, L->g.n, L->g.n3()});
      total += L->g.n3();
    }
    RngPlan plan;
    rng_plan(c->seed, total, plan);
    c->work.ensure(rng_workspace_bytes(plan));
    rng_draw_real<T>(plan, outs, c->work.p, c->work.bytes, c->st, c->hook());
    for (auto &L : c->levels) {
      L->noise_fourier = false;
      L->dc_pending = true;
    }
  } else {
    // Fourier-space draws: the Gaussians of every level in draw order into a
    // dense buffer (one stream continued across levels, or in parallel mode a
    // stream seeded seed+ks per shell and level), then placed mode by mode.
    const double sigma = double(T(1.0 / sqrt(2.0)));
    bool parallel = c->seed_mode == GIC_SEED_FOURIER_PARALLEL;
    int reverse = c->seed_mode == GIC_SEED_FOURIER_REVERSE ? 1 : 0;
    std::vector<size_t> off;
    size_t total = 0, maxslots = 0;
    for (auto &L : c->levels) {
      off.push_back(total);
      total += 2 * size_t(fourier_draw_count(L->g.n));
      maxslots = std::max(maxslots, size_t(L->g.n) * L->g.n * (L->g.n / 2 + 1));
    }
    size_t gbytes = (total * sizeof(T) + 255) / 256 * 256;
    c->stage.ensure(gbytes + maxslots * 4);
    T *gauss = (T *)c->stage.p;
    uint32_t *owner = (uint32_t *)((char *)c->stage.p + gbytes);
    if (!parallel) {
      if (total) {
        RngPlan plan;
        rng_plan(c->seed, total, plan);
        c->work.ensure(rng_workspace_bytes(plan));
        std::vector<RngLevelOut> outs{{gauss, 0, total}};
        rng_draw_real<T>(plan, outs, c->work.p, c->work.bytes, c->st, c->hook(), sigma);
      }
    } else {
      int hmax = 0;
      for (auto &L : c->levels) hmax = std::max(hmax, L->g.n / 2);
      if (hmax > 1) {
        RngPlan big;
        rng_plan(c->seed, 24 * size_t(hmax - 1) * (hmax - 1), big);
        c->work.ensure(rng_workspace_bytes(big));
      }
      for (size_t l = 0; l < c->levels.size(); ++l) {
        int h = c->levels[l]->g.n / 2;
        size_t pos = off[l];
        for (int ks = 1; ks < h; ++ks) {
          size_t cnt = 24 * size_t(ks) * ks;
          RngPlan plan;
          rng_plan(c->seed + ks, cnt, plan);
          c->work.ensure(rng_workspace_bytes(plan));
          std::vector<RngLevelOut> outs{{gauss + pos, 0, cnt}};
          rng_draw_real<T>(plan, outs, c->work.p, c->work.bytes, c->st, c->hook(), sigma);
          pos += cnt;
        }
      }
    }
    for (size_t l = 0; l < c->levels.size(); ++l) {
      Level &L = *c->levels[l];
      fourier_draw_scatter<T>(gauss + off[l], L.g.n, (T *)L.noise.p, owner, reverse, c->st);
      L.noise_fourier = true;
      L.dc_pending = false;
    }
  }
  c->drawn = true;
  if (c->fix_power) {
    // enforceUnitVariance (multilevelfield.hpp:395-407), right after the draw (ic.hpp:1343-1344)
    for (auto &Lp : c->levels) {
      ensure_noise_fourier<T>(c, *Lp);
      unit_modulus<T>((T *)Lp->noise.p, Lp->g.n, c->st);
    }
  }
}

template <typename T> void ensure_noise_fourier(gic_context *c, Level &L) { noise_to_fourier<T>(c, L, nullptr); }

template <typename T> void do_apply_ps(gic_context *c) {
  if (!c->drawn) do_draw<T>(c);
  for (auto &Lp : c->levels) {
    Level &L = *Lp;
    build_sqrtp<T>(c, L);
    if (c->hook()) c->hook()->begin("apply_sqrtp", c->st);
    noise_to_fourier<T>(c, L, (const T *)L.sqrtp.p);
    if (c->hook()) c->hook()->end("apply_sqrtp", c->st);
  }
  c->power_applied = true;
}

double kcut_for_level(gic_context *c, int level) {
  const LevelGeom &g = c->levels[level]->g;
  return 0.3 * (double(g.n) * g.kmin());   // FRACTIONAL_K_SPLIT * k_pixel (filterfamily.hpp:110-118)
}

InterpGeom interp_geom(const LevelGeom &gf, const LevelGeom &gc) {
  InterpGeom ig;
  int ratio = (int)llround(gc.dx / gf.dx);
  double dxs = gc.dx / ratio;
  ig.off[0] = (int)llround((gf.ox - gc.ox) / dxs);
  ig.off[1] = (int)llround((gf.oy - gc.oy) / dxs);
  ig.off[2] = (int)llround((gf.oz - gc.oz) / dxs);
  ig.sim = (int)llround(gf.periodic / dxs);
  ig.nsup = gc.n * ratio;
  ig.dxs = dxs;
  ig.dxc = gc.dx;
  ig.co[0] = gc.ox;
  ig.co[1] = gc.oy;
  ig.co[2] = gc.oz;
  ig.periodic = gc.periodic;
  ig.nc = gc.n;
  ig.csim = (int)llround(gc.periodic / gc.dx);
  ig.wrap = gc.n == ig.csim ? 1 : 0;
  return ig;
}

// Low-passed real-space copy of a coarse level for the interpolation onto the
// next level (addFieldFromDifferentGridWithFilter, FILTER_ON_COARSE_GRID
// branch, field.hpp:816-830): copy the field, filter in Fourier space,
// transform back.  For a zoom level the coarse fields are combined real
// fields; the copy is transformed forward with the filter fused into its
// last pass.
template <typename T> CoarseView lowpass_full(gic_context *c, const DevBuf &src, const LevelGeom &g, const T *lp) {
  size_t bytes = g.padded_reals() * sizeof(T);
  c->stage.ensure(bytes);
  GIC_HIP(hipMemcpyAsync(c->stage.p, src.p, bytes, hipMemcpyDeviceToDevice, c->st));
  fft_r2c_post<T>((T *)c->stage.p, g.n, lp, c->st, c->hook());
  fft3d<T>((T *)c->stage.p, g.n, +1, c->st, c->hook());
  CoarseView v;
  v.p = c->stage.p;
  for (int a = 0; a < 3; ++a) {
    v.org[a] = 0;
    v.len[a] = g.n;
  }
  v.sx = long(g.n) * (g.n + 2);
  v.sy = g.n + 2;
  return v;
}

// Coarse cells the tricubic stencils of a fine level touch: the fine grid
// spans coarse cells [o, o + nf/r); the stencil reaches 2 cells below and
// 2 above, plus one cell of slack for rounding in floor().
void coarse_window(const LevelGeom &gf, const LevelGeom &gc, int *w0, int *wl) {
  int r = (int)llround(gc.dx / gf.dx);
  double off[3] = {gf.ox - gc.ox, gf.oy - gc.oy, gf.oz - gc.oz};
  for (int a = 0; a < 3; ++a) {
    int o = (int)llround(off[a] / gc.dx);
    int start = o - 3;
    int len = gf.n / r + 6;
    start %= gc.n;
    if (start < 0) start += gc.n;
    w0[a] = start;
    wl[a] = len;
  }
}

// Second-order LPT (no counterpart in the reference, which stops at
// Zel'dovich, zeldovich.hpp:152-253).  With phi_ij = k_i k_j delta / k^2 the
// source is S = sum_{i<j} (phi_ii phi_jj - phi_ij^2) and
// psi2 = (3/7) i k S / k^2, i.e. D2 = -3/7 D1^2 with delta at the output
// redshift; the velocity uses f2 = 2 f1.  Each level is built from its
// combined overdensity; a zoom level keeps the high-pass part of its own
// psi2 and takes the low-pass part from the level above, as for psi.
template <typename T> void do_second_order(gic_context *c) {
  size_t nl = c->levels.size();
  CoarseView win_view[3];
  bool win_ready = false;
  for (size_t l = 0; l < nl; ++l) {
    Level &L = *c->levels[l];
    const LevelGeom &g = L.g;
    int n = g.n;
    size_t reals = g.padded_reals(), bytes = reals * sizeof(T);
    for (auto &p : L.psi2) alloc_field(c, p, g);
    c->stage.ensure(2 * bytes);
    T *D = (T *)c->stage.p;
    T *S = D + reals;
    const T *spec = (const T *)L.noise.p;
    if (!L.noise_fourier) {
      GIC_HIP(hipMemcpyAsync(D, L.noise.p, bytes, hipMemcpyDeviceToDevice, c->st));
      fft3d<T>(D, n, -1, c->st, c->hook());
      spec = D;
    }
    long h = n / 2, nent = 3 * h * h + 1;
    c->tab_a.ensure(nent * sizeof(T));
    c->tab_c.ensure(nent * sizeof(T));
    filter_table<T>(2, 0.0, g.kmin(), nent, (T *)c->tab_a.p, c->st, true);
    T *A = (T *)L.psi2[0].p, *B = (T *)L.psi2[1].p;
    // the x passes of (1/k^2) delta, i k_x (1/k^2) delta and k_x^2 (1/k^2)
    // delta serve all six phi_ab (their y passes multiply by the k_y, k_z
    // factors); the base level's first two are left by the Zel'dovich pass
    c->xhalf.ensure(3 * bytes);
    T *X0 = (T *)c->xhalf.p, *X1 = X0 + reals, *X2 = X1 + reals;
    if (c->hook()) c->hook()->begin("lpt_source", c->st);
    if (!(l == 0 && c->xhalf_base && spec == (const T *)L.noise.p)) {
      fft_c2r_xpass<T>(spec, X0, n, FFT_PRE_TABLE, nullptr, 0, g.kmin(), c->st, c->hook(), 1.0);
      fft_c2r_xpass<T>(spec, X1, n, FFT_PRE_ZA, nullptr, 0, g.kmin(), c->st, c->hook(), 1.0);
    }
    c->xhalf_base = false;
    fft_c2r_xpass<T>(spec, X2, n, FFT_PRE_KK, nullptr, 0, g.kmin(), c->st, c->hook(), 1.0);
    // each phi_ab but xx is folded into S by its transform's last pass
    auto fold = [&](const T *X, int yop, int epi) {
      fft_c2r_yz_lpt<T>(X, B, n, yop, g.kmin(), epi, S, A, c->st, c->hook());
    };
    fft_c2r_yz<T>(X2, A, n, FFT_YZ_ONE, g.kmin(), c->st, c->hook());   // A = xx
    fold(X0, FFT_YZ_KYKY, 1);   // S = xx yy, A = xx + yy
    fold(X0, FFT_YZ_KZKZ, 2);   // S += (xx + yy) zz
    fold(X1, FFT_YZ_MIKY, 3);   // S -= xy^2
    fold(X1, FFT_YZ_MIKZ, 3);   // S -= xz^2
    fold(X0, FFT_YZ_KYKZ, 3);   // S -= yz^2
    if (c->hook()) c->hook()->end("lpt_source", c->st);
    fft3d<T>(S, n, -1, c->st, c->hook());
    double kc = l >= 1 ? kcut_for_level(c, int(l) - 1) : 0.0;
    filter_table<T>(l >= 1 ? 1 : 2, kc, g.kmin(), nent, (T *)c->tab_c.p, c->st, true, 3.0 / 7.0);
    {
      // psi2_x from the x pass of i k_x f S, psi2_y and psi2_z from that of
      // f S, f = the level's filter (3/7)/k^2 (unit filter on the base level,
      // formed inside the transform)
      const T *tab = l >= 1 ? (const T *)c->tab_c.p : nullptr;
      double ik2 = l >= 1 ? 0.0 : 3.0 / 7.0;
      fft_c2r_xpass<T>(S, X1, n, FFT_PRE_ZA, tab, 0, g.kmin(), c->st, c->hook(), ik2);
      fft_c2r_yz<T>(X1, (T *)L.psi2[0].p, n, FFT_YZ_ONE, g.kmin(), c->st, c->hook());
      fft_c2r_xpass<T>(S, X0, n, FFT_PRE_TABLE, tab, 0, g.kmin(), c->st, c->hook(), ik2);
      fft_c2r_yz<T>(X0, (T *)L.psi2[1].p, n, FFT_YZ_IKY, g.kmin(), c->st, c->hook());
      fft_c2r_yz<T>(X0, (T *)L.psi2[2].p, n, FFT_YZ_IKZ, g.kmin(), c->st, c->hook());
    }
    if (l >= 1) {
      Level &C = *c->levels[l - 1];
      InterpGeom ig = interp_geom(g, C.g);
      if (l == 1 && win_ready) {
        for (int a = 0; a < 3; ++a) interp_add<T>((T *)L.psi2[a].p, n, win_view[a], ig, c->st);
      } else {
        long hc = C.g.n / 2, nc_ent = 3 * hc * hc + 1;
        c->tab_b.ensure(nc_ent * sizeof(T));
        filter_table<T>(0, kcut_for_level(c, int(l) - 1), C.g.kmin(), nc_ent, (T *)c->tab_b.p, c->st);
        for (int a = 0; a < 3; ++a) {
          CoarseView v = lowpass_full<T>(c, C.psi2[a], C.g, (const T *)c->tab_b.p);
          interp_add<T>((T *)L.psi2[a].p, n, v, ig, c->st);
        }
      }
    }
    if (l == 0 && nl >= 2 && g.covers_full()) {
      // the low-pass part that level 1 takes from the base level, straight
      // from the base source spectrum into the coarse window (pruned c2r)
      const LevelGeom &gf = c->levels[1]->g;
      int w0[3], wl[3];
      coarse_window(gf, g, w0, wl);
      if (wl[0] < n && wl[1] < n && wl[2] < n) {
        int nz = n / 2 + 1;
        size_t a_b = size_t(wl[0]) * n * nz * 2 * sizeof(T), b_b = size_t(wl[0]) * wl[1] * nz * 2 * sizeof(T);
        size_t o_b = size_t(wl[0]) * wl[1] * wl[2] * sizeof(T);
        c->lpt_win.ensure(a_b + b_b + 3 * o_b);
        T *bufA = (T *)c->lpt_win.p, *bufB = (T *)((char *)c->lpt_win.p + a_b);
        filter_table<T>(0, kcut_for_level(c, 0), g.kmin(), nent, (T *)c->tab_c.p, c->st, true, 3.0 / 7.0);
        FilterInline fil;
        fil.nf = 1;
        fil.kind[0] = 0;
        fil.kc[0] = kcut_for_level(c, 0);
        fil.scale = 3.0 / 7.0;
        fil.over_k2 = 1;
        for (int a = 0; a < 3; ++a) {
          // x: the x pass of i k_x f S; y, z: one x pass of f S, the y pass
          // multiplying by i k_y or i k_z
          T *out = (T *)((char *)c->lpt_win.p + a_b + b_b + a * o_b);
          fft_c2r_window<T>(S, n, a == 0 ? FFT_PRE_ZA : FFT_PRE_TABLE, (const T *)c->tab_c.p, 0, g.kmin(), w0, wl,
                            bufA, bufB, out, c->st, c->hook(), &fil,
                            a == 0 ? FFT_YZ_ONE : a == 1 ? FFT_YZ_IKY : FFT_YZ_IKZ, a != 2);
          win_view[a].p = out;
          for (int d = 0; d < 3; ++d) {
            win_view[a].org[d] = w0[d];
            win_view[a].len[d] = wl[d];
          }
          win_view[a].sx = long(wl[1]) * wl[2];
          win_view[a].sy = wl[2];
        }
        win_ready = true;
      }
    }
    L.have_psi2 = true;
  }
}

template <typename T> void do_zeldovich(gic_context *c) {
  if (!c->power_applied) do_apply_ps<T>(c);
  Phase ph(c, "phase:zeldovich");
  size_t nl = c->levels.size();
  // Zel'dovich per level, formed inside the first inverse pass from the
  // overdensity spectrum; zoom levels carry their high-pass filter (the cut
  // of the level above) on delta and psi.  The levels are independent here;
  // the base level goes last so that its x passes stay in xhalf for 2LPT.
  c->xhalf_base = false;
  for (size_t li = nl; li-- > 0;) {
    size_t l = li;
    Level &L = *c->levels[l];
    ensure_noise_fourier<T>(c, L);
    for (auto &p : L.psi) alloc_field(c, p, L.g);
    long h = L.g.n / 2, nent = 3 * h * h + 1;
    c->tab_a.ensure(nent * sizeof(T));
    c->tab_c.ensure(nent * sizeof(T));
    double kc = l >= 1 ? kcut_for_level(c, l - 1) : 0.0;
    filter_table<T>(l >= 1 ? 1 : 2, kc, L.g.kmin(), nent, (T *)c->tab_c.p, c->st, true);
    {
      // psi_x from the x pass of i k_x f delta; psi_y and psi_z share the x
      // pass of f delta, their y passes multiplying by i k_y and i k_z
      // (f: the high-pass filter over k^2 on a zoom level, 1/k^2 formed
      // inside the transform on the base level)
      size_t reals = L.g.padded_reals();
      c->xhalf.ensure(3 * reals * sizeof(T));
      T *X0 = (T *)c->xhalf.p, *X1 = X0 + reals;
      const T *tab = l >= 1 ? (const T *)c->tab_c.p : nullptr;
      double ik2 = l >= 1 ? 0.0 : 1.0;
      const T *spec = (const T *)L.noise.p;
      int n = L.g.n;
      fft_c2r_xpass<T>(spec, X1, n, FFT_PRE_ZA, tab, 0, L.g.kmin(), c->st, c->hook(), ik2);
      fft_c2r_yz<T>(X1, (T *)L.psi[0].p, n, FFT_YZ_ONE, L.g.kmin(), c->st, c->hook());
      fft_c2r_xpass<T>(spec, X0, n, FFT_PRE_TABLE, tab, 0, L.g.kmin(), c->st, c->hook(), ik2);
      fft_c2r_yz<T>(X0, (T *)L.psi[1].p, n, FFT_YZ_IKY, L.g.kmin(), c->st, c->hook());
      fft_c2r_yz<T>(X0, (T *)L.psi[2].p, n, FFT_YZ_IKZ, L.g.kmin(), c->st, c->hook());
      // the base level's X0 (1/k^2 delta) and X1 (i k_x / k^2 delta) are the
      // first two x passes of the second derivatives
      c->xhalf_base = l == 0 && c->lpt_order == 2;
    }
    if (l >= 1) {
      filter_table<T>(1, kc, L.g.kmin(), nent, (T *)c->tab_a.p, c->st);
      mul_table<T>((T *)L.noise.p, L.g.n, (const T *)c->tab_a.p, c->st);
      to_real<T>(c, L.noise, L.g.n, L.noise_fourier);
    }
    L.psi_fourier = false;
    L.have_psi = true;
  }
  // fine <- HP(fine) + interpolate(LP(coarse)), coarse already combined
  for (size_t l = 1; l < nl; ++l) {
    Level &F = *c->levels[l];
    Level &C = *c->levels[l - 1];
    long h = C.g.n / 2, nent = 3 * h * h + 1;
    double kc = kcut_for_level(c, l - 1);
    c->tab_b.ensure(nent * sizeof(T));
    c->tab_c.ensure(nent * sizeof(T));
    filter_table<T>(0, kc, C.g.kmin(), nent, (T *)c->tab_b.p, c->st);
    filter_table<T>(0, kc, C.g.kmin(), nent, (T *)c->tab_c.p, c->st, true);
    const T *lp = (const T *)c->tab_b.p;
    const T *lp_k2 = (const T *)c->tab_c.p;
    InterpGeom ig = interp_geom(F.g, C.g);
    int w0[3], wl[3];
    coarse_window(F.g, C.g, w0, wl);
    bool pruned = l == 1 && C.noise_fourier && C.g.covers_full() && wl[0] < C.g.n && wl[1] < C.g.n && wl[2] < C.g.n;
    DevBuf *targets[4] = {&F.noise, &F.psi[0], &F.psi[1], &F.psi[2]};
    for (int f = 0; f < 4; ++f) {
      CoarseView v;
      if (pruned) {
        // the base level's delta spectrum is intact: low-pass it (and form
        // psi from it) straight into the coarse window only
        int nz = C.g.n / 2 + 1;
        size_t a_bytes = size_t(wl[0]) * C.g.n * nz * 2 * sizeof(T);
        size_t b_bytes = size_t(wl[0]) * wl[1] * nz * 2 * sizeof(T);
        size_t o_bytes = size_t(wl[0]) * wl[1] * wl[2] * sizeof(T);
        c->stage.ensure(a_bytes + b_bytes + o_bytes);
        T *bufA = (T *)c->stage.p;
        T *bufB = (T *)((char *)c->stage.p + a_bytes);
        T *out = (T *)((char *)c->stage.p + a_bytes + b_bytes);
        FilterInline fil;   // the same low pass (over k^2 for psi), formed in the pass
        fil.nf = 1;
        fil.kind[0] = 0;
        fil.kc[0] = kc;
        fil.over_k2 = f == 0 ? 0 : 1;
        // delta: the low pass; psi_x: the x pass of i k_x LP/k^2 delta;
        // psi_y, psi_z: one x pass of LP/k^2 delta, the y pass multiplying
        // by i k_y or i k_z
        fft_c2r_window<T>((const T *)C.noise.p, C.g.n, f == 1 ? FFT_PRE_ZA : FFT_PRE_TABLE, f == 0 ? lp : lp_k2, 0,
                          C.g.kmin(), w0, wl, bufA, bufB, out, c->st, c->hook(), &fil,
                          f <= 1 ? FFT_YZ_ONE : f == 2 ? FFT_YZ_IKY : FFT_YZ_IKZ, f != 3);
        v.p = out;
        for (int a = 0; a < 3; ++a) {
          v.org[a] = w0[a];
          v.len[a] = wl[a];
        }
        v.sx = long(wl[1]) * wl[2];
        v.sy = wl[2];
      } else {
        const DevBuf &src = f == 0 ? C.noise : C.psi[f - 1];
        if (f == 0) to_real<T>(c, C.noise, C.g.n, C.noise_fourier);
        v = lowpass_full<T>(c, src, C.g, lp);
      }
      if (f == 0) to_real<T>(c, F.noise, F.g.n, F.noise_fourier);
      {
        // the fine field read and written, the coarse window read once
        double coarse = double(v.len[0]) * v.len[1] * v.len[2];
        Timed t(c, "interp_add", (2.0 * F.g.n3() + coarse) * sizeof(T));
        interp_add<T>((T *)targets[f]->p, F.g.n, v, ig, c->st);
      }
    }
  }
  if (c->lpt_order == 2) {
    Phase ph2(c, "phase:2lpt");
    do_second_order<T>(c);
  }
  c->zeldovich_done = true;
}

template <typename T> void do_copy_field(gic_context *c, int field, int level, int space, void *host, int dtype) {
  Level &L = level_at(c, level);
  DevBuf *b;
  bool *flag;
  if (field == GIC_FIELD_OVERDENSITY) {
    if (!c->drawn) do_draw<T>(c);
    b = &L.noise;
    flag = &L.noise_fourier;
  } else if (field >= 1 && field <= 3) {
    if (!L.have_psi) throw std::runtime_error("Attempt to generate particles before they have been calculated");
    b = &L.psi[field - 1];
    flag = &L.psi_fourier;
  } else if (field >= 4 && field <= 6) {
    if (!L.have_psi2) throw std::runtime_error("Second-order displacements have not been calculated (lpt_order 2)");
    if (space != GIC_REAL_SPACE) throw std::runtime_error("second-order displacements are held in real space");
    b = &L.psi2[field - 4];
    static bool real_flag;
    real_flag = false;
    flag = &real_flag;
  } else {
    throw std::runtime_error("unknown field");
  }
  int n = L.g.n;
  if (space == GIC_REAL_SPACE) {
    if (field == GIC_FIELD_OVERDENSITY) {
      if (L.dc_pending) ensure_noise_fourier<T>(c, L);
      to_real<T>(c, *b, n, *flag);
    }
    else if (*flag) {
      for (int a = 0; a < 3; ++a) {
        bool f = true;
        to_real<T>(c, L.psi[a], n, f);
      }
      *flag = false;
    }
    size_t n3 = L.g.n3();
    size_t ob = n3 * (dtype == GIC_DTYPE_F32 ? 4 : 8);
    c->stage.ensure(ob);
    if (dtype == GIC_DTYPE_F32) unpad<T, float>((const T *)b->p, n, (float *)c->stage.p, c->st);
    else unpad<T, double>((const T *)b->p, n, (double *)c->stage.p, c->st);
    GIC_HIP(hipMemcpyAsync(host, c->stage.p, ob, hipMemcpyDeviceToHost, c->st));
  } else {
    if (field == GIC_FIELD_OVERDENSITY) ensure_noise_fourier<T>(c, L);
    else if (!*flag) {
      for (int a = 0; a < 3; ++a) {
        bool f = false;
        to_fourier<T>(c, L.psi[a], n, f);
      }
      *flag = true;
    }
    size_t nreal = L.g.padded_reals();
    size_t ob = nreal * (dtype == GIC_DTYPE_F32 ? 4 : 8);
    c->stage.ensure(ob);
    if (dtype == GIC_DTYPE_F32) convert<T, float>((const T *)b->p, nreal, (float *)c->stage.p, c->st);
    else convert<T, double>((const T *)b->p, nreal, (double *)c->stage.p, c->st);
    GIC_HIP(hipMemcpyAsync(host, c->stage.p, ob, hipMemcpyDeviceToHost, c->st));
  }
  GIC_HIP(hipStreamSynchronize(c->st));
}

template <typename T> void do_set_field(gic_context *c, int field, int level, int space, const void *host, int dtype) {
  Level &L = level_at(c, level);
  if (field != GIC_FIELD_OVERDENSITY) throw std::runtime_error("only the overdensity / noise field can be set");
  alloc_field(c, L.noise, L.g);
  int n = L.g.n;
  size_t es = dtype == GIC_DTYPE_F32 ? 4 : 8;
  if (space == GIC_REAL_SPACE) {
    size_t ob = L.g.n3() * es;
    c->stage.ensure(ob);
    GIC_HIP(hipMemcpyAsync(c->stage.p, host, ob, hipMemcpyHostToDevice, c->st));
    if (dtype == GIC_DTYPE_F32) pad<float, T>((const float *)c->stage.p, n, (T *)L.noise.p, c->st);
    else pad<double, T>((const double *)c->stage.p, n, (T *)L.noise.p, c->st);
    L.noise_fourier = false;
    L.dc_pending = false;
  } else {
    size_t nreal = L.g.padded_reals();
    c->stage.ensure(nreal * es);
    GIC_HIP(hipMemcpyAsync(c->stage.p, host, nreal * es, hipMemcpyHostToDevice, c->st));
    if (dtype == GIC_DTYPE_F32) convert<float, T>((const float *)c->stage.p, nreal, (T *)L.noise.p, c->st);
    else convert<double, T>((const double *)c->stage.p, nreal, (T *)L.noise.p, c->st);
    L.noise_fourier = true;
    L.dc_pending = false;
  }
  GIC_HIP(hipStreamSynchronize(c->st));
}

// Per-|k|^2 histogram of the overdensity's power and the theory covariance
// P(k) N^3/V for dumpPowerSpectrum (cosmology/parameters.hpp:112-205).  The
// overdensity is transformed on a copy, so the field keeps its space.
template <typename T>
void do_mode_power(gic_context *c, int level, double *sum, double *cnt, double *theory, size_t nm) {
  Level &L = level_at(c, level);
  long h = L.g.n / 2;
  if (nm != size_t(3 * h * h + 1)) throw std::runtime_error("histogram length must be 3(n/2)^2+1");
  if (!c->drawn) do_draw<T>(c);
  if (!c->power_applied) do_apply_ps<T>(c);
  if (L.dc_pending) ensure_noise_fourier<T>(c, L);
  size_t bytes = L.g.padded_reals() * sizeof(T);
  size_t hb = nm * sizeof(double);
  c->stage.ensure(bytes + 2 * hb);
  T *spec = (T *)L.noise.p;
  if (!L.noise_fourier) {
    GIC_HIP(hipMemcpyAsync(c->stage.p, L.noise.p, bytes, hipMemcpyDeviceToDevice, c->st));
    spec = (T *)c->stage.p;
    fft3d<T>(spec, L.g.n, -1, c->st, c->hook());
  }
  double *dsum = (double *)((char *)c->stage.p + bytes);
  double *dcnt = dsum + nm;
  fill<double>(dsum, 2 * nm, 0.0, c->st);
  mode_power<T>(spec, L.g.n, dsum, dcnt, c->st);
  GIC_HIP(hipMemcpyAsync(sum, dsum, hb, hipMemcpyDeviceToHost, c->st));
  GIC_HIP(hipMemcpyAsync(cnt, dcnt, hb, hipMemcpyDeviceToHost, c->st));
  std::vector<T> tab(nm);
  build_sqrtp<T>(c, L);
  GIC_HIP(hipMemcpyAsync(tab.data(), L.sqrtp.p, nm * sizeof(T), hipMemcpyDeviceToHost, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
  for (size_t m = 0; m < nm; ++m) theory[m] = double(tab[m]) * double(tab[m]);
}

double velocity_ratio(const Cosmo &c) {
  double f = 1.0;
  return f * 100. * sqrt(c.Om / (c.a * c.a * c.a) + c.Ol) * sqrt(c.a);
}

template <typename T>
void do_particles(gic_context *c, int level, const uint64_t *cells, size_t start, size_t count, float *pos,
                  float *vel) {
  Level &L = level_at(c, level);
  if (!L.have_psi) throw std::runtime_error("Attempt to generate particles before they have been calculated");
  if (L.psi_fourier) {
    for (int a = 0; a < 3; ++a) {
      bool f = true;
      to_real<T>(c, L.psi[a], L.g.n, f);
    }
    L.psi_fourier = false;
  }
  size_t n3 = L.g.n3();
  if (cells) {
    for (size_t i = 0; i < count; ++i)
      if (cells[i] >= n3) throw std::out_of_range("particle cell index out of range for the level");
  } else if (start > n3 || count > n3 - start) {
    throw std::out_of_range("particle range exceeds the level's cells");
  }
  ParticleGeom pg;
  pg.dx = L.g.dx;
  pg.o[0] = L.g.ox;
  pg.o[1] = L.g.oy;
  pg.o[2] = L.g.oz;
  pg.periodic = L.g.periodic;
  pg.vratio = velocity_ratio(c->cosmo);
  pg.vratio2 = 2.0 * pg.vratio;   // f2 = 2 f1 (f1 = 1, as zeldovichVelocityToOffsetRatio)
  // device staging: cell list + outputs, processed in slices
  const size_t slice = size_t(1) << 24;
  for (size_t s0 = 0; s0 < count; s0 += slice) {
    size_t m = std::min(slice, count - s0);
    size_t need = m * 24 + (cells ? m * 8 : 0);
    c->stage.ensure(need);
    float *dpos = (float *)c->stage.p;
    float *dvel = dpos + 3 * m;
    uint64_t *dcells = nullptr;
    if (cells) {
      dcells = (uint64_t *)(dvel + 3 * m);
      GIC_HIP(hipMemcpyAsync(dcells, cells + s0, m * 8, hipMemcpyHostToDevice, c->st));
    }
    const T *q2[3] = {(const T *)L.psi2[0].p, (const T *)L.psi2[1].p, (const T *)L.psi2[2].p};
    particles<T>((const T *)L.psi[0].p, (const T *)L.psi[1].p, (const T *)L.psi[2].p, L.have_psi2 ? q2 : nullptr,
                 L.g.n, pg, dcells, start + s0, m,
                 dpos, dvel, c->st);
    GIC_HIP(hipMemcpyAsync(pos + 3 * s0, dpos, m * 12, hipMemcpyDeviceToHost, c->st));
    GIC_HIP(hipMemcpyAsync(vel + 3 * s0, dvel, m * 12, hipMemcpyDeviceToHost, c->st));
    GIC_HIP(hipStreamSynchronize(c->st));
  }
}

template <typename F> int guarded(gic_context *c, F &&f) {
  if (!c) return 2;
  try {
    GIC_HIP(hipSetDevice(c->device));
    f();
    return 0;
  } catch (std::exception &e) {
    c->err = e.what();
    return 1;
  }
}

#define DISPATCH(c, fn, ...)                  \
  do {                                        \
    if ((c)->prec == 32) fn<float>(__VA_ARGS__); \
    else fn<double>(__VA_ARGS__);             \
  } while (0)

template void build_sqrtp<float>(gic_context *, Level &);
template void build_sqrtp<double>(gic_context *, Level &);
template void build_sqrtp_dm<float>(gic_context *, Level &);
template void build_sqrtp_dm<double>(gic_context *, Level &);
template void do_draw<float>(gic_context *);
template void do_draw<double>(gic_context *);
template void ensure_noise_fourier<float>(gic_context *, Level &);
template void ensure_noise_fourier<double>(gic_context *, Level &);
template void do_zeldovich<float>(gic_context *);
template void do_zeldovich<double>(gic_context *);

template <typename T> void do_import_level(gic_context *c, int level, const double *host) {
  if (!c->drawn) do_draw<T>(c);
  if (c->power_applied) throw std::runtime_error("import_level acts on the white noise, before the output is generated");
  Level &L = level_at(c, level);
  int n = L.g.n;
  size_t n3 = L.g.n3();
  c->stage.ensure(n3 * sizeof(double));
  GIC_HIP(hipMemcpyAsync(c->stage.p, host, n3 * sizeof(double), hipMemcpyHostToDevice, c->st));
  pad<double, T>((const double *)c->stage.p, n, (T *)L.noise.p, c->st);
  // the file holds a dark-matter overdensity: back to white noise with P_dm^-1/2
  build_sqrtp_dm<T>(c, L);
  long h = n / 2, nent = 3 * h * h + 1;
  c->tab_a.ensure(nent * sizeof(T));
  inverse_table<T>((const T *)L.sqrtp_dm.p, nent, (T *)c->tab_a.p, c->st);
  fft_r2c_post<T>((T *)L.noise.p, n, (const T *)c->tab_a.p, c->st, c->hook());
  L.noise_fourier = true;
  L.dc_pending = false;
  GIC_HIP(hipStreamSynchronize(c->st));
}

}  // namespace gic

extern "C" {

int gic_create(gic_context **out, int device, int precision) {
  if (!out) return 2;
  *out = nullptr;
  if (precision != 32 && precision != 64) return 3;
  auto *c = new gic_context();
  c->device = device;
  c->prec = precision;
  try {
    int ndev = 0;
    GIC_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw std::runtime_error("no such HIP device");
    GIC_HIP(hipSetDevice(device));
    GIC_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    rng_scratch_setup(device);   // the fixed workspace of gic_gaussian_stream_noalloc on this device
  } catch (std::exception &e) {
    delete c;
    return 4;
  }
  *out = c;
  return 0;
}

void gic_destroy(gic_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  c->timer.reset();
  c->levels.clear();
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char *gic_last_error(const gic_context *c) { return c ? c->err.c_str() : "null context"; }

int gic_set_cosmology(gic_context *c, double Om, double Ol, double Ob, double hubble, double s8, double ns, double z) {
  return guarded(c, [&] {
    c->cosmo.Om = Om;
    c->cosmo.Ol = Ol;
    c->cosmo.Ob = Ob;
    c->cosmo.h = hubble;
    c->cosmo.s8 = s8;
    c->cosmo.ns = ns;
    c->cosmo.z = z;
    c->cosmo.a = 1. / (z + 1.);
  });
}

int gic_set_power_law(gic_context *c, double amplitude) {
  return guarded(c, [&] {
    if (!(amplitude > 0) || !std::isfinite(amplitude)) throw std::runtime_error("powerlaw_amplitude must be positive");
    c->kpts.clear();
    c->ns_frozen = c->cosmo.ns;
    c->amplitude = amplitude;
    c->spline_dev = SplineHost{};       // n = 0: unit transfer function for every species
    c->spline_dm_dev = SplineHost{};
    c->have_camb = true;
    for (auto &Lp : c->levels) {
      Lp->sqrtp.release();
      Lp->sqrtp_dm.release();
    }
  });
}

int gic_set_camb_table(gic_context *c, const double *v, size_t nv, int ncols) {
  return guarded(c, [&] {
    if (ncols != 7 && ncols != 13) throw std::runtime_error("CAMB transfer file doesn't have a sensible number of columns");
    if (nv < size_t(ncols) * 2) throw std::runtime_error("CAMB transfer file is too short");
    double tnorm = v[1];
    std::vector<double> kp, tall;
    for (size_t j = 0; j < nv / ncols; ++j) {
      if (v[ncols * j] > 0) {
        kp.push_back(v[ncols * j]);
        tall.push_back(v[ncols * j + 6] / tnorm);
      }
    }
    std::vector<double> tdm;
    for (size_t j = 0; j < nv / ncols; ++j)
      if (v[ncols * j] > 0) tdm.push_back(v[ncols * j + 1] / tnorm);
    c->kpts = kp;
    c->tf_all = LogSpline();
    c->tf_all.init(kp, tall);
    c->tf_dm = LogSpline();
    c->tf_dm.init(kp, tdm);
    c->ns_frozen = c->cosmo.ns;
    double gf = growth_factor(c->cosmo.a, c->cosmo.Om, c->cosmo.Ol) / growth_factor(1.0, c->cosmo.Om, c->cosmo.Ol);
    double s8pre = variance_in_sphere(c, 8.);
    double lr = (c->cosmo.s8 / s8pre) * gf;
    c->amplitude = lr * lr;
    const Spline &s = c->tf_all.s;
    int n = (int)s.x.size();
    c->spl_x.ensure(n * 8);
    c->spl_y.ensure(n * 8);
    c->spl_c.ensure(n * 8);
    GIC_HIP(hipMemcpy(c->spl_x.p, s.x.data(), n * 8, hipMemcpyHostToDevice));
    GIC_HIP(hipMemcpy(c->spl_y.p, s.y.data(), n * 8, hipMemcpyHostToDevice));
    GIC_HIP(hipMemcpy(c->spl_c.p, s.c.data(), n * 8, hipMemcpyHostToDevice));
    c->spline_dev = SplineHost{(double *)c->spl_x.p, (double *)c->spl_y.p, (double *)c->spl_c.p, n, c->tf_all.minx,
                               c->tf_all.maxx};
    const Spline &sd = c->tf_dm.s;
    int nd = (int)sd.x.size();
    c->spl_dm_x.ensure(nd * 8);
    c->spl_dm_y.ensure(nd * 8);
    c->spl_dm_c.ensure(nd * 8);
    GIC_HIP(hipMemcpy(c->spl_dm_x.p, sd.x.data(), nd * 8, hipMemcpyHostToDevice));
    GIC_HIP(hipMemcpy(c->spl_dm_y.p, sd.y.data(), nd * 8, hipMemcpyHostToDevice));
    GIC_HIP(hipMemcpy(c->spl_dm_c.p, sd.c.data(), nd * 8, hipMemcpyHostToDevice));
    c->spline_dm_dev = SplineHost{(double *)c->spl_dm_x.p, (double *)c->spl_dm_y.p, (double *)c->spl_dm_c.p, nd,
                                  c->tf_dm.minx, c->tf_dm.maxx};
    c->have_camb = true;
  });
}

int gic_init_base_grid(gic_context *c, double box, int n) {
  return guarded(c, [&] {
    if (!c->levels.empty()) throw std::runtime_error("Cannot re-initialize the base grid");
    if (c->drawn) throw std::runtime_error("Trying to initialize a grid after the random field was already drawn");
    if (box <= 0) throw std::runtime_error("Box size must be positive");
    if (n < 4 || (n & (n - 1))) throw std::runtime_error("Grid size must be a power of two >= 4");
    auto L = std::make_unique<Level>();
    L->g.n = n;
    L->g.periodic = box;
    L->g.dx = box / n;
    L->g.box = L->g.dx * n;
    c->levels.push_back(std::move(L));
  });
}

int gic_init_zoom_grid_with_origin_at(gic_context *c, int x0, int y0, int z0, int zoomfac, int n) {
  return guarded(c, [&] {
    if (c->drawn) throw std::runtime_error("Trying to initialize a grid after the random field was already drawn");
    if (c->levels.empty()) throw std::runtime_error("Cannot initialise a zoom grid before initialising the base grid");
    if (c->levels.size() >= 4) throw std::runtime_error("At most 4 levels are supported");
    if (n < 4 || (n & (n - 1))) throw std::runtime_error("Grid size must be a power of two >= 4");
    if (zoomfac < 1) throw std::runtime_error("Zoom factor must be positive");
    const LevelGeom &A = c->levels.back()->g;
    if (x0 < 0 || y0 < 0 || z0 < 0)
      throw std::runtime_error("The newly initialised zoom level is not fully contained within the previous level");
    int nc = A.n / zoomfac;
    if (!A.covers_full() && (x0 + nc > A.n || y0 + nc > A.n || z0 + nc > A.n))
      throw std::runtime_error("The newly initialised zoom level is not fully contained within the previous level");
    auto L = std::make_unique<Level>();
    double size = A.box / zoomfac;
    L->g.n = n;
    L->g.periodic = A.periodic;
    L->g.dx = size / n;
    L->g.box = L->g.dx * n;
    L->g.ox = A.ox + x0 * A.dx;
    L->g.oy = A.oy + y0 * A.dx;
    L->g.oz = A.oz + z0 * A.dx;
    double r = A.dx / L->g.dx;
    if (std::fabs(r - std::round(r)) > 1e-6 || std::round(r) < 1)
      throw std::runtime_error("Zoom cell size must divide the parent cell size");
    c->levels.push_back(std::move(L));
  });
}

int gic_num_levels(gic_context *c, int *out) {
  return guarded(c, [&] { *out = (int)c->levels.size(); });
}

int gic_level_geometry(gic_context *c, int level, int *n, double *box, double *off, double *dx) {
  return guarded(c, [&] {
    const LevelGeom &g = level_at(c, level).g;
    if (n) *n = g.n;
    if (box) *box = g.box;
    if (off) {
      off[0] = g.ox;
      off[1] = g.oy;
      off[2] = g.oz;
    }
    if (dx) *dx = g.dx;
  });
}

int gic_seed(gic_context *c, unsigned long seed, int mode) {
  return guarded(c, [&] {
    if (c->seeded) throw std::runtime_error("The random number generator has already been seeded");
    if (mode < 0 || mode > 3) throw std::runtime_error("unknown seeding mode");
    c->seed = seed;
    c->seed_mode = mode;
    c->seeded = true;
  });
}

int gic_reseed(gic_context *c, unsigned long seed, int mode) {
  return guarded(c, [&] {
    if (mode < 0 || mode > 3) throw std::runtime_error("unknown seeding mode");
    c->seed = seed;
    c->seed_mode = mode;
    c->seeded = true;
    c->drawn = c->power_applied = c->zeldovich_done = false;
    for (auto &L : c->levels) {
      L->noise_fourier = false;
      L->psi_fourier = false;
      L->have_psi = false;
    }
  });
}

int gic_draw(gic_context *c) {
  return guarded(c, [&] { DISPATCH(c, do_draw, c); });
}

int gic_zero_level(gic_context *c, int level) {
  return guarded(c, [&] {
    Level &L = level_at(c, level);
    if (!c->drawn) DISPATCH(c, do_draw, c);
    if (c->prec == 32) fill<float>((float *)L.noise.p, L.g.padded_reals(), 0.f, c->st);
    else fill<double>((double *)L.noise.p, L.g.padded_reals(), 0.0, c->st);
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_set_exact_power_spectrum(gic_context *c, int enable) {
  return guarded(c, [&] {
    if (c->drawn) throw std::runtime_error("fix_power must be set before the random field is drawn");
    c->fix_power = enable != 0;
  });
}

int gic_reverse(gic_context *c) {
  return guarded(c, [&] {
    if (!c->drawn) DISPATCH(c, do_draw, c);
    if (c->power_applied) throw std::runtime_error("reverse acts on the white noise, before the output is generated");
    for (auto &Lp : c->levels) {
      if (c->prec == 32) scale_all<float>((float *)Lp->noise.p, Lp->g.padded_reals(), -1.f, c->st);
      else scale_all<double>((double *)Lp->noise.p, Lp->g.padded_reals(), -1.0, c->st);
    }
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_reverse_small_k(gic_context *c, double kmax) {
  return guarded(c, [&] {
    if (!c->drawn) DISPATCH(c, do_draw, c);
    if (c->power_applied) throw std::runtime_error("reverse_small_k acts on the white noise, before the output is generated");
    for (auto &Lp : c->levels) {
      Level &L = *Lp;
      if (c->prec == 32) {
        ensure_noise_fourier<float>(c, L);
        reverse_small_k<float>((float *)L.noise.p, L.g.n, L.g.kmin(), kmax, c->st);
      } else {
        ensure_noise_fourier<double>(c, L);
        reverse_small_k<double>((double *)L.noise.p, L.g.n, L.g.kmin(), kmax, c->st);
      }
    }
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_import_level(gic_context *c, int level, const double *host, size_t n) {
  return guarded(c, [&] {
    Level &L = level_at(c, level);
    if (!host || n != L.g.n3()) throw std::runtime_error("import_level: the data must hold n^3 values of the level");
    DISPATCH(c, do_import_level, c, level, host);
  });
}

int gic_splice(gic_context *c, unsigned long new_seed, int *cg_iterations) {
  return guarded(c, [&] { DISPATCH(c, do_splice, c, new_seed, cg_iterations); });
}

int gic_apply_power_spectrum(gic_context *c) {
  return guarded(c, [&] {
    DISPATCH(c, do_apply_ps, c);
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_zeldovich(gic_context *c) {
  return guarded(c, [&] {
    DISPATCH(c, do_zeldovich, c);
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_copy_field(gic_context *c, int field, int level, int space, void *host, int dtype) {
  return guarded(c, [&] { DISPATCH(c, do_copy_field, c, field, level, space, host, dtype); });
}

int gic_set_field(gic_context *c, int field, int level, int space, const void *host, int dtype) {
  return guarded(c, [&] {
    DISPATCH(c, do_set_field, c, field, level, space, host, dtype);
    c->drawn = true;
  });
}

int gic_particles(gic_context *c, int level, const uint64_t *cells, size_t start, size_t count, float *pos,
                  float *vel) {
  return guarded(c, [&] { DISPATCH(c, do_particles, c, level, cells, start, count, pos, vel); });
}

int gic_mode_power(gic_context *c, int level, double *sum_power, double *count, double *theory, size_t nm) {
  return guarded(c, [&] { DISPATCH(c, do_mode_power, c, level, sum_power, count, theory, nm); });
}

int gic_particle_mass_eps(gic_context *c, int level, double *mass, double *eps) {
  return guarded(c, [&] {
    const LevelGeom &g = level_at(c, level).g;
    double boxMass = 27.744948 * c->cosmo.Om * pow(g.periodic, 3.0);
    if (mass) *mass = boxMass * pow(g.dx / g.periodic, 3.0);
    if (eps) *eps = g.dx * 1.0 * 0.01075;
  });
}

int gic_set_lpt_order(gic_context *c, int order) {
  return guarded(c, [&] {
    if (order != 1 && order != 2) throw std::runtime_error("LPT order must be 1 or 2");
    c->lpt_order = order;
  });
}

int gic_set_flags(gic_context *c, int level, const uint64_t *cells, size_t n) {
  return guarded(c, [&] {
    Level &L = level_at(c, level);
    c->flags.resize(c->levels.size());
    std::vector<uint64_t> f(cells, cells + n);
    for (uint64_t v : f)
      if (v >= L.g.n3()) throw std::out_of_range("flagged cell index out of range for the level");
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
    c->flags[level] = f;
  });
}

int gic_set_variance_filter_scale(gic_context *c, double scale) {
  return guarded(c, [&] { c->variance_scale = scale; });
}

int gic_calculate(gic_context *c, const char *name, double *out) {
  return guarded(c, [&] {
    if (c->prec == 32) *out = mod_calculate<float>(c, name);
    else *out = mod_calculate<double>(c, name);
  });
}

int gic_add_modification(gic_context *c, const char *name, int relative, double target, int init_steps,
                         double precision) {
  return guarded(c, [&] {
    if (c->prec == 32) mod_add<float>(c, name, relative != 0, target, init_steps, precision);
    else mod_add<double>(c, name, relative != 0, target, init_steps, precision);
  });
}

int gic_apply_modifications(gic_context *c, double *pre_chi2, double *post_chi2) {
  return guarded(c, [&] {
    Phase ph(c, "phase:modifications");
    DISPATCH(c, mod_apply, c, pre_chi2, post_chi2);
  });
}

int gic_clear_modifications(gic_context *c) {
  return guarded(c, [&] { mod_clear(c); });
}

int gic_chi2(gic_context *c, double *chi2, uint64_t *dof) {
  return guarded(c, [&] {
    if (c->prec == 32) *chi2 = mod_chi2<float>(c);
    else *chi2 = mod_chi2<double>(c);
    if (dof) *dof = num_dof(c);
  });
}

int gic_displacement_forward(gic_context *c, const void *const *noise, void *const *psi) {
  return guarded(c, [&] {
    if (!noise || !psi) throw std::runtime_error("displacement forward: null pointer array");
    if (c->prec == 32) do_displacement_forward<float>(c, (const float *const *)noise, (float *const *)psi);
    else do_displacement_forward<double>(c, (const double *const *)noise, (double *const *)psi);
  });
}

int gic_batch_constrained(gic_context *c, int nreal, const uint64_t *seeds, int ncons, const double *cons,
                          void *out, int lanes, int use_graph, double *elapsed_ms) {
  return guarded(c, [&] {
    if (c->prec == 32) do_batch<float>(c, nreal, seeds, ncons, cons, (float *)out, lanes, use_graph != 0, elapsed_ms);
    else do_batch<double>(c, nreal, seeds, ncons, cons, (double *)out, lanes, use_graph != 0, elapsed_ms);
  });
}

int gic_displacement_adjoint(gic_context *c, const void *const *psi_bar, void *const *noise_bar) {
  return guarded(c, [&] {
    if (!psi_bar || !noise_bar) throw std::runtime_error("displacement adjoint: null pointer array");
    if (c->prec == 32) do_displacement_adjoint<float>(c, (const float *const *)psi_bar, (float *const *)noise_bar);
    else do_displacement_adjoint<double>(c, (const double *const *)psi_bar, (double *const *)noise_bar);
  });
}

int gic_variance_gradient(gic_context *c, const void *const *noise, double *value, void *const *grad) {
  return guarded(c, [&] {
    if (!noise || !value) throw std::runtime_error("variance gradient: null pointer");
    if (c->prec == 32) *value = mod_variance_gradient<float>(c, (const float *const *)noise, (float *const *)grad);
    else *value = mod_variance_gradient<double>(c, (const double *const *)noise, (double *const *)grad);
  });
}

int gic_set_quadratic_line_search(gic_context *c, int enable) {
  return guarded(c, [&] { mod_set_line_search(c, enable != 0); });
}

int gic_velocity_ratio(gic_context *c, double *out) {
  return guarded(c, [&] { *out = velocity_ratio(c->cosmo); });
}

int gic_fft_inplace(gic_context *c, void *dev, int n, int direction) {
  return guarded(c, [&] {
    if (direction != -1 && direction != 1) throw std::runtime_error("direction must be -1 or +1");
    if (c->prec == 32) fft3d<float>((float *)dev, n, direction, c->st, c->hook());
    else fft3d<double>((double *)dev, n, direction, c->st, c->hook());
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_raw_stream(gic_context *c, unsigned long seed, uint32_t *host_out, size_t n) {
  return guarded(c, [&] {
    RngPlan plan;
    rng_plan(seed, n / 2 + 1, plan);
    c->work.ensure(rng_workspace_bytes(plan));
    DevBuf out;
    out.ensure(n * 4);
    rng_raw(plan, (uint32_t *)out.p, n, c->work.p, c->work.bytes, c->st);
    GIC_HIP(hipMemcpyAsync(host_out, out.p, n * 4, hipMemcpyDeviceToHost, c->st));
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_gaussian_stream(gic_context *c, unsigned long seed, double *host_out, size_t n) {
  return guarded(c, [&] {
    if (n == 0) return;
    RngPlan plan;
    rng_plan(seed, n, plan);
    c->work.ensure(rng_workspace_bytes(plan));
    DevBuf dense;
    dense.ensure(n * 8);
    std::vector<RngLevelOut> outs{{dense.p, 0, n}};
    rng_draw_real<double>(plan, outs, c->work.p, c->work.bytes, c->st, c->hook());
    GIC_HIP(hipMemcpyAsync(host_out, dense.p, n * 8, hipMemcpyDeviceToHost, c->st));
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_gaussian_stream_noalloc(gic_context *c, unsigned long seed, double *device_out, size_t n) {
  return guarded(c, [&] {
    if (n == 0) return;
    if (!device_out) throw std::runtime_error("gic_gaussian_stream_noalloc: no output array");
    rng_gaussians_noalloc(c->device, seed, device_out, n, c->st);
    GIC_HIP(hipStreamSynchronize(c->st));
  });
}

int gic_synchronize(gic_context *c) {
  return guarded(c, [&] { GIC_HIP(hipStreamSynchronize(c->st)); });
}

int gic_set_event_timing(gic_context *c, int enable) {
  return guarded(c, [&] {
    if (enable && !c->timer) c->timer = std::make_unique<EventTimer>();
    if (!enable) c->timer.reset();
  });
}

int gic_get_event_timing(gic_context *c, const char *kernel, double *total_ms, long *count) {
  return guarded(c, [&] {
    if (!c->timer) throw std::runtime_error("event timing is not enabled");
    c->timer->collect();
    auto it = c->timer->totals.find(kernel);
    *total_ms = it == c->timer->totals.end() ? 0.0 : it->second.first;
    *count = it == c->timer->totals.end() ? 0 : it->second.second;
  });
}

int gic_reset_event_timing(gic_context *c) {
  return guarded(c, [&] {
    if (c->timer) {
      c->timer->collect();
      c->timer->totals.clear();
      c->timer->algo_bytes.clear();
    }
  });
}

int gic_list_event_timing(gic_context *c, char *buf, size_t buflen, size_t *needed) {
  return guarded(c, [&] {
    if (!c->timer) throw std::runtime_error("event timing is not enabled");
    c->timer->collect();
    std::string s;
    char line[256];
    for (auto &kv : c->timer->totals) {
      auto ab = c->timer->algo_bytes.find(kv.first);
      double b = ab == c->timer->algo_bytes.end() ? 0.0 : ab->second;
      snprintf(line, sizeof line, "%s\t%.6f\t%ld\t%.0f\n", kv.first.c_str(), kv.second.first, kv.second.second, b);
      s += line;
    }
    if (needed) *needed = s.size() + 1;
    if (buf && buflen) {
      size_t m = std::min(buflen - 1, s.size());
      memcpy(buf, s.data(), m);
      buf[m] = 0;
    }
  });
}

int gic_device_bytes(gic_context *c, size_t *out) {
  return guarded(c, [&] { *out = total_bytes(c); });
}

}  // extern "C"
