// C entry points of the CPU oracle, for ctypes use by tests/ and bench.py's
// cpu_baseline leg.  TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <chrono>

using namespace orc;

static thread_local std::string g_err;

#define ORC_TRY(body)                \
  try {                              \
    body;                            \
    return 0;                        \
  } catch (std::exception & e) {     \
    g_err = e.what();                \
    return 1;                        \
  }

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc_ranlxs_raw(unsigned long seed, size_t n, uint32_t *out) {
  ORC_TRY({
    Ranlxs r(seed);
    for (size_t i = 0; i < n; ++i) out[i] = r.get();
  })
}

int orc_gaussians(unsigned long seed, size_t n, double *out, size_t *raw_used) {
  ORC_TRY({
    Ranlxs r(seed);
    for (size_t i = 0; i < n; ++i) out[i] = gaussian_ziggurat(r);
    if (raw_used) *raw_used = (r.nblocks) * 24 - (24 - r.pos);
  })
}

// The serial stream continued from a state at a block boundary (12 words,
// words[0] the oldest, and the carry), for windows far into a stream whose
// state a test has computed by other means.
int orc_ranlxs_raw_from_state(const int64_t *words, int carry, size_t n, uint32_t *out) {
  ORC_TRY({
    Ranlxs r(words, carry);
    for (size_t i = 0; i < n; ++i) out[i] = r.get();
  })
}

// Gaussians parsed from that state after discarding `skip` raw numbers (the
// parse then starts mid-block, possibly inside a ziggurat iteration of the
// true stream; parses from different offsets merge after a few rejections).
int orc_gaussians_from_state(const int64_t *words, int carry, size_t skip, size_t n, double *out) {
  ORC_TRY({
    Ranlxs r(words, carry);
    for (size_t i = 0; i < skip; ++i) r.get();
    for (size_t i = 0; i < n; ++i) out[i] = gaussian_ziggurat(r);
  })
}

int orc_zig_tables(double *ytab, double *wtab, uint32_t *ktab) {
  ORC_TRY({
    const ZigTables &t = ZigTables::get();
    memcpy(ytab, t.ytab, sizeof t.ytab);
    memcpy(wtab, t.wtab, sizeof t.wtab);
    memcpy(ktab, t.ktab, sizeof t.ktab);
  })
}

// in-place transforms on a padded buffer of 2 N^2 (N/2+1) doubles
int orc_fft_r2c(int n, double *buf) {
  ORC_TRY({
    std::vector<double> d(buf, buf + 2 * size_t(n) * n * (n / 2 + 1));
    fft_r2c_inplace(d, n);
    memcpy(buf, d.data(), d.size() * sizeof(double));
  })
}
int orc_fft_c2r(int n, double *buf) {
  ORC_TRY({
    std::vector<double> d(buf, buf + 2 * size_t(n) * n * (n / 2 + 1));
    fft_c2r_inplace(d, n);
    memcpy(buf, d.data(), d.size() * sizeof(double));
  })
}

int orc_growth_factor(double Om, double Ol, double a, double *out) {
  ORC_TRY({
    Cosmology c;
    c.OmegaM0 = Om;
    c.OmegaLambda0 = Ol;
    c.scalefactor = a;
    *out = growthFactor(c);
  })
}

// --- generator handle ---------------------------------------------------------
void *orc_ic_create() { return new IC(); }
void orc_ic_destroy(void *h) { delete (IC *)h; }

int orc_ic_set_cosmology(void *h, double Om, double Ol, double Ob, double hubble, double z, double s8, double ns) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    ic.cosmo.OmegaM0 = Om;
    ic.cosmo.OmegaLambda0 = Ol;
    ic.cosmo.OmegaBaryons0 = Ob;
    ic.cosmo.hubble = hubble;
    ic.cosmo.redshift = z;
    ic.cosmo.scalefactor = 1. / (z + 1.);
    ic.cosmo.sigma8 = s8;
    ic.cosmo.ns = ns;
  })
}

int orc_ic_set_camb(void *h, const double *table, size_t nvalues, int ncols) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    ic.camb = std::make_unique<CAMB>();
    ic.camb->init(ic.cosmo, std::vector<double>(table, table + nvalues), ncols);
  })
}

int orc_ic_camb_info(void *h, double *amplitude, double *sigma8pre) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    *amplitude = ic.camb->amplitude;
    *sigma8pre = ic.camb->varianceInSphere(8.);
  })
}

int orc_ic_pk(void *h, int n_k, const double *k, double *out) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    for (int i = 0; i < n_k; ++i) out[i] = ic.camb->P(k[i], true);
  })
}

int orc_ic_base_grid(void *h, double box, int n) { ORC_TRY(((IC *)h)->initBaseGrid(box, n)) }
int orc_ic_zoom_at(void *h, int x0, int y0, int z0, int fac, int n) {
  ORC_TRY(((IC *)h)->initZoomGridWithLowLeftCornerAt(x0, y0, z0, fac, n))
}
int orc_ic_seed(void *h, unsigned long seed, int mode) { ORC_TRY(((IC *)h)->setSeed(seed, (SeedMode)mode)) }
int orc_ic_draw(void *h) { ORC_TRY(((IC *)h)->draw()) }
int orc_ic_zero_level(void *h, int level) { ORC_TRY(((IC *)h)->zeroLevel(level)) }
int orc_ic_apply_ps(void *h) { ORC_TRY(((IC *)h)->applyPowerSpec()) }
int orc_ic_zeldovich(void *h) { ORC_TRY(((IC *)h)->zeldovich()) }

// which: 0 = noise/delta field, 1..3 = psi x,y,z.  fourier: 0 -> real space
// N^3 array, 1 -> padded FFTW layout (2 N^2 (N/2+1) doubles).
int orc_ic_get_field(void *h, int which, int level, int fourier, double *out) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    Field f = which == 0 ? ic.noise.at(level) : which <= 3 ? ic.psi.at(level).at(which - 1)
                                                           : ic.psi2.at(level).at(which - 4);
    if (fourier) {
      f.toFourier();
      memcpy(out, f.d.data(), f.d.size() * sizeof(double));
    } else {
      f.toReal();
      memcpy(out, f.d.data(), f.g->n3() * sizeof(double));
    }
  })
}

static void particles_level(IC &ic, int level, double *pos, double *vel) {
  size_t n3 = ic.grids.at(level)->n3();
#pragma omp parallel for
  for (size_t i = 0; i < n3; ++i) ic.particle(level, i, pos + 3 * i, vel + 3 * i);
}

int orc_ic_particles(void *h, int level, double *pos, double *vel, double *mass, double *eps) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    particles_level(ic, level, pos, vel);
    *mass = ic.particleMass(level);
    *eps = ic.particleEps(level);
  })
}

int orc_ic_set_lpt_order(void *h, int order) { ORC_TRY(((IC *)h)->lptOrder = order) }
// overwrite the (white noise or overdensity) field of a level with n^3 real values
int orc_ic_set_field(void *h, int level, const double *vals) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    Field &f = ic.noise.at(level);
    f.fourier = false;
    std::copy(vals, vals + f.g->n3(), f.d.begin());
  })
}
int orc_ic_set_flags(void *h, int level, const uint64_t *cells, size_t n) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    if (ic.flags.size() < ic.grids.size()) ic.flags.resize(ic.grids.size());
    if (level < 0 || level >= (int)ic.grids.size()) throw std::runtime_error("Specified level does not exist.");
    std::vector<size_t> f(cells, cells + n);
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
    ic.flags[level] = f;
  })
}
int orc_ic_set_variance_scale(void *h, double s) { ORC_TRY(((IC *)h)->varianceScale = s) }
int orc_ic_calculate(void *h, const char *name, double *out) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    ic.flags.resize(ic.grids.size());
    *out = ic.calculate(name);
  })
}
int orc_ic_add_modification(void *h, const char *name, int relative, double target) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    ic.flags.resize(ic.grids.size());
    ic.addModification(name, relative != 0, target);
  })
}
int orc_ic_apply_modifications(void *h, double *pre, double *post) {
  ORC_TRY(((IC *)h)->applyModifications(pre, post))
}
int orc_ic_clear_modifications(void *h) { ORC_TRY(((IC *)h)->clearModifications()) }
int orc_ic_chi2(void *h, double *out, size_t *dof) {
  ORC_TRY({
    IC &ic = *(IC *)h;
    *out = ic.chi2();
    if (dof) *dof = ic.numDof();
  })
}

int orc_ic_velocity_ratio(void *h, double *out) { ORC_TRY(*out = zeldovichVelocityToOffsetRatio(((IC *)h)->cosmo)) }

// Times the reference stage order seed -> delta, psi on one level at size n
// (serial real-space stream, r2c, sqrt P, 3 c2r, c2r of delta).  Used by
// bench.py's cpu_baseline leg.
int orc_time_pipeline(double box, int n, unsigned long seed, const double *table, size_t nvalues, int ncols,
                      double Om, double Ol, double s8, double ns, double z, double *seconds_total,
                      double *seconds_rng) {
  ORC_TRY({
    auto t0 = std::chrono::steady_clock::now();
    IC ic;
    ic.cosmo.OmegaM0 = Om;
    ic.cosmo.OmegaLambda0 = Ol;
    ic.cosmo.sigma8 = s8;
    ic.cosmo.ns = ns;
    ic.cosmo.redshift = z;
    ic.cosmo.scalefactor = 1. / (z + 1.);
    ic.camb = std::make_unique<CAMB>();
    ic.camb->init(ic.cosmo, std::vector<double>(table, table + nvalues), ncols);
    ic.initBaseGrid(box, n);
    ic.setSeed(seed, SeedMode::Real);
    auto t1 = std::chrono::steady_clock::now();
    ic.draw();
    auto t2 = std::chrono::steady_clock::now();
    ic.applyPowerSpec();
    ic.zeldovich();
    ic.noise[0].toReal();
    auto t3 = std::chrono::steady_clock::now();
    *seconds_total = std::chrono::duration<double>(t3 - t1).count();
    *seconds_rng = std::chrono::duration<double>(t2 - t1).count();
    (void)t0;
  })
}

// Median wall time (ms) of reps pairs of an in-place r2c and c2r of a
// random n^3 field on the host's threads, after one untimed pair: the FFT
// stage alone, for the CPU baseline's FFT-only comparison.
int orc_time_fft(int n, int reps, double *ms_median) {
  ORC_TRY({
    std::vector<double> d(2 * size_t(n) * n * (n / 2 + 1));
    uint64_t x = 88172645463325252ull;
    for (size_t i = 0; i < size_t(n) * n * n; ++i) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      d[i] = double(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    fft_r2c_inplace(d, n);
    fft_c2r_inplace(d, n);
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
      auto a = std::chrono::steady_clock::now();
      fft_r2c_inplace(d, n);
      fft_c2r_inplace(d, n);
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
    }
    std::sort(t.begin(), t.end());
    *ms_median = t[t.size() / 2];
  })
}

}  // extern "C"
