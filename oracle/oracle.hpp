// CPU oracle for the genetIC field-generation path.
//
// TEST INFRASTRUCTURE ONLY.  This directory restates, in plain FP64 C++, the
// reference algorithm of pynbody/genetIC for: the GSL ranlxs2 stream and
// ziggurat Gaussians, FFTW-convention r2c/c2r transforms, the CAMB power
// spectrum, Fermi level-split filters, tricubic level combination, Zel'dovich
// displacements, particle evaluation, covectors and the Hoffman-Ribak and
// quadratic modifications.  Only tests/, the smoke check and bench.py's
// cpu_baseline leg may load it, and only as the checker / CPU baseline.
// Nothing in the product (genetic_amd/) links or calls it.
//
// Pinning status of each function is recorded next to it and in DESIGN.md.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>
#include <complex>
#include <string>
#include <memory>
#include <functional>

namespace orc {

using cplx = std::complex<double>;

// ---------------------------------------------------------------- RNG -----
// GSL gsl_rng_ranlxs2 (Luescher's RANLUX, 48-bit double implementation,
// luxury p=397), reference call sites genetIC/src/simulation/field/
// randomfieldgenerator.hpp:43-45,115,229.  Pinned by the test_19/test_01c
// grid fixtures (noise phases to 1e-14).
struct Ranlxs {
  int64_t x[12];       // 48-bit words as integers, x[ir] is the oldest
  int carry = 0;
  int ir = 0, jr = 7;
  int pr = 397;
  uint32_t out[24];    // current block of 24-bit outputs
  int pos = 24;        // next output index in out[]
  uint64_t nblocks = 0;
  explicit Ranlxs(unsigned long seed, int luxury = 397);
  // the generator at a block boundary with the given 12 words (words[0] the
  // oldest) and carry: the next get() runs the 397 steps of a new block
  Ranlxs(const int64_t *words, int carry_in, int luxury = 397);
  void increment();
  uint32_t get() {                 // gsl_rng_get: 24-bit integer
    if (pos >= 24) increment();
    return out[pos++];
  }
  double uniform() { return get() / 16777216.0; }   // gsl_rng_uniform
};

// Ziggurat tables of GSL's gausszig.c, rebuilt from the Marsaglia-Tsang
// construction (128 strips, exponential-wedge base) in long double and
// rounded as GSL prints them (%.12g for ytab/wtab, truncation for ktab).
struct ZigTables {
  double ytab[128];
  double wtab[128];
  uint32_t ktab[128];
  static const ZigTables &get();
};

// gsl_ran_gaussian_ziggurat(r, 1.0) (randomfieldgenerator.hpp:149-150,191)
double gaussian_ziggurat(Ranlxs &r);

// ---------------------------------------------------------------- grid ----
// genetIC/src/simulation/grid/grid.hpp: index=(x*N+y)*N+z, centroid=(i+1/2)dx+offset
struct Grid {
  double periodic = 0;   // periodicDomainSize (box of the simulation)
  double size = 0;       // thisGridSize = dx*n
  double dx = 0;         // cellSize
  double ox = 0, oy = 0, oz = 0;  // offsetLower
  int n = 0;
  double kmin() const;                       // 2 pi / thisGridSize
  bool coversFull() const { return periodic == size; }
  int simEquivalentSize() const;
  size_t n3() const { return size_t(n) * n * n; }
  size_t padded() const { return 2 * size_t(n) * n * (n / 2 + 1); }
};

// ---------------------------------------------------------------- field ---
// Field storage follows field.hpp / fourier.hpp: the vector always has the
// padded size 2 N^2 (N/2+1); real space occupies the first N^3 entries
// (unpadded), Fourier space is FFTW's [N][N][N/2+1] interleaved complex.
struct Field {
  const Grid *g = nullptr;
  std::vector<double> d;
  bool fourier = false;
  Field() = default;
  explicit Field(const Grid &gr, bool f = false) : g(&gr), d(gr.padded(), 0.0), fourier(f) {}
  void toFourier();
  void toReal();
  cplx getK(int kx, int ky, int kz) const;
  void setK(int kx, int ky, int kz, cplx v);
  void mirror();   // ensureFourierModesAreMirrored
  // forEachFourierCell over the half space with reference loop order
  void forEachK(const std::function<cplx(cplx, int, int, int)> &fn);
  cplx accumulateK(const std::function<cplx(cplx, int, int, int)> &fn) const;
  void applyFilterFn(const std::function<double(double)> &filt);
};

// fourier.hpp:479-563 transforms on a padded vector (unnormalised FFT then /N^1.5)
void fft_r2c_inplace(std::vector<double> &d, int n);   // real (unpadded N^3 prefix) -> FFTW layout
void fft_c2r_inplace(std::vector<double> &d, int n);   // FFTW layout -> real (unpadded prefix)
void mirror_modes(std::vector<double> &d, int n);
void fft1d(cplx *a, int n, int sign);                  // unnormalised complex DFT, sign -1 forward

// ---------------------------------------------------------------- cosmology
struct Cosmology {
  double OmegaM0 = 0.3, OmegaLambda0 = 0, OmegaBaryons0 = 0, hubble = 0.701;
  double redshift = 0, scalefactor = 1, sigma8 = 0, ns = 0.96, TCMB = 2.725;
};
double growthFactor(const Cosmology &c);                 // parameters.hpp:55-70
double zeldovichVelocityToOffsetRatio(const Cosmology &c);  // parameters.hpp:83-106

// GSL natural cubic spline (gsl_interp_cspline) in log-log space,
// interpolation.hpp:20-110.
struct CubicSpline {
  std::vector<double> x, y, c;
  void init(const std::vector<double> &xs, const std::vector<double> &ys);
  double eval(double xv) const;    // zero outside [x0, xn]
};
struct LogInterpolator {
  CubicSpline s;
  double minx = 0, maxx = 0;
  void init(const std::vector<double> &x, const std::vector<double> &y);
  double operator()(double x) const;
};

// camb.hpp:173-332.  Columns: 0=k, 1=dm, 2=baryon, 6=total (species 'all').
struct CAMB {
  std::vector<double> kpts;
  LogInterpolator tf_dm, tf_all;
  double amplitude = 0, ns = 0;
  void init(const Cosmology &c, const std::vector<double> &table, int ncols);
  double P(double k, bool all = true) const;
  double varianceInSphere(double radius) const;
  static double normalisationForGrid(const Grid &g);
};
// getPowerSpectrumForGridUncached: covariance field on the grid (Fourier)
void powerSpectrumField(const CAMB &camb, const Grid &g, Field &out, bool all = true);

// ---------------------------------------------------------------- filters -
double lowPassFermi(double k, double kcut);               // filter.hpp:208-239
double highPassCov(double k, double kcut);               // filter.hpp:246-275

// ---------------------------------------------------------------- generator
enum class SeedMode { Real = 0, FourierSerial = 1, FourierParallel = 2, FourierReverse = 3 };

struct IC {
  Cosmology cosmo;
  std::unique_ptr<CAMB> camb;
  std::vector<std::unique_ptr<Grid>> grids;
  std::vector<double> weights;
  std::vector<Field> noise;               // output field (white noise, then delta)
  std::vector<std::vector<Field>> psi;    // per level: psi_x, psi_y, psi_z
  std::vector<std::vector<Field>> psi2;   // second-order displacements (lptOrder 2)
  int lptOrder = 1;
  unsigned long seed = 0;
  SeedMode seedMode = SeedMode::Real;
  bool seeded = false, drawn = false, powerApplied = false, combined = false;

  void initBaseGrid(double box, int n);
  void initZoomGridWithLowLeftCornerAt(int x0, int y0, int z0, int zoomfac, int n);
  void setSeed(unsigned long s, SeedMode m);
  void draw();
  void zeroLevel(int level);
  void applyPowerSpec();
  void zeldovich();       // ZA per level + level combination (multilevelgenerator.hpp:78-128)
  void secondOrder();     // 2LPT on the combined overdensity (no reference counterpart)
  double kcutForLevel(int level) const;   // filterfamily.hpp:110-143
  // particle evaluation on level (pos in Mpc/h wrapped, vel gadget units)
  void particle(int level, size_t id, double pos[3], double vel[3]) const;
  double particleMass(int level) const;
  double particleEps(int level) const;

  // --- modifications (modif.cpp) ---
  struct Mod {
    std::string name;
    bool quadratic = false;
    double target = 0, scale = -1, precision = 0.001;
    int initSteps = 10;
    std::vector<std::vector<size_t>> flags;
  };
  std::vector<std::vector<size_t>> flags;   // flagged cells per level, sorted
  double varianceScale = -1;
  std::vector<Mod> linear, quadratic;
  double levelFilter(int level, double k) const;
  double innerProduct(const std::vector<Field> &a, const std::vector<Field> &b) const;
  void applyMetric(std::vector<Field> &f, bool toCovector) const;
  void transferToDm(std::vector<Field> &f, bool fromWhiteNoise) const;
  int finestFlaggedCheck(const std::string &name, const std::vector<std::vector<size_t>> &fl) const;
  std::vector<Field> linearCovector(const std::string &name, const std::vector<std::vector<size_t>> &fl) const;
  std::vector<Field> pushVariance(const std::vector<Field> &field, const std::vector<std::vector<size_t>> &fl,
                                  double scale) const;
  double varianceValue(const std::vector<Field> &field, const std::vector<std::vector<size_t>> &fl,
                       double scale) const;
  void checkVarianceScale(double scale) const;
  double calculate(const std::string &name);
  void addModification(const std::string &name, bool relative, double target);
  void applyModifications(double *preChi2, double *postChi2);
  void clearModifications();
  double chi2();
  size_t numDof() const;
};

// tricubic (Catmull-Rom tensor) interpolation of a real-space field at a point
double evaluateInterpolated(const Field &f, double px, double py, double pz);
// fine <- fine + interpolate(source coarse) on every fine cell (evaluator.hpp:31-57)
void addFieldFromDifferentGrid(Field &fine, const Field &coarse);

}  // namespace orc
