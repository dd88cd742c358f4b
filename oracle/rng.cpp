// Oracle: GSL ranlxs2 + gsl_ran_gaussian_ziggurat restated from the published
// algorithms (Luescher 1994, "A portable high-quality random number generator
// for lattice field theory simulations"; Marsaglia & Tsang 2000, "The
// ziggurat method for generating random variables", with GSL's 128-strip /
// exponential-wedge variant).  GSL is not vendored in the reference; its
// distro version is unpinned (reference Dockerfile:4).  The reference calls
// these at genetIC/src/simulation/field/randomfieldgenerator.hpp:43-45,
// 149-150,191,228-229.
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace orc {

static const int64_t MASK48 = (int64_t(1) << 48) - 1;

Ranlxs::Ranlxs(unsigned long s, int luxury) {
  // Seeding: 31-bit shift register filled from the seed's low 32 bits,
  // clocked 48 times per word (Luescher's rlxd_init / GSL ranlxs_set_lux).
  if (s == 0) s = 1;
  long seed = (long)s;
  int i = (int)(seed & 0xFFFFFFFFUL);   // C conversion, as in the reference build
  int xbit[31];
  for (int k = 0; k < 31; ++k) {
    xbit[k] = i % 2;
    i /= 2;
  }
  int ibit = 0, jbit = 18;
  for (int k = 0; k < 12; ++k) {
    int64_t xv = 0;
    for (int m = 1; m <= 48; ++m) {
      int64_t y = xbit[ibit];
      xv += xv + y;
      xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
      ibit = (ibit + 1) % 31;
      jbit = (jbit + 1) % 31;
    }
    x[k] = xv;
  }
  carry = 0;
  ir = 0;
  jr = 7;
  pr = luxury;
  pos = 24;
}

Ranlxs::Ranlxs(const int64_t *words, int carry_in, int luxury) {
  for (int k = 0; k < 12; ++k) x[k] = words[k] & MASK48;
  carry = carry_in ? 1 : 0;
  ir = 0;
  jr = 7;
  pr = luxury;
  pos = 24;
}

void Ranlxs::increment() {
  // pr steps of the subtract-with-borrow recurrence x_n = x_{n-5} - x_{n-12} - c
  // in base 2^48, then emit the 12 newest words (oldest first) as low/high
  // 24-bit halves.
  for (int k = 0; k < pr; ++k) {
    int64_t d = x[jr] - x[ir] - carry;
    if (d < 0) {
      carry = 1;
      d += (int64_t(1) << 48);
    } else {
      carry = 0;
    }
    x[ir] = d;
    ir = (ir + 1) % 12;
    jr = (jr + 1) % 12;
  }
  for (int k = 0; k < 12; ++k) {
    int64_t w = x[(ir + k) % 12] & MASK48;
    out[2 * k] = uint32_t(w & 0xFFFFFF);
    out[2 * k + 1] = uint32_t(w >> 24);
  }
  pos = 0;
  ++nblocks;
}

// --- ziggurat tables --------------------------------------------------------
// Strip i spans x in [0, X[i+1]] with inner edge X[i]; X[0]=0, X[127]=R,
// X[128]=v/f(R) (virtual width of the base strip), equal strip area
// v = f(R)(R + 1/R) (exponential wedge tail).  R is fixed by requiring the top
// strip to close at f=1.  Computed here in long double.
static long double zf(long double x) { return expl(-x * x / 2); }
static long double zfinv(long double y) { return sqrtl(-2 * logl(y)); }

static long double zig_closure(long double R, long double *X) {
  long double v = zf(R) * (R + 1 / R);
  X[127] = R;
  X[128] = v / zf(R);
  for (int i = 126; i >= 1; --i) X[i] = zfinv(zf(X[i + 1]) + v / X[i + 1]);
  X[0] = 0;
  return zf(X[1]) + v / X[1] - 1;
}

static double round12(long double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.12Lg", v);
  return strtod(buf, nullptr);
}

const ZigTables &ZigTables::get() {
  static ZigTables t = [] {
    ZigTables z;
    long double X[129];
    // secant solve of the closure condition around the published R
    long double a = 3.4442864767L, b = 3.4442864768L;
    long double fa = zig_closure(a, X), fb = zig_closure(b, X);
    for (int it = 0; it < 60 && fb != fa; ++it) {
      long double c = b - fb * (b - a) / (fb - fa);
      a = b; fa = fb;
      b = c; fb = zig_closure(b, X);
      if (fabsl(b - a) < 1e-19L) break;
    }
    zig_closure(b, X);
    for (int i = 0; i < 128; ++i) {
      z.ytab[i] = round12(zf(X[i]));
      z.wtab[i] = round12(X[i + 1] / 16777216.0L);
      z.ktab[i] = (uint32_t)floorl(16777216.0L * X[i] / X[i + 1]);
    }
    z.ytab[0] = 1.0;
    z.ktab[0] = 0;
    return z;
  }();
  return t;
}

double gaussian_ziggurat(Ranlxs &r) {
  const ZigTables &T = ZigTables::get();
  const double PARAM_R = 3.44428647676;
  uint32_t i, j;
  int sign;
  double x, y;
  while (true) {
    // ranlxs2's range is 2^24-1, so GSL draws the strip byte and the
    // 24-bit abscissa from two successive outputs.
    uint32_t k1 = r.get();
    uint32_t k2 = r.get();
    i = k1 & 0xFF;
    j = k2 & 0xFFFFFF;
    sign = (i & 0x80) ? +1 : -1;
    i &= 0x7f;
    x = j * T.wtab[i];
    if (j < T.ktab[i]) break;
    if (i < 127) {
      double y0 = T.ytab[i], y1 = T.ytab[i + 1];
      double U1 = r.uniform();
      y = y1 + (y0 - y1) * U1;
    } else {
      double U1 = 1.0 - r.uniform();
      double U2 = r.uniform();
      x = PARAM_R - log(U1) / PARAM_R;
      y = exp(-PARAM_R * (x - 0.5 * PARAM_R)) * U2;
    }
    if (y < exp(-0.5 * x * x)) break;
  }
  return sign * x;
}

}  // namespace orc
