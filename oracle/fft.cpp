// Oracle: FFTW-convention 3D real transforms (genetIC/src/tools/numerics/
// fourier.hpp:300-563).  FFTW is arithmetic-equivalent to any exact DFT, so
// this is a plain radix-2 DFT along z, y, x with OpenMP over lines; the
// conventions restated are the padded [N][N][N/2+1] layout, the forward sign
// e^{-2 pi i k.x/N}, the N^-1.5 scaling both ways, and the Hermitian mirroring
// of the kz in {0, N/2} planes before c2r (positive modes win).
// Pinned by brute-force DFT and Parseval tests (tests/test_cpu_oracle.py).
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include "oracle.hpp"
#include <cmath>
#include <stdexcept>
#include <algorithm>

namespace orc {

void fft1d(cplx *a, int n, int sign) {
  if (n <= 1) return;
  if (n & (n - 1)) {
    // generic DFT (not used by genetIC grids, which are powers of two)
    std::vector<cplx> out(n);
    for (int k = 0; k < n; ++k) {
      cplx s = 0;
      for (int j = 0; j < n; ++j) s += a[j] * std::polar(1.0, sign * 2 * M_PI * double((long)k * j % n) / n);
      out[k] = s;
    }
    std::copy(out.begin(), out.end(), a);
    return;
  }
  // bit reversal
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len / 2;
    for (int k = 0; k < half; ++k) {
      cplx w = std::polar(1.0, sign * 2 * M_PI * k / len);
      for (int i = 0; i < n; i += len) {
        cplx u = a[i + k], v = a[i + k + half] * w;
        a[i + k] = u + v;
        a[i + k + half] = u - v;
      }
    }
  }
}

static inline size_t cidx(int n, int kx, int ky, int kz) {
  return (size_t(kx) * n + ky) * (n / 2 + 1) + kz;
}

void mirror_modes(std::vector<double> &d, int n) {
  // fourier.hpp:331-357: for kx in [0,N/2], ky from N/2 down to -N/2+1,
  // copy conj(mode(kx,ky,kz)) into mode(-kx,-ky,kz) for kz in {0, N/2}.
  auto loc = [n](int kx, int ky, int kz) {
    // getRealCoeffLocationAndConjugation with kz >= 0
    if (ky < 0) ky += n;
    if (kx < 0) kx += n;
    return 2 * cidx(n, kx, ky, kz);
  };
  int nyq = n / 2;
#pragma omp parallel for
  for (int kx = 0; kx < n / 2 + 1; ++kx) {
    for (int ky = n / 2; ky > -n / 2; --ky) {
      for (int kz : {0, nyq}) {
        size_t s = loc(kx, ky, kz), t = loc(-kx, -ky, kz);
        d[t] = d[s];
        d[t + 1] = -d[s + 1];
      }
    }
  }
}

void fft_r2c_inplace(std::vector<double> &d, int n) {
  const int nz = n / 2 + 1;
  std::vector<double> real(d.begin(), d.begin() + size_t(n) * n * n);
  double norm = pow(double(n), 1.5);
  // z transforms
#pragma omp parallel
  {
    std::vector<cplx> line(n);
#pragma omp for
    for (int xy = 0; xy < n * n; ++xy) {
      for (int z = 0; z < n; ++z) line[z] = real[size_t(xy) * n + z];
      fft1d(line.data(), n, -1);
      for (int kz = 0; kz < nz; ++kz) {
        d[2 * (size_t(xy) * nz + kz)] = line[kz].real();
        d[2 * (size_t(xy) * nz + kz) + 1] = line[kz].imag();
      }
    }
  }
  // y then x transforms
  for (int axis = 1; axis >= 0; --axis) {
#pragma omp parallel
    {
      std::vector<cplx> line(n);
#pragma omp for
      for (int a = 0; a < n * nz; ++a) {
        int other = a / nz, kz = a % nz;
        for (int t = 0; t < n; ++t) {
          size_t id = axis == 1 ? cidx(n, other, t, kz) : cidx(n, t, other, kz);
          line[t] = cplx(d[2 * id], d[2 * id + 1]);
        }
        fft1d(line.data(), n, -1);
        for (int t = 0; t < n; ++t) {
          size_t id = axis == 1 ? cidx(n, other, t, kz) : cidx(n, t, other, kz);
          d[2 * id] = line[t].real();
          d[2 * id + 1] = line[t].imag();
        }
      }
    }
  }
  for (auto &v : d) v /= norm;
}

void fft_c2r_inplace(std::vector<double> &d, int n) {
  const int nz = n / 2 + 1;
  double norm = pow(double(n), 1.5);
  mirror_modes(d, n);
  for (int axis = 0; axis <= 1; ++axis) {
#pragma omp parallel
    {
      std::vector<cplx> line(n);
#pragma omp for
      for (int a = 0; a < n * nz; ++a) {
        int other = a / nz, kz = a % nz;
        for (int t = 0; t < n; ++t) {
          size_t id = axis == 1 ? cidx(n, other, t, kz) : cidx(n, t, other, kz);
          line[t] = cplx(d[2 * id], d[2 * id + 1]);
        }
        fft1d(line.data(), n, +1);
        for (int t = 0; t < n; ++t) {
          size_t id = axis == 1 ? cidx(n, other, t, kz) : cidx(n, t, other, kz);
          d[2 * id] = line[t].real();
          d[2 * id + 1] = line[t].imag();
        }
      }
    }
  }
  // z: Hermitian-extend each line and take the real part of the inverse
  std::vector<double> real(size_t(n) * n * n);
#pragma omp parallel
  {
    std::vector<cplx> line(n);
#pragma omp for
    for (int xy = 0; xy < n * n; ++xy) {
      for (int kz = 0; kz < nz; ++kz)
        line[kz] = cplx(d[2 * (size_t(xy) * nz + kz)], d[2 * (size_t(xy) * nz + kz) + 1]);
      // c2r ignores the imaginary parts of the kz=0 and kz=N/2 entries
      line[0] = line[0].real();
      if (n > 1) line[n / 2] = line[n / 2].real();
      for (int kz = nz; kz < n; ++kz) line[kz] = std::conj(line[n - kz]);
      fft1d(line.data(), n, +1);
      for (int z = 0; z < n; ++z) real[size_t(xy) * n + z] = line[z].real();
    }
  }
  std::fill(d.begin(), d.end(), 0.0);
  for (size_t i = 0; i < real.size(); ++i) d[i] = real[i] / norm;
}

// --- Field ------------------------------------------------------------------
void Field::toFourier() {
  if (fourier) return;
  fft_r2c_inplace(d, g->n);
  fourier = true;
}
void Field::toReal() {
  if (!fourier) return;
  fft_c2r_inplace(d, g->n);
  fourier = false;
}
void Field::mirror() { mirror_modes(d, g->n); }

// getRealCoeffLocationAndConjugation (fourier.hpp:300-319)
static inline size_t coeffLoc(int n, int kx, int ky, int kz, bool &conj) {
  conj = false;
  if (kz < 0) {
    conj = true;
    kx = -kx;
    ky = -ky;
    kz = -kz;
  }
  if (ky < 0) ky += n;
  if (kx < 0) kx += n;
  return 2 * cidx(n, kx, ky, kz);
}

cplx Field::getK(int kx, int ky, int kz) const {
  bool c;
  size_t i = coeffLoc(g->n, kx, ky, kz, c);
  return cplx(d[i], c ? -d[i + 1] : d[i + 1]);
}
void Field::setK(int kx, int ky, int kz, cplx v) {
  bool c;
  size_t i = coeffLoc(g->n, kx, ky, kz, c);
  d[i] = v.real();
  d[i + 1] = c ? -v.imag() : v.imag();
}

// iterateFourierCellsWithAccumulation loop (fourier.hpp:110-151)
template <class F>
static void iterateK(int size, F &&cb) {
  int nyq = size / 2;
  int lneg = -size / 2 + 1;
  for (int kx = 0; kx < size / 2 + 1; kx++) {
    int ky_lower = lneg, wx = 2;
    if (kx == 0 || kx == nyq) {
      ky_lower = 0;
      wx = 1;
    }
    for (int ky = ky_lower; ky < size / 2 + 1; ky++) {
      int wy = wx, kz_lower = lneg;
      if ((kx == 0 || kx == nyq) && (ky == 0 || ky == nyq)) kz_lower = 0;
      if (ky_lower == 0 && ky != 0 && ky != nyq) wy *= 2;
      for (int kz = kz_lower; kz < size / 2 + 1; kz++) {
        int wz = wy;
        if (kz_lower == 0 && kz != 0 && kz != nyq) wz *= 2;
        cb(kx, ky, kz, wz);
      }
    }
  }
}

void Field::forEachK(const std::function<cplx(cplx, int, int, int)> &fn) {
  toFourier();
  iterateK(g->n, [&](int kx, int ky, int kz, int) { setK(kx, ky, kz, fn(getK(kx, ky, kz), kx, ky, kz)); });
}

cplx Field::accumulateK(const std::function<cplx(cplx, int, int, int)> &fn) const {
  double re = 0, im = 0;
  iterateK(g->n, [&](int kx, int ky, int kz, int w) {
    cplx r = fn(getK(kx, ky, kz), kx, ky, kz);
    re += r.real() * w;
    im += r.imag() * w;
  });
  return cplx(re, im);
}

void Field::applyFilterFn(const std::function<double(double)> &filt) {
  // field.hpp:741-746, k from comoving wavenumbers kx*kMin etc.
  double kMin = g->kmin();
  forEachK([&](cplx v, int kx, int ky, int kz) {
    double fx = kx * kMin, fy = ky * kMin, fz = kz * kMin;
    double k = sqrt(double(fx * fx + fy * fy + fz * fz));
    return v * filt(k);
  });
}

}  // namespace orc
