// Fourier-space operators, level combination and particle evaluation.
//
//   sqrt P(k) table      CAMB::operator() + getPowerSpectrumForGridUncached
//                        (genetIC/src/cosmology/camb.hpp:101-119,204-216) and the
//                        log-log cubic spline (tools/numerics/interpolation.hpp:81-110).
//                        Every mode's |k|^2 is an integer multiple of kMin^2,
//                        so one FP64 table indexed by kx^2+ky^2+kz^2 replaces the
//                        reference's per-mode spline calls and its covariance field.
//   apply transfer       Field::applyTransferFunction (field/field.hpp:394-412)
//   Zel'dovich           ZeldovichParticleGenerator::calculateOffsetFields
//                        (particles/zeldovich.hpp:155-181)
//   filters              LowPassFermiFilter / ComplementaryCovarianceFilterAdaptor
//                        (filters/filter.hpp:208-275), applied per mode
//   level combination    Field::addFieldFromDifferentGrid + evaluateInterpolated
//                        (field/field.hpp:483-617,804-812, evaluator.hpp:31-57)
//   particles            ZeldovichParticleEvaluator::getParticleNoWrap + wrapPoint
//                        (particles/zeldovich.hpp:87-109, generator.hpp:49-53)
#include "kspace.h"

namespace gic {

// mode index -> signed integer wavenumbers, with the stored N/2 read as +N/2
__device__ __forceinline__ void mode_ints(size_t idx, int n, int &ix, int &iy, int &iz) {
  int nz = n / 2 + 1;
  if ((n & (n - 1)) == 0 && idx < (size_t(1) << 32)) {
    // the transform sizes (powers of two up to 1024): one 32-bit division,
    // then a shift and a mask
    uint32_t i32 = uint32_t(idx), r = i32 / uint32_t(nz);
    iz = int(i32 - r * uint32_t(nz));
    iy = int(r & uint32_t(n - 1));
    ix = int(r >> (31 - __clz(n)));
  } else {
    iz = int(idx % nz);
    size_t r = idx / nz;
    iy = int(r % n);
    ix = int(r / n);
  }
  int h = n / 2;
  if (ix > h) ix -= n;
  if (iy > h) iy -= n;
}

// ---- spline-evaluated sqrt(P) table ----------------------------------------------
struct SplineDev {
  const double *x, *y, *c;
  int n;
  double minx, maxx;
};

__device__ double spline_eval(const SplineDev &s, double xv) {
  if (xv < s.x[0] || xv > s.x[s.n - 1]) return 0;
  int lo = 0, hi = s.n - 1;
  while (hi > lo + 1) {
    int i = (hi + lo) / 2;
    if (s.x[i] > xv) hi = i;
    else lo = i;
  }
  double xl = s.x[lo], xh = s.x[lo + 1], dx = xh - xl;
  if (!(dx > 0)) return s.y[lo];
  double dy = s.y[lo + 1] - s.y[lo], delx = xv - xl;
  double ci = s.c[lo], cip = s.c[lo + 1];
  double bi = dy / dx - dx * (cip + 2.0 * ci) / 3.0;
  double di = (cip - ci) / (3.0 * dx);
  return s.y[lo] + delx * (bi + delx * (ci + delx * di));
}

template <typename T>
__global__ void k_sqrtp_table(SplineDev sp, double amplitude, double ns, double kmin, double norm, long nent,
                              T *__restrict__ out) {
  long s = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nent) return;
  double k = sqrt(double(s)) * kmin;
  double tf = 0;
  if (k != 0) {
    if (sp.n == 0) {
      tf = 1;   // a pure power law (PowerLawPowerSpectrum, camb.hpp:145-168)
    } else {
      double lx = log(k);
      if (!(lx < sp.minx || lx > sp.maxx)) tf = exp(spline_eval(sp, lx));
    }
  }
  double P = amplitude * pow(k, ns) * tf * tf * norm;
  out[s] = T(P != 0.0 ? pow(P, 0.5) : 0.0);
}

template <typename T>
void sqrtp_table(const SplineHost &sp, double amplitude, double ns, double kmin, double norm, long nent, T *out,
                 hipStream_t st) {
  SplineDev d{sp.d_x, sp.d_y, sp.d_c, sp.n, sp.minx, sp.maxx};
  hipLaunchKernelGGL(k_sqrtp_table<T>, dim3((nent + 255) / 256), dim3(256), 0, st, d, amplitude, ns, kmin, norm, nent,
                     out);
  GIC_HIP(hipGetLastError());
}

// ---- filter tables by |k|^2 -------------------------------------------------------
// kind 0: low pass 1/(1+exp((k-kc)/(0.1 kc)))
// kind 1: high pass sqrt(1 - lp^2)
// kind 2: unity
// over_k2: divide by k^2 (zero at k = 0), the Zel'dovich kernel's 1/k^2
template <typename T>
__global__ void k_filter_table(int kind, double kc, double kmin, long nent, int over_k2, double scale,
                               T *__restrict__ out) {
  long s = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nent) return;
  double k = sqrt(double(s)) * kmin;
  double v = 1.0;
  if (kind != 2) {
    double lp = 1. / (1. + exp((k - kc) / (kc * 0.1)));
    v = kind == 0 ? lp : sqrt(1. - lp * lp);
  }
  if (over_k2) v = s == 0 ? 0.0 : v / (double(s) * kmin * kmin);
  out[s] = T(v * scale);
}

template <typename T>
void filter_table(int kind, double kc, double kmin, long nent, T *out, hipStream_t st, bool over_k2, double scale) {
  hipLaunchKernelGGL(k_filter_table<T>, dim3((nent + 255) / 256), dim3(256), 0, st, kind, kc, kmin, nent,
                     over_k2 ? 1 : 0, scale, out);
  GIC_HIP(hipGetLastError());
}

// ---- multiply a spectrum by a table indexed by |k|^2 ---------------------------------
template <typename T>
__global__ void k_mul_table(cx<T> *__restrict__ f, int n, const T *__restrict__ tab) {
  size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nm = size_t(n) * n * (n / 2 + 1);
  if (idx >= nm) return;
  int ix, iy, iz;
  mode_ints(idx, n, ix, iy, iz);
  T v = tab[ix * ix + iy * iy + iz * iz];
  cx<T> a = f[idx];
  f[idx] = cx<T>{a.x * v, a.y * v};
}

// the same for FP32 spectra, two adjacent modes per lane (one 16-byte load
// and store); n * n is even, so the modes pair up exactly
__global__ void k_mul_table2(float4 *__restrict__ f, int n, const float *__restrict__ tab) {
  size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t np = size_t(n) * n * (n / 2 + 1) / 2;
  if (p >= np) return;
  int x0, y0, z0, x1, y1, z1;
  mode_ints(2 * p, n, x0, y0, z0);
  mode_ints(2 * p + 1, n, x1, y1, z1);
  float v0 = tab[x0 * x0 + y0 * y0 + z0 * z0], v1 = tab[x1 * x1 + y1 * y1 + z1 * z1];
  float4 a = f[p];
  f[p] = float4{a.x * v0, a.y * v0, a.z * v1, a.w * v1};
}

template <typename T> void mul_table(T *field, int n, const T *tab, hipStream_t st) {
  size_t nm = size_t(n) * n * (n / 2 + 1);
  if constexpr (sizeof(T) == 4) {
    hipLaunchKernelGGL(k_mul_table2, dim3((nm / 2 + 255) / 256), dim3(256), 0, st, reinterpret_cast<float4 *>(field),
                       n, (const float *)tab);
  } else {
    hipLaunchKernelGGL(k_mul_table<T>, dim3((nm + 255) / 256), dim3(256), 0, st, reinterpret_cast<cx<T> *>(field), n,
                       tab);
  }
  GIC_HIP(hipGetLastError());
}

// ---- Fourier-space white noise ------------------------------------------------------
// drawRandomForSpecifiedGridFourier / drawOneFourierMode
// (field/randomfieldgenerator.hpp:144-171,203-260): square shells ks < N/2 in
// order; in shell ks, for k1, k2 in [-ks, ks), the modes (ks,k1,k2),
// (k1,ks,k2), (k1,k2,ks) each take two Gaussians (a, b) of the stream and
// store (a, b) -- (b, a) in reverse mode -- through setFourierCoefficient
// (tools/numerics/fourier.hpp:300-319,439-452), plus the conjugate at -k when
// a component is 0.  Later writes to a slot win; the owner pass finds, per
// slot, the last write (key 2d+1 for draw d, 2d+2 for its mirror write) and
// the write pass stores only the winners.

// draws before shell ks: sum_{j<ks} 3 (2j)^2
__device__ __forceinline__ long shell_start(long ks) { return 2 * (ks - 1) * ks * (2 * ks - 1); }

__device__ __forceinline__ void decode_draw(long d, int h, int &k1, int &k2, int &k3) {
  long lo = 1, hi = h - 1;   // shell_start(lo) <= d < shell_start(hi + 1)
  while (lo < hi) {
    long mid = (lo + hi + 1) / 2;
    if (shell_start(mid) <= d) lo = mid;
    else hi = mid - 1;
  }
  long ks = lo;
  long r = d - shell_start(ks);
  long p = r / 3;
  int which = int(r - 3 * p);
  int a = int(-ks + p / (2 * ks)), b = int(-ks + p % (2 * ks));
  if (which == 0) {
    k1 = int(ks), k2 = a, k3 = b;
  } else if (which == 1) {
    k1 = a, k2 = int(ks), k3 = b;
  } else {
    k1 = a, k2 = b, k3 = int(ks);
  }
}

// storage slot of mode (kx,ky,kz) and whether the stored value is conjugated
__device__ __forceinline__ size_t mode_slot(int kx, int ky, int kz, int n, bool &conj) {
  conj = kz < 0;
  if (conj) {
    kx = -kx;
    ky = -ky;
    kz = -kz;
  }
  if (ky < 0) ky += n;
  if (kx < 0) kx += n;
  return (size_t(kx) * n + ky) * (n / 2 + 1) + kz;
}

__global__ void k_fourier_owner(int n, long ndraw, uint32_t *__restrict__ owner) {
  long d = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (d >= ndraw) return;
  int k1, k2, k3;
  decode_draw(d, n / 2, k1, k2, k3);
  bool cj;
  atomicMax(&owner[mode_slot(k1, k2, k3, n, cj)], uint32_t(2 * d + 1));
  if (k1 == 0 || k2 == 0 || k3 == 0) atomicMax(&owner[mode_slot(-k1, -k2, -k3, n, cj)], uint32_t(2 * d + 2));
}

template <typename T>
__global__ void k_fourier_write(int n, long ndraw, const T *__restrict__ gauss, const uint32_t *__restrict__ owner,
                                cx<T> *__restrict__ out, int reverse) {
  long d = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (d >= ndraw) return;
  int k1, k2, k3;
  decode_draw(d, n / 2, k1, k2, k3);
  T a = gauss[2 * d], b = gauss[2 * d + 1];
  T re = reverse ? b : a, im = reverse ? a : b;
  bool cj;
  size_t s = mode_slot(k1, k2, k3, n, cj);
  if (owner[s] == uint32_t(2 * d + 1)) out[s] = cx<T>{re, cj ? -im : im};
  if (k1 == 0 || k2 == 0 || k3 == 0) {
    s = mode_slot(-k1, -k2, -k3, n, cj);
    if (owner[s] == uint32_t(2 * d + 2)) out[s] = cx<T>{re, cj ? im : -im};
  }
}

long fourier_draw_count(int n) {
  long h = n / 2;
  return h >= 1 ? 2 * (h - 1) * h * (2 * h - 1) : 0;
}

template <typename T>
void fourier_draw_scatter(const T *gauss, int n, T *field, uint32_t *owner, int reverse, hipStream_t st) {
  long nd = fourier_draw_count(n);
  size_t nslot = size_t(n) * n * (n / 2 + 1);
  if (nd * 2 + 2 >= (1L << 32)) throw std::runtime_error("grid too large for the Fourier-space draw");
  fill<uint32_t>(owner, nslot, 0u, st);
  fill<T>(field, 2 * nslot, T(0), st);
  if (nd == 0) return;
  hipLaunchKernelGGL(k_fourier_owner, dim3((nd + 255) / 256), dim3(256), 0, st, n, nd, owner);
  GIC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_fourier_write<T>, dim3((nd + 255) / 256), dim3(256), 0, st, n, nd, gauss, owner,
                     reinterpret_cast<cx<T> *>(field), reverse);
  GIC_HIP(hipGetLastError());
}

// ---- |delta_k|^2 histogram by integer |k|^2 ----------------------------------------
// dumpPowerSpectrum (cosmology/parameters.hpp:112-205) sums over every
// (ix,iy,iz) in [-N/2, N/2]^3 -- the Nyquist planes are visited twice -- and
// the bin of a mode depends only on m = ix^2+iy^2+iz^2.  One thread per
// visited triple adds |delta|^2 and 1 into the two FP64 bins of its m.
template <typename T>
__global__ void k_mode_power(const cx<T> *__restrict__ f, int n, double *__restrict__ sum, double *__restrict__ cnt) {
  size_t w = size_t(n) + 1;
  size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= w * w * w) return;
  int h = n / 2;
  int iz = int(idx % w) - h;
  int iy = int((idx / w) % w) - h;
  int ix = int(idx / (w * w)) - h;
  int sx = ix, sy = iy, sz = iz;
  if (sz < 0) {   // conjugate: (kx,ky,kz) = conj(-kx,-ky,-kz); the modulus is the same
    sx = -sx;
    sy = -sy;
    sz = -sz;
  }
  if (sx < 0) sx += n;
  if (sy < 0) sy += n;
  cx<T> v = f[(size_t(sx) * n + sy) * (h + 1) + sz];
  double a = double(v.x) * double(v.x) + double(v.y) * double(v.y);
  int m = ix * ix + iy * iy + iz * iz;
  atomicAdd(&sum[m], a);
  atomicAdd(&cnt[m], 1.0);
}

template <typename T> void mode_power(const T *spec, int n, double *sum, double *cnt, hipStream_t st) {
  size_t w = size_t(n) + 1, tot = w * w * w;
  hipLaunchKernelGGL(k_mode_power<T>, dim3((tot + 255) / 256), dim3(256), 0, st,
                     reinterpret_cast<const cx<T> *>(spec), n, sum, cnt);
  GIC_HIP(hipGetLastError());
}

// ---- Zel'dovich: psi_j(k) = i k_j delta(k) / k^2 -----------------------------------------
// Computed in T with k_j = int * kMin as the reference does; the component
// at k_j = +Nyquist and the k = 0 mode are zero.  An optional table factor
// (the high-pass filter of a zoom level) is applied to the three outputs.
template <typename T>
__global__ void k_zeldovich(const cx<T> *__restrict__ d, int n, T kmin, cx<T> *__restrict__ px,
                            cx<T> *__restrict__ py, cx<T> *__restrict__ pz, const T *__restrict__ post) {
  size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nm = size_t(n) * n * (n / 2 + 1);
  if (idx >= nm) return;
  int ix, iy, iz;
  mode_ints(idx, n, ix, iy, iz);
  T kx = T(ix) * kmin, ky = T(iy) * kmin, kz = T(iz) * kmin;
  T nyq = T(n / 2) * kmin;
  T kfft = kx * kx + ky * ky + kz * kz;
  cx<T> v = d[idx];
  cx<T> r{-v.y / kfft, v.x / kfft};
  T f = post ? post[ix * ix + iy * iy + iz * iz] : T(1);
  cx<T> rx{r.x * kx * f, r.y * kx * f}, ry{r.x * ky * f, r.y * ky * f}, rz{r.x * kz * f, r.y * kz * f};
  if (kx == nyq || kfft == 0) rx = cx<T>{0, 0};
  if (ky == nyq || kfft == 0) ry = cx<T>{0, 0};
  if (kz == nyq || kfft == 0) rz = cx<T>{0, 0};
  px[idx] = rx;
  py[idx] = ry;
  pz[idx] = rz;
}

template <typename T>
void zeldovich(const T *delta, int n, double kmin, T *px, T *py, T *pz, const T *post, hipStream_t st) {
  size_t nm = size_t(n) * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_zeldovich<T>, dim3((nm + 255) / 256), dim3(256), 0, st, reinterpret_cast<const cx<T> *>(delta),
                     n, T(kmin), reinterpret_cast<cx<T> *>(px), reinterpret_cast<cx<T> *>(py),
                     reinterpret_cast<cx<T> *>(pz), post);
  GIC_HIP(hipGetLastError());
}

// ---- tricubic (Catmull-Rom tensor) coarse -> fine interpolate-add -----------------------
__device__ __forceinline__ void cr_weights(double t, double *w) {
  w[0] = -0.5 * t * (t - 1) * (t - 1);
  w[1] = (2 - 5 * t * t + 3 * t * t * t) / 2;
  w[2] = -t * (-1 - 4 * t + 3 * t * t) / 2;
  w[3] = (t - 1) * t * t / 2;
}

// The tricubic stencil of fine cell (x,y,z): the 4 coarse indices per axis
// (wrapped or clamped, then relative to org mod nc) and their Catmull-Rom
// weights.  False when the cell lies outside the coarse grid.
__device__ __forceinline__ bool interp_stencil(int x, int y, int z, const InterpGeom &g, const int *org, int ci[3][4],
                                               double w[3][4]) {
  int c[3] = {x + g.off[0], y + g.off[1], z + g.off[2]};
  int p0[3];
  double t[3];
  for (int a = 0; a < 3; ++a) {
    if (c[a] > g.sim - 1) c[a] -= g.sim;
    if (c[a] < 0) c[a] += g.sim;
    if (c[a] < 0 || c[a] >= g.nsup) return false;   // outside the coarse grid: nothing added
    double cen = c[a] * g.dxs + g.co[a] + g.dxs / 2;   // supersampled centroid
    double loc = cen - g.co[a];
    loc = fmod(loc, g.periodic);
    if (loc < 0) loc += g.periodic;
    double u = loc / g.dxc - 0.5;
    p0[a] = int(floor(u));
    t[a] = u - p0[a];
  }
  for (int a = 0; a < 3; ++a) cr_weights(t[a], w[a]);
  int nc = g.nc;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 4; ++i) {
      int v = p0[a] + i - 1;
      if (g.wrap) {
        if (v > g.csim - 1) v -= g.csim;
        if (v < 0) v += g.csim;
      } else {
        if (v > nc - 1) v = nc - 1;
        if (v < 0) v = 0;
      }
      int l = v - org[a];
      if (l < 0) l += nc;
      ci[a][i] = l;
    }
  return true;
}

template <typename T>
__global__ void k_interp_add(T *__restrict__ fine, int nf, CoarseView cv, InterpGeom g) {
  const T *__restrict__ coarse = reinterpret_cast<const T *>(cv.p);
  size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nf3 = size_t(nf) * nf * nf;
  if (idx >= nf3) return;
  int x = int(idx / (size_t(nf) * nf)), y = int((idx / nf) % nf), z = int(idx % nf);
  int ci[3][4];
  double w[3][4];
  if (!interp_stencil(x, y, z, g, cv.org, ci, w)) return;
  T acc = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      T wij = T(w[0][i] * w[1][j]);
      const T *row = coarse + ci[0][i] * cv.sx + ci[1][j] * cv.sy;
      for (int k = 0; k < 4; ++k) acc += wij * T(w[2][k]) * row[ci[2][k]];
    }
  fine[(size_t(x) * nf + y) * (nf + 2) + z] += acc;
}

template <typename T> void interp_add(T *fine, int nf, const CoarseView &coarse, const InterpGeom &g, hipStream_t st) {
  size_t nf3 = size_t(nf) * nf * nf;
  hipLaunchKernelGGL(k_interp_add<T>, dim3((nf3 + 255) / 256), dim3(256), 0, st, fine, nf, coarse, g);
  GIC_HIP(hipGetLastError());
}

// Transpose of k_interp_add onto a whole coarse grid: every fine value is
// scattered, with the same 64 weights, into the padded coarse field.  The
// fine field has row stride frs (nf for a dense array, nf+2 padded).
template <typename T>
__global__ void k_interp_adjoint(const T *__restrict__ fine, int nf, long frs, T *__restrict__ coarse, InterpGeom g) {
  size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nf3 = size_t(nf) * nf * nf;
  if (idx >= nf3) return;
  int x = int(idx / (size_t(nf) * nf)), y = int((idx / nf) % nf), z = int(idx % nf);
  int ci[3][4];
  double w[3][4];
  const int org[3] = {0, 0, 0};
  if (!interp_stencil(x, y, z, g, org, ci, w)) return;
  T v = fine[(size_t(x) * nf + y) * frs + z];
  long sy = g.nc + 2, sx = long(g.nc) * sy;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      T wij = T(w[0][i] * w[1][j]) * v;
      T *row = coarse + ci[0][i] * sx + ci[1][j] * sy;
      for (int k = 0; k < 4; ++k) atomicAdd(row + ci[2][k], wij * T(w[2][k]));
    }
}

template <typename T>
void interp_adjoint(const T *fine, int nf, long frs, T *coarse, const InterpGeom &g, hipStream_t st) {
  size_t nf3 = size_t(nf) * nf * nf;
  hipLaunchKernelGGL(k_interp_adjoint<T>, dim3((nf3 + 255) / 256), dim3(256), 0, st, fine, nf, frs, coarse, g);
  GIC_HIP(hipGetLastError());
}

// ---- particles --------------------------------------------------------------------------
template <typename T>
__global__ void k_particles(const T *__restrict__ px, const T *__restrict__ py, const T *__restrict__ pz,
                            const T *__restrict__ qx, const T *__restrict__ qy, const T *__restrict__ qz, int n,
                            ParticleGeom g, const uint64_t *__restrict__ cells, size_t start, size_t count,
                            float *__restrict__ pos, float *__restrict__ vel) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  size_t id = cells ? cells[i] : start + i;
  size_t nn = size_t(n) * n;
  size_t x = id / nn, r = id - x * nn, y = r / n, z = r - y * n;
  size_t p = (x * n + y) * (n + 2) + z;
  double off[3] = {double(px[p]), double(py[p]), double(pz[p])};
  double cen[3] = {double(x) * g.dx + g.o[0] + g.dx / 2, double(y) * g.dx + g.o[1] + g.dx / 2,
                   double(z) * g.dx + g.o[2] + g.dx / 2};
  const T *q2[3] = {qx, qy, qz};
  for (int a = 0; a < 3; ++a) {
    T o = T(off[a]);
    T v = o * T(g.vratio);
    T disp = T(off[a]);
    if (qx) {   // second order: x = q + psi1 + psi2, v = r (psi1 + f2/f1 psi2)
      T o2 = q2[a][p];
      disp += o2;
      v += o2 * T(g.vratio2);
    }
    T q = disp + T(cen[a]);
    double qd = double(q);
    qd = fmod(qd, g.periodic);
    if (qd < 0) qd += g.periodic;
    pos[3 * i + a] = float(qd);
    vel[3 * i + a] = float(v);
  }
}

template <typename T>
void particles(const T *px, const T *py, const T *pz, const T *const *psi2, int n, const ParticleGeom &g,
               const uint64_t *cells, size_t start, size_t count, float *pos, float *vel, hipStream_t st) {
  if (count == 0) return;
  const T *qx = psi2 ? psi2[0] : nullptr, *qy = psi2 ? psi2[1] : nullptr, *qz = psi2 ? psi2[2] : nullptr;
  hipLaunchKernelGGL(k_particles<T>, dim3((count + 255) / 256), dim3(256), 0, st, px, py, pz, qx, qy, qz, n, g, cells,
                     start, count, pos, vel);
  GIC_HIP(hipGetLastError());
}

// ---- second-order source -------------------------------------------------------------
// elementwise over a padded real field: mode 0 S = A B, 1 S += A B, 2 S -= A A, 3 A += B
template <typename T>
__global__ void k_lpt_accum(T *__restrict__ S, T *__restrict__ A, const T *__restrict__ B, size_t n, int mode) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mode == 0) S[i] = A[i] * B[i];
  else if (mode == 1) S[i] += A[i] * B[i];
  else if (mode == 2) S[i] -= A[i] * A[i];
  else A[i] += B[i];
}
// four consecutive elements per thread (n is a multiple of 4 for even grids)
template <typename T> struct alignas(4 * sizeof(T)) Vec4 {
  T v[4];
};
template <typename T>
__global__ void k_lpt_accum4(Vec4<T> *__restrict__ S, Vec4<T> *__restrict__ A, const Vec4<T> *__restrict__ B,
                             size_t n4, int mode) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  Vec4<T> a = A[i];
  if (mode == 3) {
    Vec4<T> b = B[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) a.v[k] += b.v[k];
    A[i] = a;
    return;
  }
  Vec4<T> s;
  if (mode == 0) {
    Vec4<T> b = B[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) s.v[k] = a.v[k] * b.v[k];
  } else if (mode == 1) {
    Vec4<T> b = B[i];
    s = S[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) s.v[k] += a.v[k] * b.v[k];
  } else {
    s = S[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) s.v[k] -= a.v[k] * a.v[k];
  }
  S[i] = s;
}
template <typename T> void lpt_accum(T *S, T *A, const T *B, size_t n, int mode, hipStream_t st) {
  if (n % 4 == 0) {
    size_t n4 = n / 4;
    hipLaunchKernelGGL(k_lpt_accum4<T>, dim3((n4 + 255) / 256), dim3(256), 0, st, reinterpret_cast<Vec4<T> *>(S),
                       reinterpret_cast<Vec4<T> *>(A), reinterpret_cast<const Vec4<T> *>(B), n4, mode);
  } else {
    hipLaunchKernelGGL(k_lpt_accum<T>, dim3((n + 255) / 256), dim3(256), 0, st, S, A, B, n, mode);
  }
  GIC_HIP(hipGetLastError());
}

// ---- white-noise adjustments ------------------------------------------------------------
// enforcePowerSpectrumOneGrid (multilevelfield.hpp:498-510): every mode to
// unit modulus, keeping its phase (0 stays 0)
template <typename T> __global__ void k_unit_modulus(cx<T> *__restrict__ f, size_t nm) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  cx<T> v = f[i];
  T a = sqrt(v.x * v.x + v.y * v.y);
  f[i] = a == T(0) ? cx<T>{T(0), T(0)} : cx<T>{v.x / a, v.y / a};
}
template <typename T> void unit_modulus(T *field, int n, hipStream_t st) {
  size_t nm = size_t(n) * n * (n / 2 + 1);
  hipLaunchKernelGGL(k_unit_modulus<T>, dim3((nm + 255) / 256), dim3(256), 0, st, reinterpret_cast<cx<T> *>(field), nm);
  GIC_HIP(hipGetLastError());
}
// reverseSmallK (ic.hpp:1687-1706): modes with 0 < k^2 < kmax^2 change sign,
// k = i kMin in the field precision as the reference forms it
template <typename T> __global__ void k_reverse_small_k(cx<T> *__restrict__ f, int n, T kmin, T k2max) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nm = size_t(n) * n * (n / 2 + 1);
  if (i >= nm) return;
  int ix, iy, iz;
  mode_ints(i, n, ix, iy, iz);
  T kx = T(ix) * kmin, ky = T(iy) * kmin, kz = T(iz) * kmin;
  T k2 = kx * kx + ky * ky + kz * kz;
  if (k2 < k2max && k2 != T(0)) {
    cx<T> v = f[i];
    f[i] = cx<T>{-v.x, -v.y};
  }
}
template <typename T> void reverse_small_k(T *field, int n, double kmin, double kmax, hipStream_t st) {
  size_t nm = size_t(n) * n * (n / 2 + 1);
  T km = T(kmax);
  hipLaunchKernelGGL(k_reverse_small_k<T>, dim3((nm + 255) / 256), dim3(256), 0, st, reinterpret_cast<cx<T> *>(field),
                     n, T(kmin), km * km);
  GIC_HIP(hipGetLastError());
}
template <typename T> __global__ void k_scale_all(T *__restrict__ p, size_t n, T s) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] *= s;
}
template <typename T> void scale_all(T *p, size_t n, T s, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_all<T>, dim3((n + 255) / 256), dim3(256), 0, st, p, n, s);
  GIC_HIP(hipGetLastError());
}
// out[m] = 1 / tab[m] (0 where tab is 0): pow(P, -1/2) from a sqrt P table
template <typename T> __global__ void k_inverse_table(const T *__restrict__ tab, long n, T *__restrict__ out) {
  long m = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < n) out[m] = tab[m] == T(0) ? T(0) : T(1) / tab[m];
}
template <typename T> void inverse_table(const T *tab, long n, T *out, hipStream_t st) {
  hipLaunchKernelGGL(k_inverse_table<T>, dim3((n + 255) / 256), dim3(256), 0, st, tab, n, out);
  GIC_HIP(hipGetLastError());
}

// ---- small utilities -------------------------------------------------------------------
template <typename T> __global__ void k_fill(T *p, size_t n, T v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
template <typename T> void fill(T *p, size_t n, T v, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
  GIC_HIP(hipGetLastError());
}

// copy a padded real field to / from a dense N^3 array of another type
template <typename S, typename D>
__global__ void k_unpad(const S *__restrict__ src, int n, D *__restrict__ dst) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  size_t row = i / n, z = i - row * n;
  dst[i] = D(src[row * (n + 2) + z]);
}
template <typename S, typename D>
__global__ void k_pad(const S *__restrict__ src, int n, D *__restrict__ dst) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  size_t row = i / n, z = i - row * n;
  dst[row * (n + 2) + z] = D(src[i]);
}
template <typename S, typename D> void unpad(const S *src, int n, D *dst, hipStream_t st) {
  size_t n3 = size_t(n) * n * n;
  hipLaunchKernelGGL((k_unpad<S, D>), dim3((n3 + 255) / 256), dim3(256), 0, st, src, n, dst);
  GIC_HIP(hipGetLastError());
}
template <typename S, typename D> void pad(const S *src, int n, D *dst, hipStream_t st) {
  size_t n3 = size_t(n) * n * n;
  hipLaunchKernelGGL((k_pad<S, D>), dim3((n3 + 255) / 256), dim3(256), 0, st, src, n, dst);
  GIC_HIP(hipGetLastError());
}
template <typename S, typename D> __global__ void k_convert(const S *__restrict__ a, size_t n, D *__restrict__ b) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = D(a[i]);
}
template <typename S, typename D> void convert(const S *a, size_t n, D *b, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert<S, D>), dim3((n + 255) / 256), dim3(256), 0, st, a, n, b);
  GIC_HIP(hipGetLastError());
}

#define GIC_INST(T)                                                                                            \
  template void sqrtp_table<T>(const SplineHost &, double, double, double, double, long, T *, hipStream_t);   \
  template void filter_table<T>(int, double, double, long, T *, hipStream_t, bool, double);                   \
  template void mul_table<T>(T *, int, const T *, hipStream_t);                                               \
  template void mode_power<T>(const T *, int, double *, double *, hipStream_t);                               \
  template void fourier_draw_scatter<T>(const T *, int, T *, uint32_t *, int, hipStream_t);                   \
  template void zeldovich<T>(const T *, int, double, T *, T *, T *, const T *, hipStream_t);                  \
  template void interp_add<T>(T *, int, const CoarseView &, const InterpGeom &, hipStream_t);                 \
  template void interp_adjoint<T>(const T *, int, long, T *, const InterpGeom &, hipStream_t);               \
  template void particles<T>(const T *, const T *, const T *, const T *const *, int, const ParticleGeom &,    \
                             const uint64_t *, size_t, size_t, float *, float *, hipStream_t);                \
  template void lpt_accum<T>(T *, T *, const T *, size_t, int, hipStream_t);                                  \
  template void fill<T>(T *, size_t, T, hipStream_t);                                                       \
  template void unit_modulus<T>(T *, int, hipStream_t);                                                     \
  template void reverse_small_k<T>(T *, int, double, double, hipStream_t);                                  \
  template void scale_all<T>(T *, size_t, T, hipStream_t);                                                  \
  template void inverse_table<T>(const T *, long, T *, hipStream_t);
GIC_INST(float)
GIC_INST(double)
template void fill<uint32_t>(uint32_t *, size_t, uint32_t, hipStream_t);
template void fill<uint8_t>(uint8_t *, size_t, uint8_t, hipStream_t);
template void unpad<float, float>(const float *, int, float *, hipStream_t);
template void unpad<float, double>(const float *, int, double *, hipStream_t);
template void unpad<double, double>(const double *, int, double *, hipStream_t);
template void unpad<double, float>(const double *, int, float *, hipStream_t);
template void pad<float, float>(const float *, int, float *, hipStream_t);
template void pad<double, float>(const double *, int, float *, hipStream_t);
template void pad<double, double>(const double *, int, double *, hipStream_t);
template void pad<float, double>(const float *, int, double *, hipStream_t);
template void convert<float, double>(const float *, size_t, double *, hipStream_t);
template void convert<double, float>(const double *, size_t, float *, hipStream_t);
template void convert<double, double>(const double *, size_t, double *, hipStream_t);
template void convert<float, float>(const float *, size_t, float *, hipStream_t);

}  // namespace gic
