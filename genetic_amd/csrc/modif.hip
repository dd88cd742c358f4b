// Constrained realisations on the GPU: multi-level covectors, the metric,
// inner products, chi^2 and the Hoffman-Ribak linear and quadratic
// (filtered-variance) modifications of the white-noise field.
//
//   filters per level        filters/filterfamily.hpp:79-143
//   inner product            field/multilevelfield.hpp:248-294 (weights of fourier.hpp:110-151)
//   metric                   field/multilevelfield.hpp:318-384, field.hpp:752-800
//   cross-level transfer     field.hpp:804-849 (FILTER_ON_COARSE_GRID), evaluator.hpp:166-206
//   deinterpolation          field.hpp:420-480, multilevelgrid.hpp:246-321
//   covectors                modifications/linearmodification.hpp:22-412
//   variance                 modifications/quadraticmodification.hpp:50-221
//   manager                  modifications/modificationmanager.hpp:80-331
//   chi^2                    multilevelfield.hpp:461-487
//
// Every multi-level field is one padded buffer per level on the device; the
// inner products reduce in FP64 whatever the field precision.
#include "context.h"

#include <strings.h>

namespace gic {

// ===================================================================== kernels ==
namespace {

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
// the chi^2 window sum, which holds few registers, runs more blocks than the
// other reductions to keep more reads in flight (its partials fit the
// reduction buffer, RED_BLOCKS * GRAM_ROWS doubles)
constexpr int CHI2_BLOCKS = 4 * RED_BLOCKS;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[RED_THREADS];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// The stored modes of an n^3 half spectrum, i = (x n + y)(n/2 + 1) + z, in a
// grid-stride loop: each thread locates its first mode with one division
// and steps (row, z) by the stride after that (n is a power of two, so the
// row splits into x and y by a shift and a mask).
// (x, y, z) of stored mode i of an n^3 half spectrum (unsigned indices)
__device__ __forceinline__ void mode_xyz(size_t i, int n, int &x, int &y, int &z) {
  int nz = n / 2 + 1;
  if ((n & (n - 1)) == 0 && i < (size_t(1) << 32)) {
    uint32_t i32 = uint32_t(i), r = i32 / uint32_t(nz);
    z = int(i32 - r * uint32_t(nz));
    y = int(r & uint32_t(n - 1));
    x = int(r >> (31 - __clz(n)));
  } else {
    z = int(i % nz);
    size_t r = i / nz;
    y = int(r % n);
    x = int(r / n);
  }
}
template <class F> __device__ __forceinline__ void for_each_mode(int n, F f) {
  const int nz = n / 2 + 1;
  const long nm = long(n) * n * nz;
  const long stride = long(gridDim.x) * blockDim.x;
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const int lg = 31 - __clz(n);
  long r = i / nz;
  int z = int(i - r * nz);
  const long sr = stride / nz;
  const int sz = int(stride - sr * nz);
  for (; i < nm; i += stride) {
    f(i, int(r >> lg), int(r & (n - 1)), z);
    z += sz;
    r += sr;
    if (z >= nz) {
      z -= nz;
      ++r;
    }
  }
}

// for_each_mode over pairs of adjacent stored modes (2p, 2p + 1), for FP32
// spectra read 16 bytes (two modes) per lane: f(i, x0, y0, z0, x1, y1, z1)
// with i = 2p (n * n is even, so the modes pair up exactly).
template <class F> __device__ __forceinline__ void for_each_mode_pair(int n, F f) {
  const int nz = n / 2 + 1;
  const long np = long(n) * n * nz / 2;
  const long stride = long(gridDim.x) * blockDim.x;
  long p = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int lg = 31 - __clz(n);
  long r = (2 * p) / nz;
  int z = int(2 * p - r * nz);
  const long sr = (2 * stride) / nz;
  const int sz = int(2 * stride - sr * nz);
  for (; p < np; p += stride) {
    int z1 = z + 1;
    long r1 = r;
    if (z1 == nz) {
      z1 = 0;
      ++r1;
    }
    f(2 * p, int(r >> lg), int(r & (n - 1)), z, int(r1 >> lg), int(r1 & (n - 1)), z1);
    z += sz;
    r += sr;
    if (z >= nz) {
      z -= nz;
      ++r;
    }
  }
}

// Sum over stored modes of w * Re(conj(a) b).  The weights reproduce the
// reference's half-space loop: 2 for 0 < kz < N/2; on the kz = 0 and N/2
// planes 2 for 0 < kx < N/2, 0 for kx > N/2, and on the kx = 0, N/2 lines
// 2 / 1 / 0 for ky inside / on / above (0, N/2).
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_inner_partial(const cx<T> *__restrict__ a, const cx<T> *__restrict__ b,
                                                               int n, double *__restrict__ part) {
  int h = n / 2;
  double acc = 0;
  for_each_mode(n, [&](long i, int x, int y, int z) {
    double w;
    if (z > 0 && z < h) w = 2;
    else if (x > 0 && x < h) w = 2;
    else if (x > h) w = 0;
    else if (y > 0 && y < h) w = 2;
    else if (y > h) w = 0;
    else w = 1;
    if (w != 0) {
      cx<T> u = a[i], v = b[i];
      acc += w * (double(u.x) * double(v.x) + double(u.y) * double(v.y));
    }
  });
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double *__restrict__ part, int np,
                                                              double *__restrict__ out) {
  double acc = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) *out = s;
}

__device__ __forceinline__ double mode_weight(int x, int y, int z, int h) {
  if (z > 0 && z < h) return 2;
  if (x > 0 && x < h) return 2;
  if (x > h) return 0;
  if (y > 0 && y < h) return 2;
  if (y > h) return 0;
  return 1;
}
__device__ __forceinline__ long mode_ksq(int x, int y, int z, int n) {
  int h = n / 2;
  long kx = x > h ? x - n : x, ky = y > h ? y - n : y;
  return kx * kx + ky * ky + long(z) * z;
}

// The Gram entries of one block of up to 4 covectors a_p against up to 4
// vectors v_q and, with w given, <a_p, w>: sums over stored modes of
// weight * Re(conj(a) b), with every field read once per mode.  A block off
// the diagonal has rows 4 p + q for <a_p, v_q> and 16 + p for <a_p, w>; a
// diagonal block (DIAG) only q <= p, in rows p (p + 1) / 2 + q, and <a_p, w>
// in rows 10 + p (fewer accumulators, so that it needs no spill registers).
constexpr int GRAM_ROWS = 20;
struct GramBlock {
  const void *a[4], *v[4], *w;
  int na, nv;
};
template <bool DIAG> __host__ __device__ constexpr int gram_rows() { return DIAG ? 14 : 20; }
template <bool DIAG> __host__ __device__ constexpr int gram_row(int p, int q) {
  return DIAG ? p * (p + 1) / 2 + q : 4 * p + q;
}
template <bool DIAG> __host__ __device__ constexpr int gram_wrow(int p) { return DIAG ? 10 + p : 16 + p; }
template <typename T, bool DIAG>
__global__ __launch_bounds__(RED_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_gram_block_partial(
    GramBlock gb, int n, double *__restrict__ part) {
  constexpr int NR = gram_rows<DIAG>();
  int h = n / 2;
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = 0;
  const int na = gb.na, nv = gb.nv;
  const cx<T> *wf = reinterpret_cast<const cx<T> *>(gb.w);
  // the products of one mode: a, v, f its values in the fields
  auto mode = [&](double w, const cx<T> *a, const cx<T> *v, cx<T> f) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (p < na && q < nv && (!DIAG || q <= p))
          acc[gram_row<DIAG>(p, q)] += w * (double(a[p].x) * double(v[q].x) + double(a[p].y) * double(v[q].y));
      if (wf && p < na) acc[gram_wrow<DIAG>(p)] += w * (double(a[p].x) * double(f.x) + double(a[p].y) * double(f.y));
    }
  };
  if constexpr (sizeof(T) == 4) {
    // two modes per lane: one 16-byte load per field
    for_each_mode_pair(n, [&](long i, int x0, int y0, int z0, int x1, int y1, int z1) {
      float4 a2[4], v2[4], f2{0, 0, 0, 0};
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a2[p] = p < na ? reinterpret_cast<const float4 *>(gb.a[p])[i / 2] : float4{0, 0, 0, 0};
        v2[p] = p < nv ? reinterpret_cast<const float4 *>(gb.v[p])[i / 2] : float4{0, 0, 0, 0};
      }
      if (wf) f2 = reinterpret_cast<const float4 *>(wf)[i / 2];
      cx<T> a[4], v[4];
      double w0 = mode_weight(x0, y0, z0, h), w1 = mode_weight(x1, y1, z1, h);
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a[p] = cx<T>{a2[p].x, a2[p].y};
        v[p] = cx<T>{v2[p].x, v2[p].y};
      }
      if (w0 != 0) mode(w0, a, v, cx<T>{f2.x, f2.y});
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a[p] = cx<T>{a2[p].z, a2[p].w};
        v[p] = cx<T>{v2[p].z, v2[p].w};
      }
      if (w1 != 0) mode(w1, a, v, cx<T>{f2.z, f2.w});
    });
  } else {
    for_each_mode(n, [&](long i, int x, int y, int z) {
      double w = mode_weight(x, y, z, h);
      if (w == 0) return;
      cx<T> a[4], v[4], f{0, 0};
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        a[p] = p < na ? reinterpret_cast<const cx<T> *>(gb.a[p])[i] : cx<T>{0, 0};
        v[p] = p < nv ? reinterpret_cast<const cx<T> *>(gb.v[p])[i] : cx<T>{0, 0};
      }
      if (wf) f = wf[i];
      mode(w, a, v, f);
    });
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double s = block_sum(acc[r]);
    if (threadIdx.x == 0) part[size_t(r) * gridDim.x + blockIdx.x] = s;
    __syncthreads();
  }
}
// out[p] = sum of part[p*np_blocks ...], one block per pair
__global__ __launch_bounds__(RED_THREADS) void k_reduce_rows(const double *__restrict__ part, int nblk,
                                                             double *__restrict__ out) {
  double acc = 0;
  const double *row = part + size_t(blockIdx.x) * nblk;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc += row[i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

struct FilterSpec {
  int na = 0, nb = 0;
  int ka[6], kb[6];   // 0 low-pass Fermi, 1 sqrt(1-LP^2), 2 1-LP
  double ca[6], cb[6];
  double scale = 1.0;
  bool zero = false;
};
__device__ __forceinline__ double filter_factor(int kind, double kc, double k) {
  double lp = 1. / (1. + exp((k - kc) / (kc * 0.1)));
  if (kind == 0) return lp;
  if (kind == 1) return sqrt(1. - lp * lp);
  return 1. - lp;
}
// The filter spec at |k|^2 = ksq, rounded to T exactly as k_filter_spec
// tabulates it (the sums form it per mode rather than gather it from the
// table: a gather of 64 scattered entries per wavefront costs more than the
// arithmetic)
template <typename T> __device__ __forceinline__ T spec_value(const FilterSpec &s, double kmin, long ksq) {
  if (s.zero) return T(0);
  double k = sqrt(double(ksq)) * kmin;
  double fa = 1.0, fb = 1.0;
  for (int i = 0; i < s.na; ++i) fa *= filter_factor(s.ka[i], s.ca[i], k);
  for (int i = 0; i < s.nb; ++i) fb *= filter_factor(s.kb[i], s.cb[i], k);
  return T(fa * fb * s.scale);
}

// The chi^2 contribution of a level whose metric filters inside the window
// of the next level: with M the window mask, u = M x and F the level's
// filter, <x, B(A(x))> for A = 1 - M + F M, B = 1 - M + M F (two applications
// of Field::applyFilterInWindow, field.hpp:752-775) is
//   |x|^2 - |u|^2 + 2 F Re(u* x) - 2 F |u|^2 + F^2 |u|^2
// summed over modes, given the spectra of x and u.
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_chi2_window_partial(const cx<T> *__restrict__ xs,
                                                                     const cx<T> *__restrict__ us, FilterSpec F,
                                                                     double kmin, int n, double *__restrict__ part) {
  int h = n / 2;
  double acc = 0;
  auto term = [&](double w, double f, double ax, double ay, double bx, double by) {
    double xx = ax * ax + ay * ay;
    double uu = bx * bx + by * by;
    double ux = bx * ax + by * ay;
    return w * (xx - uu + 2 * f * ux - 2 * f * uu + f * f * uu);
  };
  if constexpr (sizeof(T) == 4) {
    // two modes per lane: one 16-byte load per spectrum
    for_each_mode_pair(n, [&](long i, int x0, int y0, int z0, int x1, int y1, int z1) {
      float4 a = reinterpret_cast<const float4 *>(xs)[i / 2], b = reinterpret_cast<const float4 *>(us)[i / 2];
      double w0 = mode_weight(x0, y0, z0, h), w1 = mode_weight(x1, y1, z1, h);
      if (w0 != 0) acc += term(w0, double(spec_value<T>(F, kmin, mode_ksq(x0, y0, z0, n))), a.x, a.y, b.x, b.y);
      if (w1 != 0) acc += term(w1, double(spec_value<T>(F, kmin, mode_ksq(x1, y1, z1, n))), a.z, a.w, b.z, b.w);
    });
  } else {
    for_each_mode(n, [&](long i, int x, int y, int z) {
      double w = mode_weight(x, y, z, h);
      if (w == 0) return;
      cx<T> a = xs[i], b = us[i];
      acc += term(w, double(spec_value<T>(F, kmin, mode_ksq(x, y, z, n))), a.x, a.y, b.x, b.y);
    });
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// sum over modes of w tab[|k|^2] Re(conj(a) b)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_inner_tab_partial(const cx<T> *__restrict__ a,
                                                                   const cx<T> *__restrict__ b,
                                                                   const T *__restrict__ tab, int n,
                                                                   double *__restrict__ part) {
  int h = n / 2;
  double acc = 0;
  for_each_mode(n, [&](long i, int x, int y, int z) {
    double w = mode_weight(x, y, z, h);
    if (w == 0) return;
    cx<T> u = a[i], v = b[i];
    acc += w * double(tab[mode_ksq(x, y, z, n)]) * (double(u.x) * v.x + double(u.y) * v.y);
  });
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// plain dot product of two dense real arrays
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_dot_partial(const T *__restrict__ a, const T *__restrict__ b, size_t n,
                                                             double *__restrict__ part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    acc += double(a[i]) * double(b[i]);
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// out = in * tab[|k|^2] over the stored modes
template <typename T>
__global__ void k_copy_mul_table(cx<T> *__restrict__ out, const cx<T> *__restrict__ in, int n,
                                 const T *__restrict__ tab) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int h = n / 2;
  size_t nm = size_t(n) * n * (h + 1);
  if (i >= nm) return;
  int x, y, z;
  mode_xyz(i, n, x, y, z);
  T v = tab[mode_ksq(x, y, z, n)];
  cx<T> a = in[i];
  out[i] = cx<T>{a.x * v, a.y * v};
}
// elementwise product of two tables
template <typename T> __global__ void k_tab_product(T *__restrict__ out, const T *__restrict__ a, long n) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] *= a[i];
}
// out[m] = -coef / k^2 (0 at k = 0), k^2 = m kmin^2: the velocity operator's factor
template <typename T> __global__ void k_tab_vel(T *__restrict__ out, double kmin, double coef, long n) {
  long m = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= n) return;
  out[m] = m == 0 ? T(0) : T(-coef / (double(m) * kmin * kmin)) * out[m];
}

// y += sum_j s_j x_j for up to 8 fields in one pass
struct AxpyList {
  const void *x[8];
  double s[8];
  int k;
};
template <typename T> __global__ void k_axpy_multi(T *__restrict__ y, AxpyList L, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T v = y[i];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (j < L.k) v += T(L.s[j]) * reinterpret_cast<const T *>(L.x[j])[i];
  y[i] = v;
}
// the same for FP32 with four values per lane (16-byte accesses); n4 = n / 4
__global__ void k_axpy_multi4(float *__restrict__ y, AxpyList L, size_t n4) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  float4 v = reinterpret_cast<float4 *>(y)[i];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (j < L.k) {
      float s = float(L.s[j]);
      float4 x = reinterpret_cast<const float4 *>(L.x[j])[i];
      v.x += s * x.x;
      v.y += s * x.y;
      v.z += s * x.z;
      v.w += s * x.w;
    }
  reinterpret_cast<float4 *>(y)[i] = v;
}
template <typename T> __global__ void k_axpy(T *__restrict__ y, const T *__restrict__ x, T s, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] += s * x[i];
}
template <typename T> __global__ void k_scale(T *__restrict__ x, T s, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s;
}

// product filters as tables over |k|^2 = m kMin^2: A(k) * B(k) * scale
template <typename T> __global__ void k_filter_spec(FilterSpec s, double kmin, long nent, T *__restrict__ out) {
  long m = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= nent) return;
  out[m] = spec_value<T>(s, kmin, m);
}

// pow(P_all, -1/2) * pow(P_dm, 1/2) from the two sqrt tables (0 where P = 0)
template <typename T>
__global__ void k_ratio_table(const T *__restrict__ all, const T *__restrict__ dm, long nent, T *__restrict__ out) {
  long m = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= nent) return;
  T a = all[m];
  out[m] = (a != T(0) ? T(1) / a : T(0)) * dm[m];
}

// densityToPotential (cosmology/parameters.hpp:209-225) -- zero where ky == 0,
// as the reference tests k2 (the y component) -- and the velocity covector
// -s delta / k^2 (linearmodification.hpp:223-263)
template <typename T> __global__ void k_mode_op(cx<T> *__restrict__ f, int n, double kmin, int kind, double coef) {
  int h = n / 2;
  size_t nm = size_t(n) * n * (h + 1);
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  int ix, iy, iz;
  mode_xyz(i, n, ix, iy, iz);
  if (ix > h) ix -= n;
  if (iy > h) iy -= n;
  T k1 = T(ix * kmin), k2 = T(iy * kmin), k3 = T(iz * kmin);
  T ksq = k1 * k1 + k2 * k2 + k3 * k3;
  cx<T> v = f[i];
  T c = T(coef);
  if (kind == 1) {
    T inv = T(1.0) / ksq;
    if (k2 == T(0)) inv = T(0);
    f[i] = cx<T>{(-c * v.x) * inv, (-c * v.y) * inv};
  } else {
    if (ksq == T(0)) f[i] = cx<T>{T(0), T(0)};
    else f[i] = cx<T>{(-c * v.x) / ksq, (-c * v.y) / ksq};
  }
}

// Window<T> of a grid in physical coordinates (simulation/window.hpp)
struct WinDev {
  double L, lo[3], hi[3];
};
__device__ __forceinline__ bool win_contains(const WinDev &w, double x, double y, double z) {
  double p[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    bool in = w.hi[d] >= w.lo[d] ? (p[d] < w.hi[d] && p[d] >= w.lo[d]) : (p[d] < w.hi[d] || p[d] >= w.lo[d]);
    if (!in) return false;
  }
  return true;
}
// cell centroid along one axis; the window tests of every kernel use this one
// expression so that they agree cell for cell
__device__ __forceinline__ double centroid1(int x, double dx, double o) { return x * dx + o + dx / 2; }
__device__ __forceinline__ bool win_axis(const WinDev &w, int d, double p) {
  return w.hi[d] >= w.lo[d] ? (p < w.hi[d] && p >= w.lo[d]) : (p < w.hi[d] || p >= w.lo[d]);
}
// per axis d, out[d*n + x] = 1 if cells with coordinate x pass the window's test on that axis
__global__ void k_axis_member(int n, double dx, double ox, double oy, double oz, WinDev w, uint8_t *__restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  int d = t / n, x = t - d * n;
  double o = d == 0 ? ox : d == 1 ? oy : oz;
  out[t] = win_axis(w, d, centroid1(x, dx, o)) ? 1 : 0;
}
// zero the cells (padded real layout) whose centroid is inside / outside w
template <typename T>
__global__ void k_window_zero(T *__restrict__ f, int n, double dx, double ox, double oy, double oz, WinDev w,
                              int inside) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  int z = int(i % n);
  size_t r = i / n;
  int y = int(r % n), x = int(r / n);
  bool c = win_contains(w, centroid1(x, dx, ox), centroid1(y, dx, oy), centroid1(z, dx, oz));
  if (c == (inside != 0)) f[(size_t(x) * n + y) * (n + 2) + z] = T(0);
}

// keep only the listed cells (windowOperator, quadraticmodification.hpp:180-186)
template <typename T> __global__ void k_mask_keep(T *__restrict__ f, int n, const uint8_t *__restrict__ mask) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  if (!mask[i]) {
    size_t x = i / (size_t(n) * n), y = (i / n) % n, z = i % n;
    f[(x * n + y) * (n + 2) + z] = T(0);
  }
}
__global__ void k_mask_set(uint8_t *__restrict__ mask, const uint64_t *__restrict__ idx, size_t m) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m) mask[idx[i]] = 1;
}

__device__ __forceinline__ size_t padded_of(uint64_t c, int n) {
  size_t x = c / (size_t(n) * n), y = (c / n) % n, z = c % n;
  return (x * n + y) * (n + 2) + z;
}

// varianceOperator (quadraticmodification.hpp:202-218): sum over the
// flagged cells, then x = (x - sum/m) / m on each of them
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_flag_sum(const T *__restrict__ f, int n, const uint64_t *__restrict__ idx,
                                                          size_t m, double *__restrict__ part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += size_t(gridDim.x) * blockDim.x)
    acc += double(f[padded_of(idx[i], n)]);
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <typename T>
__global__ void k_flag_center(T *__restrict__ f, int n, const uint64_t *__restrict__ idx, size_t m,
                              const double *__restrict__ sum) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T mean = T(*sum) / T(m);
  size_t p = padded_of(idx[i], n);
  T v = f[p] - mean;
  f[p] = v / T(m);
}

// fine -> coarse: the mean of the contained fine subcells of each coarse
// cell whose first or last subcell lies in the fine grid (SubSampleEvaluator
// over a SectionOfGrid of the fine grid)
struct AvgGeom {
  int r, cell_off[3], simf, nsec, nf, nc;
};
__device__ __forceinline__ bool sub_under(const AvgGeom &g, int sx, int sy, int sz, int *u) {
  if (sx < 0 || sy < 0 || sz < 0 || sx >= g.nsec || sy >= g.nsec || sz >= g.nsec) return false;
  int s[3] = {sx, sy, sz};
  for (int a = 0; a < 3; ++a) {
    int v = s[a] + g.cell_off[a];
    if (v > g.simf - 1) v -= g.simf;
    if (v < 0) v += g.simf;
    if (v < 0 || v >= g.nf) return false;
    u[a] = v;
  }
  return true;
}
template <typename T>
__global__ void k_fine_to_coarse(T *__restrict__ coarse, const T *__restrict__ fine, AvgGeom g) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nc3 = size_t(g.nc) * g.nc * g.nc;
  if (i >= nc3) return;
  int Z = int(i % g.nc);
  size_t rr = i / g.nc;
  int Y = int(rr % g.nc), X = int(rr / g.nc);
  int u[3];
  int r = g.r;
  bool c0 = sub_under(g, X * r, Y * r, Z * r, u);
  bool c1 = sub_under(g, X * r + r - 1, Y * r + r - 1, Z * r + r - 1, u);
  if (!c0 && !c1) return;
  T sum = T(0);
  int cnt = 0;
  for (int a = 0; a < r; ++a)
    for (int b = 0; b < r; ++b)
      for (int c = 0; c < r; ++c)
        if (sub_under(g, X * r + a, Y * r + b, Z * r + c, u)) {
          sum += fine[(size_t(u[0]) * g.nf + u[1]) * (g.nf + 2) + u[2]];
          ++cnt;
        }
  coarse[(size_t(X) * g.nc + Y) * (g.nc + 2) + Z] += sum / T(cnt);
}

// The same average for the coarse cells of a box [b0, b0+bl) (mod nc), into
// the dense box buffer (bl[0]*bl[1]*bl[2]); cells outside the box have no
// fine subcells.
struct Box3 {
  int b0[3], bl[3];
};
template <typename T>
__global__ void k_fine_to_coarse_box(T *__restrict__ box, const T *__restrict__ fine, AvgGeom g, Box3 B) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nb = size_t(B.bl[0]) * B.bl[1] * B.bl[2];
  if (i >= nb) return;
  int k = int(i % B.bl[2]);
  size_t rr = i / B.bl[2];
  int j = int(rr % B.bl[1]), l = int(rr / B.bl[1]);
  int X = (B.b0[0] + l) % g.nc, Y = (B.b0[1] + j) % g.nc, Z = (B.b0[2] + k) % g.nc;
  int u[3];
  int r = g.r;
  bool c0 = sub_under(g, X * r, Y * r, Z * r, u);
  bool c1 = sub_under(g, X * r + r - 1, Y * r + r - 1, Z * r + r - 1, u);
  if (!c0 && !c1) return;
  T sum = T(0);
  int cnt = 0;
  for (int a = 0; a < r; ++a)
    for (int b = 0; b < r; ++b)
      for (int c = 0; c < r; ++c)
        if (sub_under(g, X * r + a, Y * r + b, Z * r + c, u)) {
          sum += fine[(size_t(u[0]) * g.nf + u[1]) * (g.nf + 2) + u[2]];
          ++cnt;
        }
  box[i] += sum / T(cnt);
}

// Field::deInterpolate of every fine cell (value / r^3) into the coarse grid,
// accumulated in the reference's contiguous real layout (n^3 cells plus the
// tail the reference's vector has): coordinates wrap by the grid size only
// below 0 or above n, as there, so the same cells alias.
struct DeinterpGeom {
  int nf, nc, simc;
  double fdx, fo[3], cdx, co[3], periodic, inv_vol;
};
__device__ __forceinline__ double crw(int i, double t) {
  double t2 = t * t, t3 = t2 * t;
  if (i == 0) return -0.5 * t3 + t2 - 0.5 * t;
  if (i == 1) return 1.5 * t3 - 2.5 * t2 + 1.0;
  if (i == 2) return -1.5 * t3 + 2.0 * t2 + 0.5 * t;
  return 0.5 * t3 - 0.5 * t2;
}
template <typename T, typename S>
__global__ void k_deinterp(const S *__restrict__ fine, int fine_padded, DeinterpGeom g, T *__restrict__ acc,
                           size_t acc_len, size_t acc_lo, unsigned *__restrict__ dropped) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nf3 = size_t(g.nf) * g.nf * g.nf;
  if (i >= nf3) return;
  int z = int(i % g.nf);
  size_t rr = i / g.nf;
  int y = int(rr % g.nf), x = int(rr / g.nf);
  double v = fine_padded ? double(fine[(size_t(x) * g.nf + y) * (g.nf + 2) + z]) : double(fine[i]);
  if (v == 0) return;
  double val = v * g.inv_vol;
  int cf[3] = {x, y, z};
  int p0[3];
  double t[3];
  for (int a = 0; a < 3; ++a) {
    double loc = cf[a] * g.fdx + g.fo[a] + g.fdx / 2 - g.co[a];
    double w = fmod(loc, g.periodic);
    if (w < 0) w += g.periodic;
    double u = w / g.cdx - 0.5;
    p0[a] = int(floor(u));
    t[a] = u - p0[a];
  }
  long n = g.nc;
  int ix[3][4];
  for (int a = 0; a < 3; ++a)
    for (int q = 0; q < 4; ++q) {
      int c = p0[a] + q - 1;
      if (c < 0) c += int(n);
      if (c > n) c -= int(n);
      ix[a][q] = c;
    }
  long key[3];
  for (int a = 0; a < 3; ++a) {
    long c = ix[a][0];
    if (c > g.simc - 1) c -= g.simc;
    if (c < 0) c += g.simc;
    key[a] = c;
  }
  long kidx = (key[0] * n + key[1]) * n + key[2];
  double wx[4], wy[4], wz[4];
  for (int q = 0; q < 4; ++q) {
    wx[q] = crw(q, t[0]);
    wy[q] = crw(q, t[1]);
    wz[q] = crw(q, t[2]);
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        long idx = kidx + (ix[0][a] - ix[0][0]) * n * n + (ix[1][b] - ix[1][0]) * n + (ix[2][c] - ix[2][0]);
        // acc holds the contiguous entries [acc_lo, acc_lo + acc_len)
        long rel = idx - long(acc_lo);
        if (rel >= 0 && size_t(rel) < acc_len) atomicAdd(&acc[rel], T(val * (wx[a] * wy[b] * wz[c])));
        else if (dropped && idx >= 0 && size_t(idx) < size_t(g.nc) * g.nc * g.nc) atomicAdd(dropped, 1u);
      }
}


// localisation covectors of the finest level (linearmodification.hpp:60-70,
// 198-226, 330-370), one thread per flagged cell, summed in FP64
struct LocArgs {
  int kind, n, sim, dir;
  double w, a, b, dx, o[3], periodic, qc[3];
};
__device__ __forceinline__ long loc_step(long index, int n, int sim, int dir, int s) {
  long q[3] = {index / (long(n) * n), (index / n) % n, index % n};
  q[dir] += s;
  for (int a = 0; a < 3; ++a) {
    if (q[a] > sim - 1) q[a] -= sim;
    if (q[a] < 0) q[a] += sim;
  }
  return (q[0] * n + q[1]) * n + q[2];
}
__device__ __forceinline__ void loc_add(double *dense, long n3, long i, double v) {
  if (i >= 0 && i < n3) atomicAdd(&dense[i], v);   // steps leaving a zoom grid are dropped
}
__global__ void k_localise(const uint64_t *__restrict__ cells, size_t m, LocArgs A, double *__restrict__ dense) {
  size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= m) return;
  long c = long(cells[t]);
  long n3 = long(A.n) * A.n * A.n;
  if (A.kind <= 1) {
    loc_add(dense, n3, c, A.w);
  } else if (A.kind <= 4) {
    long m1 = loc_step(c, A.n, A.sim, A.dir, -1), p1 = loc_step(c, A.n, A.sim, A.dir, +1);
    long m2 = loc_step(m1, A.n, A.sim, A.dir, -1), p2 = loc_step(p1, A.n, A.sim, A.dir, +1);
    loc_add(dense, n3, m2, A.a);
    loc_add(dense, n3, m1, A.b);
    loc_add(dense, n3, p1, -A.b);
    loc_add(dense, n3, p2, -A.a);
  } else {
    int dir = A.dir, dp1 = (dir + 1) % 3, dp2 = (dir + 2) % 3;
    long q[3] = {c / (long(A.n) * A.n), (c / A.n) % A.n, c % A.n};
    double dq[3], cr[3];
    for (int d = 0; d < 3; ++d) {
      double p = double(q[d]) * A.dx + A.o[d] + A.dx / 2;
      double r = A.qc[d] - p;
      if (r > A.periodic / 2) r -= A.periodic;
      if (r < -A.periodic / 2) r += A.periodic;
      dq[d] = r;
    }
    cr[dir] = 0;
    cr[dp2] = dq[dp1];
    cr[dp1] = -dq[dp2];
    for (int d = 0; d < 3; ++d) {
      if (d == dir) continue;
      long m1 = loc_step(c, A.n, A.sim, d, -1), m2 = loc_step(m1, A.n, A.sim, d, -1);
      long p1 = loc_step(c, A.n, A.sim, d, +1), p2 = loc_step(p1, A.n, A.sim, d, +1);
      loc_add(dense, n3, m2, cr[d] * A.a);
      loc_add(dense, n3, m1, cr[d] * A.b);
      loc_add(dense, n3, p1, -cr[d] * A.b);
      loc_add(dense, n3, p2, -cr[d] * A.a);
    }
  }
}
template <typename T> __global__ void k_dense_to_padded(T *__restrict__ f, const double *__restrict__ d, int n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  size_t x = i / (size_t(n) * n), y = (i / n) % n, z = i % n;
  f[(x * n + y) * (n + 2) + z] = T(d[i]);
}

int blocks(size_t n) { return int((n + 255) / 256); }

}  // namespace

// ============================================================ multi-level fields ==
struct ModRec {
  std::string name;
  bool quadratic = false;
  double target = 0, scale = -1, precision = 0.001;
  int init_steps = 10;
  std::vector<std::vector<uint64_t>> flags;
};

// Device memory reused across the many temporary fields of the modification
// algebra: every buffer goes back to the free list when its field dies and
// is taken again by the next request of a size it fits (all work runs on
// the context's one stream, so reuse is ordered after the last use).
struct BufPool {
  std::multimap<size_t, void *> free_;
  std::map<void *, size_t> size_of;
  void *get(size_t b) {
    auto it = free_.lower_bound(b);
    if (it != free_.end() && it->first <= 2 * b) {
      void *p = it->second;
      free_.erase(it);
      return p;
    }
    void *p = nullptr;
    GIC_HIP(hipMalloc(&p, b));
    size_of[p] = b;
    return p;
  }
  void put(void *p) {
    if (p) free_.emplace(size_of.at(p), p);
  }
  ~BufPool() {
    for (auto &kv : size_of) (void)hipFree(kv.first);
  }
};

}  // namespace gic

struct ModState {
  std::vector<gic::ModRec> linear, quadratic;
  bool line_search = false;   // quadratic steps solved exactly along the gradient
  gic::DevBuf part, scalar;
  gic::BufPool pool;
};

gic_context::gic_context() = default;
gic_context::~gic_context() = default;

namespace gic {

namespace {

size_t table_len(const LevelGeom &g) {
  long h = g.n / 2;
  return size_t(3 * h * h + 1);
}

// A pool buffer that returns itself on destruction
struct PBuf {
  BufPool *pool = nullptr;
  void *p = nullptr;
  PBuf() = default;
  PBuf(BufPool &pl, size_t b) : pool(&pl), p(pl.get(b)) {}
  PBuf(PBuf &&o) noexcept : pool(o.pool), p(o.p) { o.p = nullptr; }
  PBuf &operator=(PBuf &&o) noexcept {
    if (this != &o) {
      if (pool) pool->put(p);
      pool = o.pool;
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  PBuf(const PBuf &) = delete;
  ~PBuf() {
    if (pool) pool->put(p);
  }
};

// A multi-level field: one padded buffer per level, allocated on first use
template <typename T> struct MLF {
  gic_context *c = nullptr;
  std::vector<PBuf> lv;
  std::vector<bool> fourier;
  MLF() = default;
  explicit MLF(gic_context *cc) : c(cc) {
    lv.resize(c->levels.size());
    fourier.assign(c->levels.size(), false);
  }
  MLF(MLF &&) noexcept = default;
  MLF &operator=(MLF &&) noexcept = default;
  size_t nl() const { return lv.size(); }
  const LevelGeom &g(size_t l) const { return c->levels[l]->g; }
  size_t reals(size_t l) const { return g(l).padded_reals(); }
  bool has(size_t l) const { return lv[l].p != nullptr; }
  T *p(size_t l) {
    if (!lv[l].p) lv[l] = PBuf(c->mod->pool, reals(l) * sizeof(T));
    return (T *)lv[l].p;
  }
  const T *p(size_t l) const { return (const T *)lv[l].p; }
  void copy_level(const MLF &o, size_t l) {
    Timed t(c, "copy_field", 2.0 * reals(l) * sizeof(T));
    GIC_HIP(hipMemcpyAsync(p(l), o.p(l), reals(l) * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    fourier[l] = o.fourier[l];
  }
  MLF clone() const {
    MLF m(c);
    for (size_t l = 0; l < nl(); ++l)
      if (has(l)) m.copy_level(*this, l);
    return m;
  }
  void zero(size_t l) {
    fill<T>(p(l), reals(l), T(0), c->st);
    fourier[l] = false;
  }
  void zero() {
    for (size_t l = 0; l < nl(); ++l) zero(l);
  }
  void toF(size_t l) {
    if (fourier[l]) return;
    fft3d<T>(p(l), g(l).n, -1, c->st, c->hook());
    fourier[l] = true;
  }
  void toR(size_t l) {
    if (!fourier[l]) return;
    fft3d<T>(p(l), g(l).n, +1, c->st, c->hook());
    fourier[l] = false;
  }
  void toF() {
    for (size_t l = 0; l < nl(); ++l) toF(l);
  }
  void axpy_level(size_t l, const T *x, double s) {
    size_t n = reals(l);
    hipLaunchKernelGGL(k_axpy<T>, dim3(blocks(n)), dim3(256), 0, c->st, p(l), x, T(s), n);
  }
  void axpy(const MLF &x, double s) {   // this += s x, same spaces
    for (size_t l = 0; l < nl(); ++l) axpy_level(l, x.p(l), s);
  }
  void scale(double s) {
    for (size_t l = 0; l < nl(); ++l) {
      size_t n = reals(l);
      hipLaunchKernelGGL(k_scale<T>, dim3(blocks(n)), dim3(256), 0, c->st, p(l), T(s), n);
    }
  }
};

// The output field's own buffers (the noise of every level) as a multi-level
// field, for work that reads it or updates it in place (chi^2, calculate, the
// modifications themselves) without the copies in and out of the pool.  The
// levels' Fourier flags are written back when the view goes out of scope.
template <typename T> struct OutputView {
  gic_context *c;
  MLF<T> m;
  explicit OutputView(gic_context *cc) : c(cc), m(cc) {
    for (size_t l = 0; l < m.nl(); ++l) {
      Level &L = *c->levels[l];
      if (L.dc_pending) ensure_noise_fourier<T>(c, L);
      m.lv[l].p = L.noise.p;   // no pool: the buffer stays the level's
      m.fourier[l] = L.noise_fourier;
    }
  }
  OutputView(const OutputView &) = delete;
  ~OutputView() {
    for (size_t l = 0; l < m.nl(); ++l) {
      m.lv[l].p = nullptr;
      c->levels[l]->noise_fourier = m.fourier[l];
    }
  }
};

bool combined(gic_context *c) { return c->zeldovich_done && c->levels.size() >= 2; }

// filter of level l (FilterFamily): prod_{j<l} HP(kc_j) * LP(kc_l) (not on the last)
void level_factors(gic_context *c, size_t l, int &nf, int *kind, double *kc, bool &zero) {
  size_t nl = c->levels.size();
  nf = 0;
  zero = false;
  if (combined(c)) {
    zero = l + 1 < nl;
    return;
  }
  for (size_t j = 0; j < l; ++j) {
    kind[nf] = 1;
    kc[nf++] = kcut_for_level(c, int(j));
  }
  if (l + 1 < nl) {
    kind[nf] = 0;
    kc[nf++] = kcut_for_level(c, int(l));
  }
}

FilterSpec spec_levels(gic_context *c, long la, long lb, double scale) {
  FilterSpec s;
  bool za = false, zb = false;
  if (la >= 0) level_factors(c, size_t(la), s.na, s.ka, s.ca, za);
  if (lb >= 0) level_factors(c, size_t(lb), s.nb, s.kb, s.cb, zb);
  s.zero = za || zb;
  s.scale = scale;
  return s;
}

// the table of a filter spec on the k of a level (in c->tab_a)
template <typename T> const T *spec_table(gic_context *c, const LevelGeom &g, const FilterSpec &s) {
  size_t nent = table_len(g);
  c->tab_a.ensure(nent * sizeof(T));
  hipLaunchKernelGGL(k_filter_spec<T>, dim3(blocks(nent)), dim3(256), 0, c->st, s, g.kmin(), long(nent),
                     (T *)c->tab_a.p);
  return (const T *)c->tab_a.p;
}

// the same product for the transforms to form in their passes (FP32)
FilterInline inline_of(const FilterSpec &s) {
  FilterInline f;
  if (s.na + s.nb > FILTER_INLINE_MAX) return f;   // nf = -1: use the table
  f.nf = 0;
  for (int i = 0; i < s.na; ++i) {
    f.kind[f.nf] = s.ka[i];
    f.kc[f.nf++] = s.ca[i];
  }
  for (int i = 0; i < s.nb; ++i) {
    f.kind[f.nf] = s.kb[i];
    f.kc[f.nf++] = s.cb[i];
  }
  f.scale = s.scale;
  f.zero = s.zero ? 1 : 0;
  return f;
}

// multiply level l of m (Fourier) by the filter spec evaluated on that level's k
template <typename T> void apply_spec(MLF<T> &m, size_t l, const FilterSpec &s) {
  const T *tab = spec_table<T>(m.c, m.g(l), s);
  m.toF(l);
  mul_table<T>(m.p(l), m.g(l).n, tab, m.c->st);
}

// sum_l sum_k w Re(conj(a) b), both in Fourier space
template <typename T> double inner(gic_context *c, MLF<T> &a, MLF<T> &b) {
  a.toF();
  b.toF();
  c->mod->part.ensure(RED_BLOCKS * sizeof(double));
  c->mod->scalar.ensure(sizeof(double) * 8);
  size_t nl = a.nl();
  for (size_t l = 0; l < nl; ++l) {
    Timed t(c, "inner", 2.0 * a.g(l).nmodes() * sizeof(cx<T>));
    hipLaunchKernelGGL(k_inner_partial<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st,
                       reinterpret_cast<const cx<T> *>(a.p(l)), reinterpret_cast<const cx<T> *>(b.p(l)), a.g(l).n,
                       (double *)c->mod->part.p);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, c->st, (const double *)c->mod->part.p,
                       RED_BLOCKS, (double *)c->mod->scalar.p + l);
  }
  double v[8];
  GIC_HIP(hipMemcpyAsync(v, c->mod->scalar.p, nl * sizeof(double), hipMemcpyDeviceToHost, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
  double total = 0;
  for (size_t l = 0; l < nl; ++l) total += v[l];
  return total;
}

// FP64 reductions: launch a partial kernel into part, finish in scalar[slot]
double finish_reduction(gic_context *c, int slot, int nblk = RED_BLOCKS) {
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, c->st, (const double *)c->mod->part.p, nblk,
                     (double *)c->mod->scalar.p + slot);
  GIC_HIP(hipGetLastError());
  double v;
  GIC_HIP(hipMemcpyAsync(&v, (double *)c->mod->scalar.p + slot, sizeof(double), hipMemcpyDeviceToHost, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
  return v;
}
void ensure_red(gic_context *c) {
  c->mod->part.ensure(size_t(RED_BLOCKS) * GRAM_ROWS * sizeof(double));
  c->mod->scalar.ensure(sizeof(double) * GRAM_ROWS);
}

// a cyclic run [start, start+len) (mod n) of the indices where member[] is 1;
// false if the members are not one run
bool cyclic_run(const uint8_t *member, int n, int &start, int &len) {
  int cnt = 0;
  for (int i = 0; i < n; ++i) cnt += member[i];
  if (cnt == 0) return false;
  if (cnt == n) {
    start = 0;
    len = n;
    return true;
  }
  int s = -1;
  for (int i = 0; i < n; ++i)
    if (member[i] && !member[(i + n - 1) % n]) {
      if (s >= 0) return false;   // two runs
      s = i;
    }
  start = s;
  len = cnt;
  for (int k = 0; k < len; ++k)
    if (!member[(s + k) % n]) return false;
  return true;
}

// The box of cells of level g whose centroids lie in window w (the cells
// k_window_zero treats as inside), per axis a cyclic run; false if empty or
// not a single run on some axis.
bool window_box(gic_context *c, const LevelGeom &g, const WinDev &w, Box3 &B) {
  PBuf mem(c->mod->pool, size_t(3) * g.n);
  hipLaunchKernelGGL(k_axis_member, dim3((3 * g.n + 255) / 256), dim3(256), 0, c->st, g.n, g.dx, g.ox, g.oy, g.oz, w,
                     (uint8_t *)mem.p);
  GIC_HIP(hipGetLastError());
  std::vector<uint8_t> h(size_t(3) * g.n);
  GIC_HIP(hipMemcpyAsync(h.data(), mem.p, h.size(), hipMemcpyDeviceToHost, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
  for (int d = 0; d < 3; ++d)
    if (!cyclic_run(h.data() + size_t(d) * g.n, g.n, B.b0[d], B.bl[d])) return false;
  return true;
}

AvgGeom avg_geom(const LevelGeom &gt, const LevelGeom &gs) {
  AvgGeom a;
  a.r = (int)llround(gt.dx / gs.dx);
  a.cell_off[0] = (int)llround((gt.ox - gs.ox) / gs.dx);
  a.cell_off[1] = (int)llround((gt.oy - gs.oy) / gs.dx);
  a.cell_off[2] = (int)llround((gt.oz - gs.oz) / gs.dx);
  a.simf = (int)llround(gs.periodic / gs.dx);
  a.nsec = gt.n * a.r;
  a.nf = gs.n;
  a.nc = gt.n;
  return a;
}
// coarse cells of level gt with subcells on the finer level gs (the cells
// k_fine_to_coarse writes), per axis; the same integer tests on the host
bool fine_footprint(const LevelGeom &gt, const LevelGeom &gs, std::vector<uint8_t> &member) {
  AvgGeom a = avg_geom(gt, gs);
  member.assign(size_t(3) * gt.n, 0);
  for (int d = 0; d < 3; ++d)
    for (int X = 0; X < gt.n; ++X)
      for (int q = 0; q < a.r; ++q) {
        int sidx = X * a.r + q;
        if (sidx < 0 || sidx >= a.nsec) continue;
        int v = sidx + a.cell_off[d];
        if (v > a.simf - 1) v -= a.simf;
        if (v < 0) v += a.simf;
        if (v >= 0 && v < a.nf) {
          member[size_t(d) * gt.n + X] = 1;
          break;
        }
      }
  return true;
}
// the box of level lt receiving averages from every finer level
bool fine_box(gic_context *c, size_t lt, Box3 &B) {
  const LevelGeom &gt = c->levels[lt]->g;
  std::vector<uint8_t> un(size_t(3) * gt.n, 0), m;
  bool any = false;
  for (size_t s = lt + 1; s < c->levels.size(); ++s) {
    fine_footprint(gt, c->levels[s]->g, m);
    for (size_t i = 0; i < un.size(); ++i) un[i] |= m[i];
    any = true;
  }
  if (!any) return false;
  for (int d = 0; d < 3; ++d)
    if (!cyclic_run(un.data() + size_t(d) * gt.n, gt.n, B.b0[d], B.bl[d])) return false;
  return true;
}
size_t box_cells(const Box3 &B) { return size_t(B.bl[0]) * B.bl[1] * B.bl[2]; }
bool box_pruned(const Box3 &B, int n) { return B.bl[0] < n || B.bl[1] < n || B.bl[2] < n; }

WinDev grid_window(const LevelGeom &g) {   // Grid::getWindow (grid.hpp:502-504)
  WinDev w;
  w.L = g.periodic;
  double lo[3] = {g.ox, g.oy, g.oz};
  auto wrap = [&](double v) {
    if (v < 0) v += w.L;
    if (v >= w.L) v -= w.L;
    return v;
  };
  for (int d = 0; d < 3; ++d) {
    double a = lo[d], b = lo[d] + g.box;
    w.lo[d] = wrap(a);
    w.hi[d] = std::nextafter(wrap(std::nextafter(b, -INFINITY)), INFINITY);
  }
  return w;
}

template <typename T> void window_zero(MLF<T> &m, size_t l, const WinDev &w, bool inside) {
  m.toR(l);
  const LevelGeom &g = m.g(l);
  size_t n3 = g.n3();
  hipLaunchKernelGGL(k_window_zero<T>, dim3(blocks(n3)), dim3(256), 0, m.c->st, m.p(l), g.n, g.dx, g.ox, g.oy, g.oz,
                     w, inside ? 1 : 0);
}

// applyFilterInWindow (field.hpp:752-775) on level l of m
template <typename T> void filter_in_window(MLF<T> &m, size_t l, const FilterSpec &s, const WinDev &w, bool first) {
  MLF<T> in(m.c);
  in.copy_level(m, l);
  if (first) {
    window_zero(in, l, w, false);
    apply_spec(in, l, s);
  } else {
    apply_spec(in, l, s);
    window_zero(in, l, w, false);
  }
  window_zero(m, l, w, true);
  if (in.fourier[l]) m.toF(l);
  else m.toR(l);
  m.axpy_level(l, in.p(l), 1.0);
}

double level_weight(gic_context *c, size_t l) {
  double r = c->levels[l]->g.dx / c->levels[0]->g.dx;
  return l == 0 ? 1.0 : pow(r, 3.0);
}

// R (real, level lt) += the level-ls field f transferred with filter spec s
// (addFieldFromDifferentGridWithFilter, FILTER_ON_COARSE_GRID): from a
// coarser level by tricubic interpolation -- through a pruned inverse
// transform of the coarse window when the coarse level covers the box -- and
// from a finer level by averaging its cells.
template <typename T> void add_transfer(gic_context *c, T *R, size_t lt, MLF<T> &f, size_t ls, const FilterSpec &s) {
  const LevelGeom &gt = c->levels[lt]->g, &gs = c->levels[ls]->g;
  f.toF(ls);
  if (gs.dx > gt.dx) {
    int w0[3], wl[3];
    coarse_window(gt, gs, w0, wl);
    const T *tab = spec_table<T>(c, gs, s);
    FilterInline fil = inline_of(s);
    CoarseView v;
    PBuf scratch;
    if (gs.covers_full() && wl[0] < gs.n && wl[1] < gs.n && wl[2] < gs.n) {
      int nz = gs.n / 2 + 1;
      size_t a_b = size_t(wl[0]) * gs.n * nz * 2 * sizeof(T);
      size_t b_b = size_t(wl[0]) * wl[1] * nz * 2 * sizeof(T);
      size_t o_b = size_t(wl[0]) * wl[1] * wl[2] * sizeof(T);
      scratch = PBuf(c->mod->pool, a_b + b_b + o_b);
      T *bufA = (T *)scratch.p;
      T *bufB = (T *)((char *)scratch.p + a_b);
      T *out = (T *)((char *)scratch.p + a_b + b_b);
      fft_c2r_window<T>(f.p(ls), gs.n, FFT_PRE_TABLE, tab, 0, gs.kmin(), w0, wl, bufA, bufB, out, c->st, c->hook(),
                        &fil);
      v.p = out;
      for (int a = 0; a < 3; ++a) {
        v.org[a] = w0[a];
        v.len[a] = wl[a];
      }
      v.sx = long(wl[1]) * wl[2];
      v.sy = wl[2];
    } else {
      scratch = PBuf(c->mod->pool, gs.padded_reals() * sizeof(T));
      fft_c2r_from<T>(f.p(ls), (T *)scratch.p, gs.n, FFT_PRE_TABLE, tab, 0, gs.kmin(), c->st, c->hook());
      v.p = scratch.p;
      for (int a = 0; a < 3; ++a) {
        v.org[a] = 0;
        v.len[a] = gs.n;
      }
      v.sx = long(gs.n) * (gs.n + 2);
      v.sy = gs.n + 2;
    }
    interp_add<T>(R, gt.n, v, interp_geom(gt, gs), c->st);
  } else {
    const T *tab = spec_table<T>(c, gs, s);
    PBuf tmp(c->mod->pool, gs.padded_reals() * sizeof(T));
    fft_c2r_from<T>(f.p(ls), (T *)tmp.p, gs.n, FFT_PRE_TABLE, tab, 0, gs.kmin(), c->st, c->hook());
    AvgGeom a;
    a.r = (int)llround(gt.dx / gs.dx);
    a.cell_off[0] = (int)llround((gt.ox - gs.ox) / gs.dx);
    a.cell_off[1] = (int)llround((gt.oy - gs.oy) / gs.dx);
    a.cell_off[2] = (int)llround((gt.oz - gs.oz) / gs.dx);
    a.simf = (int)llround(gs.periodic / gs.dx);
    a.nsec = gt.n * a.r;
    a.nf = gs.n;
    a.nc = gt.n;
    size_t nc3 = gt.n3();
    hipLaunchKernelGGL(k_fine_to_coarse<T>, dim3(blocks(nc3)), dim3(256), 0, c->st, R, (const T *)tmp.p, a);
  }
  GIC_HIP(hipGetLastError());
}

// The cross-level terms of level l for a field f (Fourier on every level):
// sum over s != l of the level-s field filtered by F_l F_s sqrt(w_l/w_s) and
// carried to level l (FILTER_ON_COARSE_GRID).  When every other level is
// finer and level l covers the box, they only reach the box of cells under
// the finer grids: returned there as a dense box (and true); otherwise in R
// as a full padded real field of level l.
template <typename T> bool cross_terms(gic_context *c, MLF<T> &f, size_t l, PBuf &Rbox, Box3 &B, MLF<T> &R) {
  size_t nl = f.nl();
  const LevelGeom &g = c->levels[l]->g;
  bool box = l == 0 && g.covers_full() && fine_box(c, l, B) && box_pruned(B, g.n);
  if (box) {
    size_t nb = box_cells(B);
    Rbox = PBuf(c->mod->pool, nb * sizeof(T));
    fill<T>((T *)Rbox.p, nb, T(0), c->st);
    for (size_t s = l + 1; s < nl; ++s) {
      const LevelGeom &gs = c->levels[s]->g;
      double sq = sqrt(level_weight(c, l) / level_weight(c, s));
      f.toF(s);
      const T *tab = spec_table<T>(c, gs, spec_levels(c, long(l), long(s), sq));
      PBuf tmp(c->mod->pool, gs.padded_reals() * sizeof(T));
      fft_c2r_from<T>(f.p(s), (T *)tmp.p, gs.n, FFT_PRE_TABLE, tab, 0, gs.kmin(), c->st, c->hook());
      hipLaunchKernelGGL(k_fine_to_coarse_box<T>, dim3(blocks(nb)), dim3(256), 0, c->st, (T *)Rbox.p,
                         (const T *)tmp.p, avg_geom(g, gs), B);
      GIC_HIP(hipGetLastError());
    }
    return true;
  }
  R.zero(l);
  for (size_t s = 0; s < nl; ++s) {
    if (s == l) continue;
    double sq = sqrt(level_weight(c, l) / level_weight(c, s));
    add_transfer<T>(c, R.p(l), l, f, s, spec_levels(c, long(l), long(s), sq));
  }
  return false;
}

// scratch for the pruned transforms of a box on a level of size n
template <typename T> struct BoxBufs {
  PBuf a, b;
  BoxBufs(gic_context *c, int n, const Box3 &B) {
    size_t nz = size_t(n / 2 + 1);
    a = PBuf(c->mod->pool, size_t(B.bl[0]) * n * nz * 2 * sizeof(T));
    b = PBuf(c->mod->pool, size_t(B.bl[0]) * std::max(B.bl[1], n) * nz * 2 * sizeof(T));
  }
};

// applyMetric (multilevelfield.hpp:335-384); the cross-level terms of a
// level are gathered in one real buffer (or box) and added in Fourier space
template <typename T> MLF<T> metric(MLF<T> &f, bool to_covector) {
  gic_context *c = f.c;
  size_t nl = f.nl();
  f.toF();
  MLF<T> out(c);
  for (size_t l = 0; l < nl; ++l) {
    const LevelGeom &g = c->levels[l]->g;
    if (to_covector && l + 1 < nl) {
      out.copy_level(f, l);
      WinDev w = grid_window(c->levels[l + 1]->g);
      FilterSpec fl = spec_levels(c, long(l), -1, 1.0);
      filter_in_window(out, l, fl, w, true);
      filter_in_window(out, l, fl, w, false);
      out.toF(l);
    }
    bool self_done = to_covector && l + 1 < nl;
    auto self_term = [&]() {   // out = F_l^2 f in one pass
      const T *tab = spec_table<T>(c, g, spec_levels(c, long(l), long(l), 1.0));
      size_t nm = size_t(g.n) * g.n * (g.n / 2 + 1);
      hipLaunchKernelGGL(k_copy_mul_table<T>, dim3(blocks(nm)), dim3(256), 0, c->st, reinterpret_cast<cx<T> *>(out.p(l)),
                         reinterpret_cast<const cx<T> *>(f.p(l)), g.n, tab);
      out.fourier[l] = true;
    };
    if (nl == 1 && !self_done) self_term();
    if (nl > 1) {
      PBuf Rbox;
      Box3 B;
      MLF<T> R(c);
      if (cross_terms<T>(c, f, l, Rbox, B, R)) {
        BoxBufs<T> bb(c, g.n, B);
        if (self_done) {
          fft_r2c_window<T>((const T *)Rbox.p, g.n, B.b0, B.bl, (T *)bb.b.p, (T *)bb.a.p, out.p(l), nullptr, true,
                            c->st, c->hook());
        } else {
          // out = F_l^2 f + r2c(box), the self term formed by the transform's last pass
          FilterSpec f2 = spec_levels(c, long(l), long(l), 1.0);
          const T *tab = spec_table<T>(c, g, f2);
          FilterInline fil = inline_of(f2);
          fft_r2c_window<T>((const T *)Rbox.p, g.n, B.b0, B.bl, (T *)bb.b.p, (T *)bb.a.p, out.p(l), tab, false,
                            c->st, c->hook(), f.p(l), &fil, g.kmin());
          out.fourier[l] = true;
        }
      } else {
        if (!self_done) self_term();
        R.toF(l);
        out.axpy_level(l, R.p(l), 1.0);
      }
    }
  }
  return out;
}

// applyTransferRatio(field species -> dm): white noise gets sqrt P_dm; an
// output field (species 'all') gets sqrt(P_dm / P_all)
template <typename T> void transfer_to_dm(MLF<T> &m) {
  gic_context *c = m.c;
  for (size_t l = 0; l < m.nl(); ++l) {
    Level &L = *c->levels[l];
    build_sqrtp_dm<T>(c, L);
    m.toF(l);
    if (!c->power_applied) {
      mul_table<T>(m.p(l), L.g.n, (const T *)L.sqrtp_dm.p, c->st);
    } else {
      build_sqrtp<T>(c, L);
      size_t nent = table_len(L.g);
      c->tab_b.ensure(nent * sizeof(T));
      hipLaunchKernelGGL(k_ratio_table<T>, dim3(blocks(nent)), dim3(256), 0, c->st, (const T *)L.sqrtp.p,
                         (const T *)L.sqrtp_dm.p, long(nent), (T *)c->tab_b.p);
      mul_table<T>(m.p(l), L.g.n, (const T *)c->tab_b.p, c->st);
    }
  }
}

struct DevList {
  PBuf idx;
  size_t n = 0;
};
void upload_list(gic_context *c, const std::vector<uint64_t> &idx, DevList &out) {
  out.n = idx.size();
  if (!out.n) return;
  out.idx = PBuf(c->mod->pool, out.n * 8);
  GIC_HIP(hipMemcpyAsync(out.idx.p, idx.data(), out.n * 8, hipMemcpyHostToDevice, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
}

int mod_kind(const std::string &name) {
  static const char *names[] = {"overdensity", "potential", "vx", "vy", "vz", "Lx", "Ly", "Lz"};
  for (int i = 0; i < 8; ++i)
    if (!strcasecmp(name.c_str(), names[i])) return i;
  return -1;
}

// LinearModification::getCovector with generateMultilevelCovectorFromHiresCovector
template <typename T> MLF<T> linear_covector(gic_context *c, const std::string &name,
                                             const std::vector<std::vector<uint64_t>> &flags) {
  int kind = mod_kind(name);
  if (kind < 0) throw std::runtime_error(name + " is an unknown modification name");
  size_t nl = c->levels.size(), L = nl - 1;
  if (flags.size() < nl || flags[L].empty())
    throw std::runtime_error(
        "Linear modifications use exclusively information from the highest resolution grid but no cells have been "
        "flagged on this grid.\nChange flagged cells selection to use linear modifications.");
  if (kind >= 2 && kind <= 4 && combined(c))
    throw std::runtime_error(
        "Cannot create velocity covector for recombined fields (i.e. after output generation is complete)");
  const LevelGeom &g = c->levels[L]->g;
  const std::vector<uint64_t> &cells = flags[L];
  LocArgs A{};
  A.kind = kind;
  A.n = g.n;
  A.sim = (int)llround(g.periodic / g.dx);
  A.dx = g.dx;
  A.o[0] = g.ox;
  A.o[1] = g.oy;
  A.o[2] = g.oz;
  A.periodic = g.periodic;
  A.w = 1.0 / double(cells.size());
  if (kind >= 2 && kind <= 4) {
    A.dir = kind - 2;
    A.a = -A.w / 12. / g.dx;
    A.b = A.w * 2. / 3. / g.dx;
  } else if (kind >= 5) {
    A.dir = kind - 5;
    // getCentreWrapped of the flagged cells (grid.hpp:466-494)
    auto cen = [&](uint64_t i, int d) {
      uint64_t q = d == 0 ? i / (uint64_t(g.n) * g.n) : d == 1 ? (i / g.n) % g.n : i % g.n;
      return double(q) * g.dx + (d == 0 ? g.ox : d == 1 ? g.oy : g.oz) + g.dx / 2;
    };
    for (int d = 0; d < 3; ++d) {
      double p0 = cen(cells[0], d), s = 0;
      for (size_t i = 1; i < cells.size(); ++i) {
        double r = cen(cells[i], d) - p0;
        if (r > g.periodic / 2) r -= g.periodic;
        if (r < -g.periodic / 2) r += g.periodic;
        s += r;
      }
      double v = fmod(s / double(cells.size()) + p0, g.periodic);
      if (v < 0) v += g.periodic;
      A.qc[d] = v;
    }
    double norm = double(cells.size()) * g.dx;
    A.a = 1. / 12. / norm;
    A.b = -2. / 3. / norm;
  }
  MLF<T> m(c);
  {
    DevList dl;
    upload_list(c, cells, dl);
    PBuf dense(c->mod->pool, g.n3() * sizeof(double));
    fill<double>((double *)dense.p, g.n3(), 0.0, c->st);
    hipLaunchKernelGGL(k_localise, dim3(blocks(dl.n)), dim3(256), 0, c->st, (const uint64_t *)dl.idx.p, dl.n, A,
                       (double *)dense.p);
    hipLaunchKernelGGL(k_dense_to_padded<T>, dim3(blocks(g.n3())), dim3(256), 0, c->st, m.p(L),
                       (const double *)dense.p, g.n);
    GIC_HIP(hipStreamSynchronize(c->st));   // the cell list is released here
  }
  // the factors every level gets after deinterpolation, as one |k|^2 table:
  // weight w_l / w_L, the level filter, the transfer to dark matter and, for
  // velocities and angular momenta, -coef / k^2
  const Cosmo &cc = c->cosmo;
  double coef = kind == 1 ? 3. / 2. * cc.Om / cc.a * 100. * 100. / (299792.) / (299792.)
                          : 1.0 * 100. * sqrt(cc.Om / (cc.a * cc.a * cc.a) + cc.Ol) * sqrt(cc.a);
  auto level_table = [&](size_t l, PBuf &tab) {
    const LevelGeom &gl = m.g(l);
    size_t nent = table_len(gl);
    double sw = level_weight(c, l) / level_weight(c, nl - 1);
    FilterSpec fs = spec_levels(c, long(l), -1, sw);
    tab = PBuf(c->mod->pool, nent * sizeof(T));
    hipLaunchKernelGGL(k_filter_spec<T>, dim3(blocks(nent)), dim3(256), 0, c->st, fs, gl.kmin(), long(nent),
                       (T *)tab.p);
    Level &Lv = *c->levels[l];
    build_sqrtp_dm<T>(c, Lv);
    if (!c->power_applied) {
      hipLaunchKernelGGL(k_tab_product<T>, dim3(blocks(nent)), dim3(256), 0, c->st, (T *)tab.p,
                         (const T *)Lv.sqrtp_dm.p, long(nent));
    } else {
      build_sqrtp<T>(c, Lv);
      PBuf ratio(c->mod->pool, nent * sizeof(T));
      hipLaunchKernelGGL(k_ratio_table<T>, dim3(blocks(nent)), dim3(256), 0, c->st, (const T *)Lv.sqrtp.p,
                         (const T *)Lv.sqrtp_dm.p, long(nent), (T *)ratio.p);
      hipLaunchKernelGGL(k_tab_product<T>, dim3(blocks(nent)), dim3(256), 0, c->st, (T *)tab.p, (const T *)ratio.p,
                         long(nent));
    }
    if (kind >= 2)
      hipLaunchKernelGGL(k_tab_vel<T>, dim3(blocks(nent)), dim3(256), 0, c->st, (T *)tab.p, gl.kmin(), coef,
                         long(nent));
    GIC_HIP(hipGetLastError());
  };
  std::vector<bool> done(nl, false);
  // deinterpolate down the levels, one level at a time, in the contiguous
  // layout.  The scatter accumulates in FP64 for either precision and is
  // rounded to T once: the 64-tap sums of the velocity and angular-momentum
  // stencils cancel heavily, and FP32 atomic sums would leave the coarse
  // levels' covector both imprecise and dependent on the order of the adds,
  // so a covector rebuilt later (calculate) would differ from the one the
  // modification solved with.
  if (L > 0) {
    PBuf accs[2];
    const T *src_t = m.p(L);        // the finest level: padded, type T
    const double *src_d = nullptr;  // coarser levels: contiguous FP64 sums
    int src_padded = 1;
    auto deinterp = [&](const DeinterpGeom &dg, size_t nf3, double *acc, size_t len, size_t lo, unsigned *drop) {
      if (src_d)
        hipLaunchKernelGGL((k_deinterp<double, double>), dim3(blocks(nf3)), dim3(256), 0, c->st, src_d, src_padded,
                           dg, acc, len, lo, drop);
      else
        hipLaunchKernelGGL((k_deinterp<double, T>), dim3(blocks(nf3)), dim3(256), 0, c->st, src_t, src_padded, dg,
                           acc, len, lo, drop);
      GIC_HIP(hipGetLastError());
    };
    PBuf dropped(c->mod->pool, sizeof(unsigned));
    for (long l = long(L) - 1; l >= 0; --l) {
      const LevelGeom &gh = m.g(size_t(l) + 1), &gl = m.g(size_t(l));
      DeinterpGeom dg;
      dg.nf = gh.n;
      dg.nc = gl.n;
      dg.simc = (int)llround(gl.periodic / gl.dx);
      dg.fdx = gh.dx;
      dg.fo[0] = gh.ox;
      dg.fo[1] = gh.oy;
      dg.fo[2] = gh.oz;
      dg.cdx = gl.dx;
      dg.co[0] = gl.ox;
      dg.co[1] = gl.oy;
      dg.co[2] = gl.oz;
      dg.periodic = gl.periodic;
      double r = double(llround(gl.dx / gh.dx));
      dg.inv_vol = 1.0 / (r * r * r);
      size_t nf3 = gh.n3();
      size_t nn = size_t(gl.n) * gl.n;
      if (l == 0 && gl.covers_full()) {
        // the base level: the finer grid's footprint (with the stencil's
        // reach) is a range of x planes; deinterpolate into those planes
        // only and transform them with a pruned r2c that also applies the
        // level's factors.  Any value outside the range sends us back to
        // the whole-level path.
        long x_lo = long(std::floor((gh.ox - gl.ox) / gl.dx)) - 3;
        long x_hi = long(std::ceil((gh.ox + gh.n * gh.dx - gl.ox) / gl.dx)) + 3;
        if (x_lo >= 0 && x_hi < gl.n - 1 && x_hi - x_lo + 1 < gl.n) {
          int wl0 = int(x_hi - x_lo + 1);
          size_t len = size_t(wl0) * nn;
          PBuf wacc(c->mod->pool, len * sizeof(double));
          fill<double>((double *)wacc.p, len, 0.0, c->st);
          fill<unsigned>((unsigned *)dropped.p, 1, 0u, c->st);
          deinterp(dg, nf3, (double *)wacc.p, len, size_t(x_lo) * nn, (unsigned *)dropped.p);
          PBuf win(c->mod->pool, len * sizeof(T));
          convert<double, T>((const double *)wacc.p, len, (T *)win.p, c->st);
          unsigned nd = 0;
          GIC_HIP(hipMemcpyAsync(&nd, dropped.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
          GIC_HIP(hipStreamSynchronize(c->st));
          if (nd == 0) {
            PBuf tab;
            level_table(0, tab);
            int w0[3] = {int(x_lo), 0, 0}, wl[3] = {wl0, gl.n, gl.n};
            Box3 B{{w0[0], 0, 0}, {wl[0], wl[1], wl[2]}};
            BoxBufs<T> bb(c, gl.n, B);
            fft_r2c_window<T>((const T *)win.p, gl.n, w0, wl, (T *)bb.b.p, (T *)bb.a.p, m.p(0), (const T *)tab.p,
                              false, c->st, c->hook());
            m.fourier[0] = true;
            done[0] = true;
            continue;   // buffers return to the pool; later users are ordered on the stream
          }
        }
      }
      size_t len = gl.n3() + 2 * nn;
      PBuf &acc = accs[l & 1];
      acc = PBuf(c->mod->pool, len * sizeof(double));
      fill<double>((double *)acc.p, len, 0.0, c->st);
      deinterp(dg, nf3, (double *)acc.p, len, size_t(0), (unsigned *)nullptr);
      pad<double, T>((const double *)acc.p, gl.n, m.p(size_t(l)), c->st);
      src_d = (const double *)acc.p;   // the next level reads this level's contiguous FP64 values
      src_padded = 0;
    }
  }
  // every other level: transform, then the level's factors in one pass
  for (size_t l = 0; l < nl; ++l) {
    if (done[l]) continue;
    PBuf tab;
    level_table(l, tab);
    m.toF(l);
    mul_table<T>(m.p(l), m.g(l).n, (const T *)tab.p, c->st);
  }
  if (kind == 1) {
    for (size_t l = 0; l < nl; ++l) {
      const LevelGeom &gg = m.g(l);
      size_t nm = size_t(gg.n) * gg.n * (gg.n / 2 + 1);
      hipLaunchKernelGGL(k_mode_op<T>, dim3(blocks(nm)), dim3(256), 0, c->st, reinterpret_cast<cx<T> *>(m.p(l)),
                         gg.n, gg.kmin(), 1, coef);
    }
  }
  GIC_HIP(hipGetLastError());
  return m;
}

// FilteredVarianceModification::pushMultiLevelFieldThroughMatrix
template <typename T> MLF<T> push_variance(MLF<T> &field, const std::vector<std::vector<uint64_t>> &flags,
                                           double scale) {
  gic_context *c = field.c;
  MLF<T> f = field.clone();
  transfer_to_dm(f);
  FilterSpec hp;
  hp.na = 1;
  hp.ka[0] = 2;
  hp.ca[0] = 2 * M_PI / scale;
  for (size_t l = 0; l < f.nl(); ++l) {
    const LevelGeom &g = f.g(l);
    size_t n3 = g.n3();
    static const std::vector<uint64_t> none;
    const std::vector<uint64_t> &cells = l < flags.size() ? flags[l] : none;
    PBuf mask(c->mod->pool, n3);
    fill<uint8_t>((uint8_t *)mask.p, n3, 0, c->st);
    DevList dl;
    upload_list(c, cells, dl);
    if (dl.n)
      hipLaunchKernelGGL(k_mask_set, dim3(blocks(dl.n)), dim3(256), 0, c->st, (uint8_t *)mask.p,
                         (const uint64_t *)dl.idx.p, dl.n);
    auto window = [&]() {
      f.toR(l);
      hipLaunchKernelGGL(k_mask_keep<T>, dim3(blocks(n3)), dim3(256), 0, c->st, f.p(l), g.n, (const uint8_t *)mask.p);
    };
    window();
    apply_spec(f, l, hp);
    window();
    if (dl.n) {
      c->mod->part.ensure(RED_BLOCKS * sizeof(double));
      c->mod->scalar.ensure(sizeof(double) * 8);
      hipLaunchKernelGGL(k_flag_sum<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, (const T *)f.p(l), g.n,
                         (const uint64_t *)dl.idx.p, dl.n, (double *)c->mod->part.p);
      hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, c->st, (const double *)c->mod->part.p,
                         RED_BLOCKS, (double *)c->mod->scalar.p);
      hipLaunchKernelGGL(k_flag_center<T>, dim3(blocks(dl.n)), dim3(256), 0, c->st, f.p(l), g.n,
                         (const uint64_t *)dl.idx.p, dl.n, (const double *)c->mod->scalar.p);
    }
    apply_spec(f, l, hp);
    window();
    GIC_HIP(hipStreamSynchronize(c->st));   // the cell list is released here
  }
  MLF<T> cov = metric(f, true);
  transfer_to_dm(cov);
  return cov;
}

void check_variance_scale(gic_context *c, double scale) {
  if (scale < 0.0)
    throw std::runtime_error(
        "Trying to calculate filtered variance without initialising variance filtering scale. Use filtering_scale "
        "command to do this.");
  if (scale < c->levels.back()->g.dx)
    throw std::runtime_error("Variance high-pass filtering scale is smaller than the smallest pixel size.");
}

template <typename T> double variance_value(MLF<T> &field, const std::vector<std::vector<uint64_t>> &flags,
                                            double scale) {
  MLF<T> p = push_variance(field, flags, scale);
  return inner(field.c, p, field);
}

// <w|M|w>, the chi^2 of getChi2 (multilevelfield.hpp:461-487), for a field
// in Fourier space.  Only the scalar is needed, so no covector is formed:
// per level, the windowed filter term reduces to one pass over the spectra
// of x and of its window u = M x (k_chi2_window_partial), the last level's
// term to sum F^2 |x|^2, and the cross terms to a dot product with x in
// real space on the box they reach (or in Fourier space on a zoom level).
// Falls back to the metric when a window is not a single box.
template <typename T> double chi2_of(gic_context *c, MLF<T> &field) {
  size_t nl = field.nl();
  field.toF();
  ensure_red(c);
  std::vector<Box3> wins(nl);
  for (size_t l = 0; l + 1 < nl; ++l)
    if (!window_box(c, c->levels[l]->g, grid_window(c->levels[l + 1]->g), wins[l]) ||
        !box_pruned(wins[l], c->levels[l]->g.n)) {
      MLF<T> cov = metric(field, true);
      return inner(c, cov, field);
    }
  double total = 0;
  for (size_t l = 0; l < nl; ++l) {
    const LevelGeom &g = c->levels[l]->g;
    const cx<T> *xs = reinterpret_cast<const cx<T> *>(field.p(l));
    PBuf ubox;   // x on the window box of the next level (u = M x there)
    int nred = RED_BLOCKS;
    if (l + 1 < nl) {
      const Box3 &M = wins[l];
      size_t nb = box_cells(M), nz = size_t(g.n / 2 + 1);
      ubox = PBuf(c->mod->pool, nb * sizeof(T));
      {
        PBuf A(c->mod->pool, size_t(M.bl[0]) * g.n * nz * 2 * sizeof(T));
        PBuf Bb(c->mod->pool, size_t(M.bl[0]) * M.bl[1] * nz * 2 * sizeof(T));
        fft_c2r_window<T>(field.p(l), g.n, FFT_PRE_NONE, nullptr, 0, g.kmin(), M.b0, M.bl, (T *)A.p, (T *)Bb.p,
                          (T *)ubox.p, c->st, c->hook());
      }
      PBuf uspec(c->mod->pool, g.padded_reals() * sizeof(T));
      {
        BoxBufs<T> bb(c, g.n, M);
        fft_r2c_window<T>((const T *)ubox.p, g.n, M.b0, M.bl, (T *)bb.b.p, (T *)bb.a.p, (T *)uspec.p, nullptr, false,
                          c->st, c->hook());
      }
      FilterSpec F = spec_levels(c, long(l), -1, 1.0);
      Timed t(c, "chi2_window", 2.0 * g.nmodes() * sizeof(cx<T>));
      hipLaunchKernelGGL(k_chi2_window_partial<T>, dim3(CHI2_BLOCKS), dim3(RED_THREADS), 0, c->st, xs,
                         reinterpret_cast<const cx<T> *>(uspec.p), F, g.kmin(), g.n, (double *)c->mod->part.p);
      nred = CHI2_BLOCKS;
    } else {
      const T *F2 = spec_table<T>(c, g, spec_levels(c, long(l), long(l), 1.0));
      Timed t(c, "inner", double(g.nmodes()) * sizeof(cx<T>));
      hipLaunchKernelGGL(k_inner_tab_partial<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, xs, xs, F2, g.n,
                         (double *)c->mod->part.p);
    }
    total += finish_reduction(c, 0, nred);
    if (nl > 1) {
      PBuf Rbox;
      Box3 B;
      MLF<T> R(c);
      if (cross_terms<T>(c, field, l, Rbox, B, R)) {
        size_t nb = box_cells(B), nz = size_t(g.n / 2 + 1);
        const T *xv;
        PBuf xbox;
        bool same = ubox.p && l + 1 < nl;
        for (int d = 0; same && d < 3; ++d) same = B.b0[d] == wins[l].b0[d] && B.bl[d] == wins[l].bl[d];
        if (same) {
          xv = (const T *)ubox.p;   // the same cells: x there is u
        } else {
          xbox = PBuf(c->mod->pool, nb * sizeof(T));
          PBuf A(c->mod->pool, size_t(B.bl[0]) * g.n * nz * 2 * sizeof(T));
          PBuf Bb(c->mod->pool, size_t(B.bl[0]) * B.bl[1] * nz * 2 * sizeof(T));
          fft_c2r_window<T>(field.p(l), g.n, FFT_PRE_NONE, nullptr, 0, g.kmin(), B.b0, B.bl, (T *)A.p, (T *)Bb.p,
                            (T *)xbox.p, c->st, c->hook());
          xv = (const T *)xbox.p;
        }
        hipLaunchKernelGGL(k_dot_partial<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, xv, (const T *)Rbox.p,
                           nb, (double *)c->mod->part.p);
      } else {
        R.toF(l);
        Timed t(c, "inner", 2.0 * g.nmodes() * sizeof(cx<T>));
        hipLaunchKernelGGL(k_inner_partial<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, xs,
                           reinterpret_cast<const cx<T> *>(R.p(l)), g.n, (double *)c->mod->part.p);
      }
      total += finish_reduction(c, 0);
    }
  }
  return total;
}

template <typename T> void ensure_ready(gic_context *c) {
  if (c->levels.empty()) throw std::runtime_error("No grid has yet been initialised");
  if (!c->mod) c->mod = std::make_unique<ModState>();
  if (!c->drawn) do_draw<T>(c);
  c->flags.resize(c->levels.size());
}

}  // namespace

// ============================================================== entry points ==
template <typename T> double mod_calculate(gic_context *c, const std::string &name) {
  ensure_ready<T>(c);
  OutputView<T> out(c);
  MLF<T> &field = out.m;
  field.toF();
  if (!strcasecmp(name.c_str(), "variance")) {
    check_variance_scale(c, c->variance_scale);
    return variance_value(field, c->flags, c->variance_scale);
  }
  MLF<T> cov = linear_covector<T>(c, name, c->flags);
  return inner(c, cov, field);
}

template <typename T> void mod_add(gic_context *c, const std::string &name, bool relative, double target,
                                   int init_steps, double precision) {
  ensure_ready<T>(c);
  ModRec m;
  m.name = name;
  m.flags = c->flags;
  m.scale = c->variance_scale;
  m.init_steps = init_steps;
  m.precision = precision;
  m.quadratic = !strcasecmp(name.c_str(), "variance");
  if (!m.quadratic && mod_kind(name) < 0) throw std::runtime_error(name + " is an unknown modification name");
  if (m.quadratic) check_variance_scale(c, m.scale);
  if (relative) target *= mod_calculate<T>(c, name);
  m.target = target;
  (m.quadratic ? c->mod->quadratic : c->mod->linear).push_back(m);
}

void mod_clear(gic_context *c) {
  if (c->mod) {
    c->mod->linear.clear();
    c->mod->quadratic.clear();
  }
}

template <typename T> double mod_chi2(gic_context *c) {
  ensure_ready<T>(c);
  if (c->power_applied) throw std::runtime_error("Cannot calculate the chi^2 after the output field has been generated");
  OutputView<T> out(c);
  out.m.toF();
  return chi2_of(c, out.m);
}

size_t num_dof(gic_context *c) {
  size_t dof = 0;
  for (size_t i = 0; i < c->levels.size(); ++i) {
    dof += c->levels[i]->g.n3();
    if (i > 0) {
      size_t f = (size_t)llround(c->levels[i - 1]->g.dx / c->levels[i]->g.dx);
      dof -= c->levels[i]->g.n3() / (f * f * f);
    }
  }
  return dof;
}

// ModificationManager::applyModifications (modificationmanager.hpp:80-331)
template <typename T> void mod_apply(gic_context *c, double *pre, double *post) {
  ensure_ready<T>(c);
  double chi_pre;
  {
    Phase ph(c, "phase:chi2_pre");
    chi_pre = mod_chi2<T>(c);
  }
  // the modifications update the noise in place
  std::unique_ptr<OutputView<T>> out(new OutputView<T>(c));
  MLF<T> &field = out->m;
  field.toF();
  std::vector<MLF<T>> alphas;
  std::vector<double> targets;
  {
    Phase ph(c, "phase:covectors");
    for (auto &m : c->mod->linear) {
      alphas.push_back(linear_covector<T>(c, m.name, m.flags));
      targets.push_back(m.target);
    }
  }
  auto cov_dot = [&](MLF<T> &a, MLF<T> &b) {
    MLF<T> bv = metric(b, false);
    return inner(c, a, bv);
  };
  // orthonormaliseModifications + applyLinearModif (modificationmanager.hpp:
  // 160-172,266-283) as one FP64 Gram solve.  The covector-covector product
  // goes through the inverse metric, which is linear: with v_i = M^-1 a_i,
  // G_ij = <a_i, v_j>, the Gram-Schmidt of the reference is the Cholesky
  // factor L of G, and the correction it adds is sum_i v_i [G^-1 (d - <a, w>)]_i.
  size_t nm = alphas.size();
  std::vector<MLF<T>> vecs;
  {
    Phase ph(c, "phase:metric_vectors");
    for (auto &a : alphas) vecs.push_back(metric(a, false));
  }
  std::vector<double> G(nm * nm), L(nm * nm, 0.0), rhs(nm);
  {
    Phase ph(c, "phase:gram");
    // every <alpha_i, v_j> (j <= i) and <alpha_i, w>, a block of up to 4
    // alphas against 4 vectors (and w) per pass and level
    if (nm > 0) {
      field.toF();
      for (size_t i = 0; i < nm; ++i) {
        alphas[i].toF();
        vecs[i].toF();
      }
      ensure_red(c);
      std::vector<double> gsum(nm * nm, 0.0), dsum(nm, 0.0);
      std::vector<double> part(GRAM_ROWS);
      for (size_t ib = 0; ib < nm; ib += 4)
        for (size_t jb = 0; jb <= ib; jb += 4) {
          size_t ie = std::min(nm, ib + 4), je = std::min(nm, jb + 4);
          for (size_t l = 0; l < field.nl(); ++l) {
            GramBlock gb{};
            for (size_t i = ib; i < ie; ++i) gb.a[i - ib] = alphas[i].p(l);
            for (size_t j = jb; j < je; ++j) gb.v[j - jb] = vecs[j].p(l);
            gb.w = jb == 0 ? field.p(l) : nullptr;
            gb.na = int(ie - ib);
            gb.nv = int(je - jb);
            const bool diag = ib == jb;
            const int nr = diag ? gram_rows<true>() : gram_rows<false>();
            int nfields = gb.na + gb.nv + (gb.w ? 1 : 0);
            Timed t(c, "multi_inner", double(nfields) * field.g(l).nmodes() * sizeof(cx<T>));
            if (diag)
              hipLaunchKernelGGL((k_gram_block_partial<T, true>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, gb,
                                 field.g(l).n, (double *)c->mod->part.p);
            else
              hipLaunchKernelGGL((k_gram_block_partial<T, false>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->st, gb,
                                 field.g(l).n, (double *)c->mod->part.p);
            hipLaunchKernelGGL(k_reduce_rows, dim3(nr), dim3(RED_THREADS), 0, c->st, (const double *)c->mod->part.p,
                               RED_BLOCKS, (double *)c->mod->scalar.p);
            GIC_HIP(hipGetLastError());
            GIC_HIP(hipMemcpyAsync(part.data(), c->mod->scalar.p, nr * sizeof(double), hipMemcpyDeviceToHost, c->st));
            GIC_HIP(hipStreamSynchronize(c->st));
            for (size_t i = ib; i < ie; ++i) {
              int p = int(i - ib);
              for (size_t j = jb; j < je && j <= i; ++j) {
                int q = int(j - jb);
                gsum[i * nm + j] += part[diag ? gram_row<true>(p, q) : gram_row<false>(p, q)];
              }
              if (jb == 0) dsum[i] += part[diag ? gram_wrow<true>(p) : gram_wrow<false>(p)];
            }
          }
        }
      for (size_t i = 0; i < nm; ++i) {
        for (size_t j = 0; j <= i; ++j) G[i * nm + j] = G[j * nm + i] = gsum[i * nm + j];
        rhs[i] = targets[i] - dsum[i];
      }
    } else {
      for (size_t i = 0; i < nm; ++i) {
        for (size_t j = 0; j <= i; ++j) G[i * nm + j] = G[j * nm + i] = inner(c, alphas[i], vecs[j]);
        rhs[i] = targets[i] - inner(c, alphas[i], field);
      }
    }
  }
  for (size_t i = 0; i < nm; ++i) {   // G = L L^T
    for (size_t j = 0; j <= i; ++j) {
      double v = G[i * nm + j];
      for (size_t k = 0; k < j; ++k) v -= L[i * nm + k] * L[j * nm + k];
      if (i == j) {
        if (!(v > 0)) throw std::runtime_error("Modifications are not linearly independent");
        L[i * nm + i] = sqrt(v);
      } else {
        L[i * nm + j] = v / L[j * nm + j];
      }
    }
  }
  std::vector<double> y(nm), coef(nm);
  for (size_t i = 0; i < nm; ++i) {
    double v = rhs[i];
    for (size_t k = 0; k < i; ++k) v -= L[i * nm + k] * y[k];
    y[i] = v / L[i * nm + i];
  }
  for (size_t ii = nm; ii-- > 0;) {
    double v = y[ii];
    for (size_t k = ii + 1; k < nm; ++k) v -= L[k * nm + ii] * coef[k];
    coef[ii] = v / L[ii * nm + ii];
  }
  {
    Phase ph(c, "phase:update");
    for (size_t i0 = 0; i0 < nm; i0 += 8) {
      size_t k = std::min<size_t>(8, nm - i0);
      for (size_t l = 0; l < field.nl(); ++l) {
        AxpyList al{};
        for (size_t j = 0; j < k; ++j) {
          al.x[j] = vecs[i0 + j].p(l);
          al.s[j] = coef[i0 + j];
        }
        al.k = int(k);
        size_t n = field.reals(l);
        Timed t(c, "axpy_multi", double(k + 2) * n * sizeof(T));
        if (sizeof(T) == 4 && n % 4 == 0)
          hipLaunchKernelGGL(k_axpy_multi4, dim3(blocks(n / 4)), dim3(256), 0, c->st, (float *)field.p(l), al, n / 4);
        else
          hipLaunchKernelGGL(k_axpy_multi<T>, dim3(blocks(n)), dim3(256), 0, c->st, field.p(l), al, n);
      }
    }
    GIC_HIP(hipGetLastError());
  }
  alphas.clear();
  if (!c->mod->quadratic.empty() && nm) {
    // the orthonormal vectors of the linear family: rows of L^-1 applied to v
    std::vector<double> Li(nm * nm, 0.0);
    for (size_t col = 0; col < nm; ++col)
      for (size_t i = 0; i < nm; ++i) {
        double v = i == col ? 1.0 : 0.0;
        for (size_t k = 0; k < i; ++k) v -= L[i * nm + k] * Li[k * nm + col];
        Li[i * nm + col] = v / L[i * nm + i];
      }
    std::vector<MLF<T>> ortho;
    for (size_t i = 0; i < nm; ++i) {
      MLF<T> o = vecs[i].clone();
      o.scale(Li[i * nm + i]);
      for (size_t j = 0; j < i; ++j) o.axpy(vecs[j], Li[i * nm + j]);
      ortho.push_back(std::move(o));
    }
    vecs = std::move(ortho);
  }
  for (auto &m : c->mod->quadratic) {
    // performIterations (modificationmanager.hpp:246-283): the value at f is
    // <Q f, f>, so one push gives both the value and the gradient direction
    auto iterate = [&](MLF<T> &f, int nsteps) {
      double start = variance_value(f, m.flags, m.scale);
      double stepv = (m.target - start) / nsteps;
      for (int i = 0; i < nsteps; ++i) {
        MLF<T> pushed = push_variance(f, m.flags, m.scale);
        double cur = inner(c, pushed, f);
        MLF<T> qf;
        if (c->mod->line_search) qf = pushed.clone();
        double norm = sqrt(cov_dot(pushed, pushed));
        for (auto &a : vecs) {
          double r = inner(c, pushed, a);
          pushed.axpy(a, -r);
        }
        double n2 = sqrt(cov_dot(pushed, pushed));
        pushed.scale(1.0 / n2);
        double target_next = start + (i + 1) * stepv;
        MLF<T> pv = metric(pushed, false);
        double mult;
        if (!c->mod->line_search) {
          mult = 0.5 * (target_next - cur) / norm;
        } else {
          // exact line search (no counterpart in the reference): Q is
          // symmetric, so along v the value is the quadratic
          // cur + 2 t <Q f, v> + t^2 <Q v, v>; take the root of
          // value = target_next nearest t = 0, or the vertex if none
          MLF<T> qv = push_variance(pv, m.flags, m.scale);
          double qa = inner(c, qv, pv), qb = 2 * inner(c, qf, pv), qc = cur - target_next;
          double disc = qb * qb - 4 * qa * qc;
          if (qa == 0) mult = qb != 0 ? -qc / qb : 0.0;
          else if (disc < 0) mult = -qb / (2 * qa);
          else {
            double sq = sqrt(disc);
            double q = -0.5 * (qb + (qb >= 0 ? sq : -sq));   // stable pair of roots q/qa, qc/q
            double r1 = q / qa, r2 = q != 0 ? qc / q : r1;
            mult = std::fabs(r1) < std::fabs(r2) ? r1 : r2;
          }
        }
        f.axpy(pv, mult);
      }
    };
    int init = m.init_steps;
    if (c->mod->line_search) {
      iterate(field, init);   // every step lands on its target: no trial run
      continue;
    }
    MLF<T> test = field.clone();
    iterate(test, init);
    double achieved = std::fabs(variance_value(test, m.flags, m.scale) - m.target);
    double want = m.target * m.precision;
    int nsteps = init * (int)std::ceil(std::sqrt(achieved / want));
    iterate(field, nsteps > init ? nsteps : init);
  }
  field.toF();
  for (size_t l = 0; l < field.nl(); ++l) c->levels[l]->dc_pending = false;
  out.reset();   // the levels' Fourier flags written back
  double chi_post;
  {
    Phase ph(c, "phase:chi2_post");
    chi_post = mod_chi2<T>(c);
  }
  if (pre) *pre = chi_pre;
  if (post) *post = chi_post;
}

// The filtered variance V(w) = <Q w, w> of the flagged region for an
// arbitrary white-noise field w (dense n_l^3 device arrays per level) and its
// gradient dV/dw = 2 Q w in real space (Q is symmetric in the white-noise
// inner product, which is the plain real-space dot product).  No
// counterpart in the reference, which only evaluates V
// (QuadraticModification::calculateCurrentValue, quadraticmodification.hpp:50-62).
template <typename T> double mod_variance_gradient(gic_context *c, const T *const *noise, T *const *grad) {
  if (c->levels.empty()) throw std::runtime_error("No grid has yet been initialised");
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
  if (!c->mod) c->mod = std::make_unique<ModState>();
  c->flags.resize(c->levels.size());
  check_variance_scale(c, c->variance_scale);
  bool pa = c->power_applied, zd = c->zeldovich_done;
  c->power_applied = c->zeldovich_done = false;   // the operator acts on white noise
  double v = 0;
  try {
    MLF<T> f(c);
    for (size_t l = 0; l < f.nl(); ++l) {
      if (!noise[l]) throw std::runtime_error("variance gradient: null device pointer");
      pad<T, T>(noise[l], f.g(l).n, f.p(l), c->st);
      f.fourier[l] = false;
    }
    MLF<T> p = push_variance(f, c->flags, c->variance_scale);
    v = inner(c, p, f);
    if (grad) {
      for (size_t l = 0; l < p.nl(); ++l) {
        if (!grad[l]) throw std::runtime_error("variance gradient: null device pointer");
        p.toR(l);
        size_t n3 = p.g(l).n3();
        unpad<T, T>(p.p(l), p.g(l).n, grad[l], c->st);
        hipLaunchKernelGGL(k_scale<T>, dim3(blocks(n3)), dim3(256), 0, c->st, grad[l], T(2), n3);
      }
      GIC_HIP(hipStreamSynchronize(c->st));
    }
  } catch (...) {
    c->power_applied = pa;
    c->zeldovich_done = zd;
    throw;
  }
  c->power_applied = pa;
  c->zeldovich_done = zd;
  return v;
}

void mod_set_line_search(gic_context *c, bool on) {
  if (!c->mod) c->mod = std::make_unique<ModState>();
  c->mod->line_search = on;
}

template double mod_variance_gradient<float>(gic_context *, const float *const *, float *const *);
template double mod_variance_gradient<double>(gic_context *, const double *const *, double *const *);
template double mod_calculate<float>(gic_context *, const std::string &);
template double mod_calculate<double>(gic_context *, const std::string &);
template void mod_add<float>(gic_context *, const std::string &, bool, double, int, double);
template void mod_add<double>(gic_context *, const std::string &, bool, double, int, double);
template void mod_apply<float>(gic_context *, double *, double *);
template void mod_apply<double>(gic_context *, double *, double *);
template double mod_chi2<float>(gic_context *);
template double mod_chi2<double>(gic_context *);

}  // namespace gic
