// Fourier-space, level-combination and particle kernels.
#pragma once
#include "gic_common.h"

namespace gic {

// Natural cubic spline of log T(k) against log k, coefficients on the device.
struct SplineHost {
  double *d_x = nullptr, *d_y = nullptr, *d_c = nullptr;
  int n = 0;
  double minx = 0, maxx = 0;
};

struct InterpGeom {
  int off[3];        // fine-grid origin in supersampled coarse cells
  int sim;           // box size in supersampled cells (wrap period)
  int nsup;          // coarse grid size in supersampled cells
  double dxs;        // supersampled cell size
  double dxc;        // coarse cell size
  double co[3];      // coarse grid offsetLower
  double periodic;   // box size
  int nc;            // coarse grid cells per side
  int csim;          // box size in coarse cells
  int wrap;          // coarse grid covers the box: wrap, else clamp
};

// Where the coarse values live: element (cx,cy,cz) of the coarse grid (after
// wrapping/clamping) is p[lx*sx + ly*sy + lz] with l = (c - org) mod n < len.
// A full padded field is org = 0, len = n, sx = n*(n+2), sy = n+2.
struct CoarseView {
  const void *p;
  int org[3];
  int len[3];
  long sx, sy;
};

struct ParticleGeom {
  double dx;
  double o[3];
  double periodic;
  double vratio;    // velocity per unit first-order displacement
  double vratio2;   // velocity per unit second-order displacement
};

template <typename T>
void sqrtp_table(const SplineHost &sp, double amplitude, double ns, double kmin, double norm, long nent, T *out,
                 hipStream_t st);
// kind 0 low pass, 1 high pass (sqrt(1-lp^2)), 2 unity; over_k2 divides by k^2
template <typename T>
void filter_table(int kind, double kc, double kmin, long nent, T *out, hipStream_t st, bool over_k2 = false,
                  double scale = 1.0);
template <typename T> void mul_table(T *field, int n, const T *tab, hipStream_t st);
// sum[m] += |spec_k|^2, cnt[m] += 1 over (ix,iy,iz) in [-N/2,N/2]^3, m = |k_int|^2
template <typename T> void mode_power(const T *spec, int n, double *sum, double *cnt, hipStream_t st);
template <typename T>
void zeldovich(const T *delta, int n, double kmin, T *px, T *py, T *pz, const T *post, hipStream_t st);
template <typename T> void interp_add(T *fine, int nf, const CoarseView &coarse, const InterpGeom &g, hipStream_t st);
// transpose of interp_add with the coarse field whole and padded: coarse += I^T fine
// (fine row stride frs: nf dense, nf+2 padded)
template <typename T>
void interp_adjoint(const T *fine, int nf, long frs, T *coarse, const InterpGeom &g, hipStream_t st);
// psi2: the three second-order displacement fields, or nullptr for Zel'dovich
template <typename T>
void particles(const T *px, const T *py, const T *pz, const T *const *psi2, int n, const ParticleGeom &g,
               const uint64_t *cells, size_t start, size_t count, float *pos, float *vel, hipStream_t st);
// elementwise source accumulation: mode 0 S = A B, 1 S += A B, 2 S -= A A, 3 A += B
template <typename T> void lpt_accum(T *S, T *A, const T *B, size_t n, int mode, hipStream_t st);
template <typename T> void fill(T *p, size_t n, T v, hipStream_t st);
// white noise adjustments: every stored mode to unit modulus; sign flip of the
// modes with 0 < k^2 < kmax^2; p *= s over n values; out = 1/tab (0 where tab is 0)
template <typename T> void unit_modulus(T *field, int n, hipStream_t st);
template <typename T> void reverse_small_k(T *field, int n, double kmin, double kmax, hipStream_t st);
template <typename T> void scale_all(T *p, size_t n, T s, hipStream_t st);
template <typename T> void inverse_table(const T *tab, long n, T *out, hipStream_t st);
// Fourier-space white noise: number of (a, b) draws of an n^3 level, and the
// placement of gauss[0 .. 2*count) (already scaled by 1/sqrt 2) into the
// padded spectrum 'field'; owner holds n*n*(n/2+1) words of scratch.
long fourier_draw_count(int n);
template <typename T>
void fourier_draw_scatter(const T *gauss, int n, T *field, uint32_t *owner, int reverse, hipStream_t st);
template <typename S, typename D> void unpad(const S *src, int n, D *dst, hipStream_t st);
template <typename S, typename D> void pad(const S *src, int n, D *dst, hipStream_t st);
template <typename S, typename D> void convert(const S *a, size_t n, D *b, hipStream_t st);

}  // namespace gic
