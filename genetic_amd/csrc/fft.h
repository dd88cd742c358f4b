// 3D FFT entry points and the per-kernel event timer hook.
#pragma once
#include "gic_common.h"

namespace gic {

// Called around every kernel launch when event timing is enabled; the
// implementation records hipEvents on the launch stream.
struct TimingHook {
  virtual ~TimingHook() = default;
  virtual void begin(const char *name, hipStream_t st) = 0;
  virtual void end(const char *name, hipStream_t st) = 0;
  // the algorithmic HBM bytes of one launch under this name: each input
  // element read once, each output element written once (reads of an
  // accumulated output counted too); tables and indices are not counted
  virtual void bytes(const char *, double) {}
};

// In-place 3D transform of a padded field (N*N*(N+2) reals of type T).
// dir < 0: r2c (real -> FFTW layout), dir > 0: c2r (with Hermitian mirroring
// of the kz = 0, N/2 planes).  Both directions scale by N^-1.5, as
// FieldFourierManager::performTransform does.
template <typename T> void fft3d(T *d, int n, int dir, hipStream_t st, TimingHook *hook);

// k-space operators applied inside the x pass
enum { FFT_PRE_NONE = 0, FFT_PRE_TABLE = 1, FFT_PRE_ZA = 2, FFT_PRE_KK = 3 };

// A product of Fermi level filters, formed inside a transform's pass instead
// of read from a |k|^2 table (used by FP32 transforms, whose table reads are
// scattered): scale * prod_i f_i(k), f = LP, sqrt(1 - LP^2) or 1 - LP for
// kind 0, 1, 2 with LP = 1 / (1 + exp((k - kc) / (0.1 kc))); over_k2 divides
// by k^2 (0 at k = 0); zero makes every factor 0.  nf < 0: not given.
// the most factors a FilterInline may hold (more: use a table)
constexpr int FILTER_INLINE_MAX = 4;
struct FilterInline {
  int nf = -1;
  int kind[8];
  double kc[8];
  double scale = 1.0;
  int over_k2 = 0;
  int zero = 0;
};

// r2c whose last pass multiplies each mode by post[kx^2+ky^2+kz^2] (if non-null).
// With src non-null the real input is read from the dense n^3 array src (z
// fastest, no padding) and the spectrum lands in d.
template <typename T>
void fft_r2c_post(T *d, int n, const T *post, hipStream_t st, TimingHook *hook, const T *src = nullptr);
// In-place c2r of d; with dst non-null the real result goes to the dense n^3
// array dst instead (d is then scratch).
// r2c of nb padded n^3 fields stored back to back, each mode times post[|k|^2] if post is given
template <typename T> void fft_r2c_batch(T *d, int n, int nb, const T *post, hipStream_t st, TimingHook *hook);
// builds the twiddle table of size n (a synchronous upload), so that later
// launches, also those recorded into a graph, need no host work
template <typename T> void fft_prepare(int n);
// r2c of nb fields that vanish outside boxes of wl x wl cells in (x, y), all
// z: box b starts at (winx[b], winy[b]) mod n (device int arrays).  rows
// holds the boxes' rows as dense rows of n reals, row (b wl + i) wl + j for
// x = winx[b]+i, y = winy[b]+j; nrows is a multiple of
// fft_box_rows_multiple(n) and rows past nb wl^2 are transformed but unused.
// zbuf (nrows (n/2+1) complex) and ybuf (nb wl n (n/2+1) complex) are
// scratch; field b's spectrum times post[|k|^2] goes to spec + b spec_stride.
template <typename T>
void fft_r2c_boxes(const T *rows, int n, long nrows, int nb, int wl, const int *winx, const int *winy, T *zbuf, T *ybuf,
                   T *spec, long spec_stride, const T *post, hipStream_t st, TimingHook *hook);
template <typename T> int fft_box_rows_multiple(int n);
template <typename T> void fft_c2r_to(T *d, T *dst, int n, hipStream_t st, TimingHook *hook);
// Transpose of one Zel'dovich component (FFT_PRE_ZA followed by c2r): the r2c
// of src (dense n^3, or d in place when null; d is overwritten) whose last
// pass forms conj(i k_c tab[|k|^2]) times each mode, zero at k_c = +Nyquist,
// and adds it to the spectrum acc (stores it when !accumulate).
template <typename T>
void fft_r2c_za_adjoint(const T *src, T *d, T *acc, int n, const T *tab, int comp, double kmin, bool accumulate,
                        hipStream_t st, TimingHook *hook);
// out <- c2r(op(spec)), out-of-place from a spectrum that is left untouched.
// op: FFT_PRE_TABLE multiplies by tab[|k|^2]; FFT_PRE_ZA forms the Zel'dovich
// component comp, i k_c tab[|k|^2], with tab = filter(k)/k^2 (0 at k=0) and
// zero at k_c = +Nyquist.  FFT_PRE_KK forms k_a k_b tab[|k|^2] for comp = a + 3 b
// (zero at a +Nyquist component when a != b).
// With dst non-null the real result goes to the dense n^3 array dst and out is scratch.
// With tab null and inv_k2 non-zero the operator's factor is inv_k2 / (|k|^2 kmin^2)
// (0 at k = 0), formed in FP64 as filter_table forms it for a unit filter.
// fil (optional) replaces tab for FP32 transforms.
template <typename T>
void fft_c2r_from(const T *spec, T *out, int n, int pre, const T *tab, int comp, double kmin, hipStream_t st,
                  TimingHook *hook, T *dst = nullptr, double inv_k2 = 0, const FilterInline *fil = nullptr);
// The second derivative phi_ab = c2r(k_a k_b tab[|k|^2] spec) (comp = a + 3b,
// as FFT_PRE_KK) folded into the 2LPT source by the last pass instead of being
// stored: epi 1: S = A phi, A += phi;  2: S += A phi;  3: S -= phi^2
// (S, A padded real fields; scratch holds the intermediate passes).
template <typename T>
void fft_c2r_lpt(const T *spec, T *scratch, int n, const T *tab, int comp, double kmin, int epi, T *S, T *A,
                 hipStream_t st, TimingHook *hook, double inv_k2 = 0);
// The same transforms split after their first pass, so that one x pass serves
// several outputs (c2r is linear and a factor of k_y, k_z commutes with the
// transform along x):
//   half <- the x pass of c2r(op(spec)) alone (op, comp, tab, inv_k2 as for
//           fft_c2r_from; half is a padded field holding complex x-k_y-k_z data);
//   out  <- the y and z passes of c2r from half, the y pass multiplying each
//           mode by m(k_y, k_z) as it reads it (FFT_YZ_*; dst as for
//           fft_c2r_from);
//   fft_c2r_yz_lpt: the same, its last pass folding the result into the 2LPT
//           source as fft_c2r_lpt does (out is scratch).
// With X0 = x pass of tab spec, X1 = x pass of i k_x tab spec (FFT_PRE_ZA comp 0)
// and X2 = x pass of k_x^2 tab spec (FFT_PRE_KK comp 0): the Zel'dovich
// components are yz(X1, ONE), yz(X0, IKY), yz(X0, IKZ) and the second
// derivatives yz(X2, ONE), yz(X1, MIKY), yz(X1, MIKZ), yz(X0, KYKY),
// yz(X0, KYKZ), yz(X0, KZKZ), with the +Nyquist rules of FFT_PRE_ZA/KK.
enum {
  FFT_YZ_ONE = 0,    // 1
  FFT_YZ_IKY = 1,    // i k_y (0 at k_y = +Nyquist)
  FFT_YZ_IKZ = 2,    // i k_z (0 at k_z = +Nyquist)
  FFT_YZ_MIKY = 3,   // -i k_y (0 at k_y = +Nyquist)
  FFT_YZ_MIKZ = 4,   // -i k_z (0 at k_z = +Nyquist)
  FFT_YZ_KYKY = 5,   // k_y^2
  FFT_YZ_KYKZ = 6,   // k_y k_z (0 at either +Nyquist)
  FFT_YZ_KZKZ = 7    // k_z^2
};
template <typename T>
void fft_c2r_xpass(const T *spec, T *half, int n, int pre, const T *tab, int comp, double kmin, hipStream_t st,
                   TimingHook *hook, double inv_k2 = 0);
template <typename T>
void fft_c2r_yz(const T *half, T *out, int n, int yop, double kmin, hipStream_t st, TimingHook *hook,
                T *dst = nullptr);
template <typename T>
void fft_c2r_yz_lpt(const T *half, T *scratch, int n, int yop, double kmin, int epi, T *S, T *A, hipStream_t st,
                    TimingHook *hook);
// Spectrum (r2c, N^-1.5) of a field that is zero outside the real-space window
// [w0, w0+wl) (per axis, wrapping mod n), given as the dense window win
// (wl[0]*wl[1]*wl[2] reals, z fastest): pruned z and y passes and a full x
// pass that scales by post[|k|^2] (if non-null) and stores into, or with
// accumulate adds to, the padded spectrum spec.  bufB holds
// wl[0]*wl[1]*(n+2) reals, bufA wl[0]*n*(n/2+1) complex.
// With aux (a spectrum of the same layout) the result is spec = post * aux +
// r2c(win) instead: post then scales aux, not the transform.
template <typename T>
void fft_r2c_window(const T *win, int n, const int *w0, const int *wl, T *bufB, T *bufA, T *spec, const T *post,
                    bool accumulate, hipStream_t st, TimingHook *hook, const T *aux = nullptr,
                    const FilterInline *fil = nullptr, double kmin = 0);
// Only the real-space window [w0, w0+wl) (per axis, wrapping mod n) of
// c2r(op(spec)), via pruned passes: bufA holds wl[0]*n*(n/2+1) complex,
// bufB wl[0]*wl[1]*(n/2+1) complex, out wl[0]*wl[1]*wl[2] reals.  yop
// (FFT_YZ_*) multiplies the y pass's input as fft_c2r_yz does; with xpass
// false bufA already holds the x pass of an earlier call on the same
// spectrum, operator and window, and only the y and z passes run.
template <typename T>
void fft_c2r_window(const T *spec, int n, int pre, const T *tab, int comp, double kmin, const int *w0, const int *wl,
                    T *bufA, T *bufB, T *out, hipStream_t st, TimingHook *hook, const FilterInline *fil = nullptr,
                    int yop = FFT_YZ_ONE, bool xpass = true);

}  // namespace gic
