// 3D real<->complex FFT for the padded FFTW layout, written for gfx950.
//
// Replaces the FFTW calls behind FieldFourierManager::performTransform
// (genetIC/src/tools/numerics/fourier.hpp:479-563): in-place r2c / c2r over
// [N][N][N+2] reals <-> [N][N][N/2+1] complex, forward sign e^{-2 pi i k.x/N}.
// The reference's pad/unpad copies (fourier.hpp:368-405) disappear because
// the real-space field stays padded in HBM.
//
// One kernel launch per axis:
//   z  : contiguous rows, N/2-point complex Stockham FFT of the packed
//        (even, odd) samples plus the r2c post-/c2r pre-twiddle;
//   y,x: strided lines, a tile of C adjacent kz columns is staged in LDS so
//        every global access is a C*sizeof(complex)-byte contiguous segment.
// Stages are radix-8 (plus one radix-4 or radix-16 stage) Stockham autosort steps;
// the first stage reads HBM, the last writes HBM, the rest go through LDS.
// Twiddles come from an N-entry table generated in FP64 on the host.
#include "gic_common.h"
#include "fft.h"
#include "fft_tiles.h"
#include <algorithm>
#include <vector>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

namespace gic {

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// WIDE: where the radix-8 stages leave a factor 2, the first stage takes 16
// points instead (one stage, and one pass through LDS, fewer).  It pays
// where the first stage has no |k|^2 operator (none, or a y pass's k_y / k_z
// factor); with a filter or table the 16 slots' factors take more registers
// than two blocks per CU allow.
template <int L, bool WIDE = false> struct RadixPlan {
  static constexpr int lg = ilog2(L);
  static constexpr int rem = lg % 3;
  static constexpr bool r16 = WIDE && rem == 1 && lg >= 4;
  static constexpr int nstages = r16 ? (lg - 4) / 3 + 1 : lg / 3 + (rem ? 1 : 0);
  static constexpr int radix(int s) { return s == 0 ? (r16 ? 16 : rem ? (1 << rem) : 8) : 8; }
  static constexpr int pbefore(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= radix(i);
    return p;
  }
};

// multiply by i*DIR
template <int DIR, typename C> __device__ __forceinline__ C mul_i(C a) {
  return DIR > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}

template <int R, int DIR, typename T> struct Dft;

template <int DIR, typename T> struct Dft<1, DIR, T> {
  __device__ __forceinline__ static void run(cx<T> *) {}
};
template <int DIR, typename T> struct Dft<2, DIR, T> {
  __device__ __forceinline__ static void run(cx<T> *x) {
    cx<T> a = x[0];
    x[0] = cadd(a, x[1]);
    x[1] = csub(a, x[1]);
  }
};
template <int DIR, typename T> struct Dft<4, DIR, T> {
  __device__ __forceinline__ static void run(cx<T> *x) {
    cx<T> a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    cx<T> a2 = cadd(x[1], x[3]), a3 = mul_i<DIR>(csub(x[1], x[3]));
    x[0] = cadd(a0, a2);
    x[2] = csub(a0, a2);
    x[1] = cadd(a1, a3);
    x[3] = csub(a1, a3);
  }
};
template <int DIR, typename T> struct Dft<8, DIR, T> {
  __device__ __forceinline__ static void run(cx<T> *x) {
    cx<T> e[4] = {x[0], x[2], x[4], x[6]};
    cx<T> o[4] = {x[1], x[3], x[5], x[7]};
    Dft<4, DIR, T>::run(e);
    Dft<4, DIR, T>::run(o);
    const T h = T(0.70710678118654752440084436210484903928);
    // w8^1 = (1 + DIR i)/sqrt2, w8^2 = DIR i, w8^3 = (-1 + DIR i)/sqrt2
    cx<T> o1 = o[1], o3 = o[3];
    cx<T> t1 = cx<T>{h * (o1.x - DIR * o1.y), h * (o1.y + DIR * o1.x)};
    cx<T> t2 = mul_i<DIR>(o[2]);
    cx<T> t3 = cx<T>{h * (-o3.x - DIR * o3.y), h * (-o3.y + DIR * o3.x)};
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], t1);
    x[5] = csub(e[1], t1);
    x[2] = cadd(e[2], t2);
    x[6] = csub(e[2], t2);
    x[3] = cadd(e[3], t3);
    x[7] = csub(e[3], t3);
  }
};

// 16 points as 4 x 4: DFT-4 of the four decimated columns, the W16^(jk)
// twiddles, DFT-4 of the rows; output k + 4m in natural order
template <int DIR, typename T> struct Dft<16, DIR, T> {
  __device__ __forceinline__ static cx<T> rot(cx<T> a, T c, T s) {   // a * (c + DIR s i)
    return cx<T>{a.x * c - T(DIR) * s * a.y, a.y * c + T(DIR) * s * a.x};
  }
  __device__ __forceinline__ static void run(cx<T> *x) {
    const T c1 = T(0.92387953251128675612818318939678828682), s1 = T(0.38268343236508977172845998403039886676);
    const T h = T(0.70710678118654752440084436210484903928);
    cx<T> a[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int m = 0; m < 4; ++m) a[j][m] = x[j + 4 * m];
      Dft<4, DIR, T>::run(a[j]);
    }
    a[1][1] = rot(a[1][1], c1, s1);    // W^1
    a[1][2] = rot(a[1][2], h, h);      // W^2
    a[1][3] = rot(a[1][3], s1, c1);    // W^3
    a[2][1] = rot(a[2][1], h, h);      // W^2
    a[2][2] = mul_i<DIR>(a[2][2]);     // W^4
    a[2][3] = rot(a[2][3], -h, h);     // W^6
    a[3][1] = rot(a[3][1], s1, c1);    // W^3
    a[3][2] = rot(a[3][2], -h, h);     // W^6
    a[3][3] = rot(a[3][3], -c1, -s1);  // W^9
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      cx<T> b[4] = {a[0][k], a[1][k], a[2][k], a[3][k]};
      Dft<4, DIR, T>::run(b);
#pragma unroll
      for (int m = 0; m < 4; ++m) x[k + 4 * m] = b[m];
    }
  }
};

// XCD-aware tile order: consecutive tiles land on the same XCD (the
// dispatcher places workgroup b on XCD b % 8), so tiles that share 128-byte
// lines are served by one L2.
__device__ __forceinline__ int xcd_tile(int b, int nblk) {
  if ((nblk & 7) != 0) return b;
  return (b & 7) * (nblk >> 3) + (b >> 3);
}

// One Stockham stage.  Butterfly q is mapped to (line, i): LINE_FAST puts
// adjacent q on adjacent lines (strided passes), otherwise on adjacent points.
// An operator on the first stage's loads has a real factor per (line, point)
// that depends only on the mode (factor) and combines it with the loaded
// value (apply).
struct NoOp {};

template <typename T, int L, int NL, int THREADS, int S, int DIR, bool LINE_FAST, bool WIDE, class Ld, class St,
          class Op = NoOp>
__device__ __forceinline__ void stage(const cx<T> *__restrict__ tw, int twmul, Ld ld, St st, bool sync_between,
                                      Op op = Op()) {
  using P = RadixPlan<L, WIDE>;
  constexpr int R = P::radix(S);
  constexpr int p = P::pbefore(S);
  constexpr int NB = NL * (L / R);
  constexpr int PER = (NB + THREADS - 1) / THREADS;
  const int tid = threadIdx.x;
  cx<T> v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    int q = tid + b * THREADS;
    if (NB % THREADS == 0 || q < NB) {
      int line = LINE_FAST ? q % NL : q / (L / R);
      int i = LINE_FAST ? q / NL : q % (L / R);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = ld(line, i + r * (L / R));
    }
  }
  // operators: every factor is formed while the loads are in flight (they
  // depend only on the mode), and only then applied to the loaded values
  if constexpr (!std::is_same<Op, NoOp>::value) {
    // all slots' factors in one call; an idle slot (no butterfly) takes
    // line 0 and point r L/R, inside the tile, and its factor is unused
    int ln[PER * R], pp[PER * R];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
      int q = tid + b * THREADS;
      bool has = NB % THREADS == 0 || q < NB;
      int line = has ? (LINE_FAST ? q % NL : q / (L / R)) : 0;
      int i = has ? (LINE_FAST ? q / NL : q % (L / R)) : 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        ln[b * R + r] = line;
        pp[b * R + r] = i + r * (L / R);
      }
    }
    // in groups of four slots: enough chains to interleave, few enough
    // registers for two blocks per CU.  Each group is applied as soon as it
    // is formed, so that only its four factors are held and it waits only on
    // its own loads (the later slots' loads stay in flight)
    constexpr int NSL = PER * R, CH = NSL % 4 == 0 ? 4 : NSL;
#pragma unroll
    for (int c0 = 0; c0 < NSL; c0 += CH) {
      int lc[CH], pc[CH];
      T fc[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        lc[j] = ln[c0 + j];
        pc[j] = pp[c0 + j];
      }
      op.factors(lc, pc, fc);
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int b = (c0 + j) / R, r = (c0 + j) % R;
        if (NB % THREADS == 0 || tid + b * THREADS < NB) v[b][r] = op.apply(v[b][r], fc[j]);
      }
    }
  }
  if (sync_between) __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    int q = tid + b * THREADS;
    if (NB % THREADS == 0 || q < NB) {
      int line = LINE_FAST ? q % NL : q / (L / R);
      int i = LINE_FAST ? q / NL : q % (L / R);
      int k = i % p;
      if (p > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = tw[(r * k * (L / (p * R))) * twmul];
          if (DIR > 0) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      Dft<R, DIR, T>::run(v[b]);
      int j = (i - k) * R + k;
#pragma unroll
      for (int s = 0; s < R; ++s) st(line, j + s * p, v[b][s]);
    }
  }
}

// Run all stages: stage 0 loads with gld, the last stage stores with gst, the
// others go through LDS laid out by lidx(line, point).  GLD_LDS / GST_LDS say
// whether gld / gst themselves touch LDS, in which case the stage must put a
// barrier between its loads and its stores.
template <typename T, int L, int NL, int THREADS, int DIR, bool LINE_FAST, bool GLD_LDS, bool GST_LDS, int S,
          bool WIDE, class GLd, class GSt, class LIdx, class Op = NoOp>
__device__ __forceinline__ void run_stages(cx<T> *lds, const cx<T> *tw, int twmul, GLd gld, GSt gst, LIdx lidx,
                                           Op op = Op()) {
  using P = RadixPlan<L, WIDE>;
  constexpr int NS = P::nstages;
  auto lds_ld = [&](int line, int pt) { return lds[lidx(line, pt)]; };
  auto lds_st = [&](int line, int pt, cx<T> v) { lds[lidx(line, pt)] = v; };
  if constexpr (NS == 1) {
    stage<T, L, NL, THREADS, 0, DIR, LINE_FAST, WIDE>(tw, twmul, gld, gst, GLD_LDS && GST_LDS, op);
  } else if constexpr (S == 0) {
    stage<T, L, NL, THREADS, 0, DIR, LINE_FAST, WIDE>(tw, twmul, gld, lds_st, GLD_LDS, op);
    __syncthreads();
    run_stages<T, L, NL, THREADS, DIR, LINE_FAST, GLD_LDS, GST_LDS, 1, WIDE>(lds, tw, twmul, gld, gst, lidx);
  } else if constexpr (S == NS - 1) {
    stage<T, L, NL, THREADS, S, DIR, LINE_FAST, WIDE>(tw, twmul, lds_ld, gst, GST_LDS);
  } else {
    stage<T, L, NL, THREADS, S, DIR, LINE_FAST, WIDE>(tw, twmul, lds_ld, lds_st, true);
    __syncthreads();
    run_stages<T, L, NL, THREADS, DIR, LINE_FAST, GLD_LDS, GST_LDS, S + 1, WIDE>(lds, tw, twmul, gld, gst, lidx);
  }
}

// ---- strided c2c pass (y or x) ------------------------------------------------
// Lines of length L along one axis; a tile holds C adjacent kz columns of one
// `outer` index.  Input and output may differ (out-of-place) and the output
// may keep only a window of the line (pruned inverse transforms, where the
// line's output is a spatial coordinate).  Optional k-space operators:
//   PRE  (inverse x pass, input indexed by kx=pt, ky=outer, kz=col):
//        PRE_TABLE: v *= tab[kx^2+ky^2+kz^2]
//        PRE_ZA:    v <- i k_c v / k^2 (zero at k_c = +Nyquist and k = 0), times tab if given
//        PRE_KK:    v <- k_a k_b tab[|k|^2] v (second derivatives; a != b zero at a Nyquist component)
//   POST (forward x pass, output indexed by kx=pt, ky=outer, kz=col):
//        POST_TABLE: v *= tab[kx^2+ky^2+kz^2]
//        POST_ZA_ADJ: out (+)= -i k_c tab[|k|^2] v, zero at k_c = +Nyquist: the
//                     transpose of PRE_ZA, accumulated over components in out
//        POST_AUX:   out = tab[|k|^2] aux + v, aux a spectrum laid out as out
// PRE_YZ: the y pass of a field already transformed along x multiplies each
// mode by the FFT_YZ_* factor a.comp of (k_y, k_z)
enum { PRE_NONE = FFT_PRE_NONE, PRE_TABLE = FFT_PRE_TABLE, PRE_ZA = FFT_PRE_ZA, PRE_KK = FFT_PRE_KK, PRE_YZ = 4 };
enum { POST_NONE = 0, POST_TABLE = 1, POST_ZA_ADJ = 2, POST_AUX = 3 };

// FilterInline prepared for single precision on the host: per filter the
// cut kc and 1 / (0.1 kc), the scale, kMin and 1 / kMin^2
struct FilterF {
  int nf, zero, over_k2;
  int kind[8];
  float kc[8], inv_w[8];
  float scale, kmin, inv_kmin2;
};

template <typename T> struct StridedArgs {
  const cx<T> *in;
  cx<T> *out;
  long in_ls, in_os, out_ls, out_os;
  int ncol, nkzb;
  int n;            // full grid size, for signed wavenumbers
  const T *tab;     // |k|^2-indexed factor table (or nullptr)
  int comp;         // Zel'dovich component 0,1,2
  T kmin;
  int win0, wlen;   // output keeps line points (pt - win0) mod n < wlen
  T scale;
  int accum;        // POST ops: add to the output instead of storing
  const cx<T> *aux;        // POST_AUX: the spectrum added after its table
  double inv_k2, kmin_d;   // PRE ops without a table: factor inv_k2 / (|k|^2 kmin_d^2)
  FilterInline fil;        // FP32: the table's filter product formed in the pass (fil.nf >= 0)
  int in_win0, in_wlen;   // if 0 < in_wlen < n: the input holds only line points (pt - in_win0) mod n < in_wlen,
                          // stored at index (pt - in_win0) mod n
  FilterF ff;             // FP32: fil in single precision with its divisions turned into products (set by strided())
  float inv_k2_f;         // FP32: inv_k2 / kMin^2 (set by strided())
  int nbatch, nouter;     // nbatch > 1: that many fields, in_bs / out_bs elements apart, nouter outer lines each
  long in_bs, out_bs;
  const int *in_win0_b;   // if given: in_win0 of each field of the batch (device array)
};

__device__ __forceinline__ int signed_k(int i, int n) { return i > n / 2 ? i - n : i; }

// The factor of FilterInline at |k|^2 = ksq (filter_table's product) in
// single precision, with hardware reciprocal and square root (1 ulp each)
// and the host's reciprocals: no division per mode.
__device__ __forceinline__ float filter_fast_f(const FilterF &F, int ksq) {
  if (F.zero) return 0.f;
  float fk = float(ksq);
  float k = __builtin_amdgcn_sqrtf(fk) * F.kmin;
  float v = F.scale;
  // a fixed bound with a uniform exit: the parameters are indexed by
  // constants and stay in scalar registers
#pragma unroll
  for (int i = 0; i < FILTER_INLINE_MAX; ++i) {
    if (i >= F.nf) break;
    float lp = __builtin_amdgcn_rcpf(1.f + __expf((k - F.kc[i]) * F.inv_w[i]));
    v *= F.kind[i] == 0 ? lp : F.kind[i] == 1 ? __builtin_amdgcn_sqrtf(fmaxf(1.f - lp * lp, 0.f)) : 1.f - lp;
  }
  if (F.over_k2) v = ksq == 0 ? 0.f : v * F.inv_kmin2 * __builtin_amdgcn_rcpf(fk);
  return v;
}
// filter_fast_f for NS values of |k|^2 at once, one factor at a time across
// them: the parameter tests are taken once for all values and their
// exp / rcp / sqrt chains interleave.  The same arithmetic per value.
template <int NS>
__device__ __forceinline__ void filter_fast_f_n(const FilterF &F, const int (&ksq)[NS], float (&out)[NS]) {
  if (F.zero) {
#pragma unroll
    for (int s = 0; s < NS; ++s) out[s] = 0.f;
    return;
  }
  float k[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    k[s] = __builtin_amdgcn_sqrtf(float(ksq[s])) * F.kmin;
    out[s] = F.scale;
  }
#pragma unroll
  for (int i = 0; i < FILTER_INLINE_MAX; ++i) {
    if (i >= F.nf) break;
    const float kc = F.kc[i], iw = F.inv_w[i];
    const int kind = F.kind[i];
    float lp[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) lp[s] = __builtin_amdgcn_rcpf(1.f + __expf((k[s] - kc) * iw));
    if (kind == 0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) out[s] *= lp[s];
    } else if (kind == 1) {
#pragma unroll
      for (int s = 0; s < NS; ++s) out[s] *= __builtin_amdgcn_sqrtf(fmaxf(1.f - lp[s] * lp[s], 0.f));
    } else {
#pragma unroll
      for (int s = 0; s < NS; ++s) out[s] *= 1.f - lp[s];
    }
  }
  if (F.over_k2) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
      out[s] = ksq[s] == 0 ? 0.f : out[s] * F.inv_kmin2 * __builtin_amdgcn_rcpf(float(ksq[s]));
  }
}
template <typename T, class A> __device__ __forceinline__ T kfactor(const A &a, int ksq) {
  if constexpr (sizeof(T) == 4) {
    if (a.ff.nf >= 0) return filter_fast_f(a.ff, ksq);
    if (!a.tab) return ksq == 0 ? 0.f : a.inv_k2_f * __builtin_amdgcn_rcpf(float(ksq));
  }
  return a.tab ? a.tab[ksq] : T(ksq == 0 ? 0.0 : (1.0 / (double(ksq) * a.kmin_d * a.kmin_d)) * a.inv_k2);
}

// PRE_ZA multiplies by i * k_c * tab[|k|^2], where the caller's table holds
// filter(k) / k^2 (0 at k = 0); the component at k_c = +Nyquist is zero.
// PASS only names the pass (0 x, 1 y, 2 windowed x, 3 windowed y) so that
// profiles tell the launches apart.
// The part of the field one tile of a strided pass works on: C adjacent kz
// columns of one outer line (of one field of a batch).
template <typename T> struct TileG {
  int col0, nc, ky, ky2, iw0;
  const cx<T> *ib;
  cx<T> *ob;
};
template <typename T, int C> __device__ __forceinline__ TileG<T> tile_geom(const StridedArgs<T> &a, int tile) {
  TileG<T> g;
  int outer = tile / a.nkzb, kzb = tile - outer * a.nkzb;
  g.col0 = kzb * C;
  g.nc = min(C, a.ncol - g.col0);
  long boff_in = 0, boff_out = 0;
  int b = 0;
  if (a.nbatch > 1) {
    b = outer / a.nouter;
    outer -= b * a.nouter;
    boff_in = b * a.in_bs;
    boff_out = b * a.out_bs;
  }
  g.iw0 = a.in_win0_b ? a.in_win0_b[b] : a.in_win0;
  g.ib = a.in + boff_in + outer * a.in_os + g.col0;
  g.ob = a.out + boff_out + outer * a.out_os + g.col0;
  g.ky = signed_k(outer, a.n);
  g.ky2 = g.ky * g.ky;
  return g;
}
// input value at (line, point) of a tile: zero past the last column and, for
// a pruned input, outside its window
template <typename T, int L>
__device__ __forceinline__ cx<T> tile_in(int in_wlen, long in_ls, const TileG<T> &g, int line, int pt) {
  if (line >= g.nc) return cx<T>{0, 0};
  if (in_wlen > 0 && in_wlen < L) {
    int o = pt - g.iw0;
    if (o < 0) o += L;
    if (o >= in_wlen) return cx<T>{0, 0};
    return g.ib[o * in_ls + line];
  }
  return g.ib[pt * in_ls + line];
}
// the operator's parameters, copied out of the kernel arguments once per
// tile (read through a reference they were reloaded for every mode)
template <typename T> struct PreArgs {
  const T *tab;
  int comp, ncol, n;
  T kmin;
  double inv_k2, kmin_d;
  FilterF ff;
  float inv_k2_f;
};
template <typename T, int PRE> struct PreOp {
  PreArgs<T> a;
  int col0, ky, ky2;
  // the factor of |k|^2: the caller's table, its filter formed here (FP32),
  // or inv_k2 / (|k|^2 kMin^2) formed exactly as filter_table forms it;
  // times k_c (Zel'dovich, zero at k_c = +Nyquist) or k_a k_b (second
  // derivatives, zero at a +Nyquist component when a != b)
  __device__ __forceinline__ T factor(int line, int pt) const {
    if constexpr (PRE == PRE_YZ) {   // a y pass: points run over k_y, columns over k_z
      int kyy = signed_k(pt, a.n), kz = col0 + line, nyq = a.n / 2;
      if (kz >= a.ncol) return T(0);
      T fy = T(kyy) * a.kmin, fz = T(kz) * a.kmin;
      switch (a.comp) {
        case FFT_YZ_IKY: return kyy == nyq ? T(0) : fy;
        case FFT_YZ_IKZ: return kz == nyq ? T(0) : fz;
        case FFT_YZ_MIKY: return kyy == nyq ? T(0) : -fy;
        case FFT_YZ_MIKZ: return kz == nyq ? T(0) : -fz;
        case FFT_YZ_KYKY: return fy * fy;
        case FFT_YZ_KYKZ: return (kyy == nyq || kz == nyq) ? T(0) : fy * fz;
        case FFT_YZ_KZKZ: return fz * fz;
        default: return T(1);
      }
    }
    int kx = signed_k(pt, a.n), kz = col0 + line;
    if (kz >= a.ncol) return T(0);   // past the last column (its loads are zero): no table read
    int ksq = kx * kx + ky2 + kz * kz;
    return finish(kx, kz, kfactor<T>(a, ksq));
  }
  // factor() for NS slots at once; FP32 filters go through filter_fast_f_n
  template <int NS>
  __device__ __forceinline__ void factors(const int (&line)[NS], const int (&pt)[NS], T (&f)[NS]) const {
    if constexpr (PRE != PRE_YZ && sizeof(T) == 4) {
      if (a.ff.nf >= 0) {
        int ksq[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          int kx = signed_k(pt[s], a.n), kz = col0 + line[s];
          ksq[s] = kx * kx + ky2 + kz * kz;
        }
        float base[NS];
        filter_fast_f_n<NS>(a.ff, ksq, base);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          int kx = signed_k(pt[s], a.n), kz = col0 + line[s];
          f[s] = kz >= a.ncol ? T(0) : finish(kx, kz, base[s]);
        }
        return;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) f[s] = factor(line[s], pt[s]);
  }
  // the |k|^2 factor f times the operator's k-components (PRE_ZA, PRE_KK)
  __device__ __forceinline__ T finish(int kx, int kz, T f) const {
    if constexpr (PRE == PRE_ZA) {
      int kc = a.comp == 0 ? kx : a.comp == 1 ? ky : kz;
      return kc == a.n / 2 ? T(0) : f * (T(kc) * a.kmin);
    } else if constexpr (PRE == PRE_KK) {
      int ca = a.comp % 3, cb = a.comp / 3;
      int k1 = ca == 0 ? kx : ca == 1 ? ky : kz;
      int k2 = cb == 0 ? kx : cb == 1 ? ky : kz;
      if (ca != cb && (k1 == a.n / 2 || k2 == a.n / 2)) return T(0);
      return f * ((T(k1) * a.kmin) * (T(k2) * a.kmin));
    } else {
      return f;
    }
  }
  // Zel'dovich and the first-derivative y factors: i f v; otherwise f v
  __device__ __forceinline__ cx<T> apply(cx<T> v, T f) const {
    if constexpr (PRE == PRE_ZA) {
      return cx<T>{-v.y * f, v.x * f};
    } else if constexpr (PRE == PRE_YZ) {
      if (a.comp >= FFT_YZ_IKY && a.comp <= FFT_YZ_MIKZ) return cx<T>{-v.y * f, v.x * f};
      return cx<T>{v.x * f, v.y * f};
    } else {
      return cx<T>{v.x * f, v.y * f};
    }
  }
};
// One tile of a strided pass; ld supplies the first stage's inputs.
template <typename T, int L, int C, int DIR, int THREADS, int PRE, int POST, class Ld>
__device__ __forceinline__ void strided_tile(const StridedArgs<T> &a, const cx<T> *__restrict__ tw, cx<T> *lds,
                                             const TileG<T> &g, Ld gld) {
  const int col0 = g.col0, nc = g.nc, ky = g.ky, ky2 = g.ky2;
  cx<T> *ob = g.ob;
  PreOp<T, PRE> op{PreArgs<T>{a.tab, a.comp, a.ncol, a.n, a.kmin, a.inv_k2, a.kmin_d, a.ff, a.inv_k2_f}, col0, ky, ky2};
  constexpr bool WIDE = PRE == PRE_NONE || PRE == PRE_YZ;
  // the store path's parameters, read once per tile
  const int wlen = a.wlen, win0 = a.win0, accum = a.accum;
  const long out_ls = a.out_ls;
  const T scale = a.scale;
  const cx<T> *aux = a.aux, *out0 = a.out;
  auto gst = [&](int line, int pt, cx<T> v) {
    if (line >= nc) return;
    int o = pt;
    if (wlen < L) {
      o = pt - win0;
      if (o < 0) o += L;
      if (o >= wlen) return;
    }
    T s = scale;
    if constexpr (POST == POST_TABLE) {
      int kx = signed_k(pt, op.a.n), kz = col0 + line;
      s *= kfactor<T>(op.a, kx * kx + ky2 + kz * kz);
    }
    if constexpr (POST == POST_NONE || POST == POST_TABLE) {
      if (accum) {
        cx<T> &dst = ob[o * out_ls + line];
        cx<T> prev = dst;
        dst = cx<T>{prev.x + v.x * s, prev.y + v.y * s};
        return;
      }
    }
    if constexpr (POST == POST_AUX) {
      // the added spectrum is read here, at the store: fetched before the
      // first stage it held 16 more registers through the transform and left
      // one block per CU instead of two
      int kx = signed_k(pt, op.a.n), kz = col0 + line;
      T f = kfactor<T>(op.a, kx * kx + ky2 + kz * kz);
      cx<T> x = aux[long(o) * out_ls + line + (ob - out0)];
      ob[o * out_ls + line] = cx<T>{x.x * f + v.x * s, x.y * f + v.y * s};
      return;
    }
    if constexpr (POST == POST_ZA_ADJ) {
      int kx = signed_k(pt, op.a.n), kz = col0 + line;
      int kc = op.a.comp == 0 ? kx : op.a.comp == 1 ? ky : kz;
      cx<T> r{0, 0};
      if (kc != op.a.n / 2) {
        T f = s * op.a.tab[kx * kx + ky2 + kz * kz] * (T(kc) * op.a.kmin);
        r = cx<T>{v.y * f, -v.x * f};
      }
      cx<T> &dst = ob[o * out_ls + line];
      if (accum) {
        cx<T> prev = dst;
        r = cx<T>{prev.x + r.x, prev.y + r.y};
      }
      dst = r;
      return;
    }
    ob[o * out_ls + line] = cx<T>{v.x * s, v.y * s};
  };
  auto lidx = [](int line, int pt) { return pt * C + line; };
  if constexpr (PRE == PRE_NONE) {
    run_stages<T, L, C, THREADS, DIR, true, false, false, 0, WIDE>(lds, tw, 1, gld, gst, lidx);
  } else {
    run_stages<T, L, C, THREADS, DIR, true, false, false, 0, WIDE>(lds, tw, 1, gld, gst, lidx, op);
  }
}

// One tile per block, tiles in XCD order.
template <typename T, int L, int C, int DIR, int THREADS, int PRE, int POST, int PASS>
__global__ __launch_bounds__(THREADS) void k_fft_strided(StridedArgs<T> a, const cx<T> *__restrict__ tw) {
  extern __shared__ char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  // the L twiddles in LDS after the tile: the stages after the first read
  // them there instead of through the vector memory path (the first stage
  // needs none, and the barrier after it publishes the copy)
  cx<T> *ltw = lds + C * L;
  for (int i = int(threadIdx.x); i < L; i += THREADS) ltw[i] = tw[i];
  TileG<T> g = tile_geom<T, C>(a, xcd_tile(blockIdx.x, gridDim.x));
  const int in_wlen = a.in_wlen;   // read once, not per load
  const long in_ls = a.in_ls;
  strided_tile<T, L, C, DIR, THREADS, PRE, POST>(
      a, ltw, lds, g, [&](int line, int pt) { return tile_in<T, L>(in_wlen, in_ls, g, line, pt); });
}

// ---- z pass, r2c --------------------------------------------------------------
// ROWS rows of N reals (row stride N+2) per block.  Complex FFT of length
// M=N/2 on z_j = x_{2j} + i x_{2j+1}, then X_k = F_k + W^k G_k.
// LDS row of one z-pass line (M + 1 complex): one pad element after every 16
// points breaks the power-of-two strides of the Stockham stores, which
// otherwise put up to 8 lanes of a wavefront on the same bank
template <int M> struct ZRow {
  static constexpr int PAD = M >= 64 ? 16 : 0;
  static constexpr int len = M + 1 + (PAD ? M / PAD : 0);
  __device__ __forceinline__ static int idx(int line, int pt) {
    if constexpr (PAD) return line * len + pt + pt / PAD;
    else return line * len + pt;
  }
};

template <typename T, int N, int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_fft_z_r2c(T *__restrict__ data, long nrows, T scale,
                                                       const cx<T> *__restrict__ twN, const T *__restrict__ src) {
  constexpr int M = N / 2;
  extern __shared__ char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  // the M-point transform's twiddles (every other entry of the N-point
  // table) in LDS after the rows, for the stages after the first (which
  // needs none; the barriers before the second stage publish them)
  cx<T> *ltw = lds + ROWS * ZRow<M>::len;
  for (int i = int(threadIdx.x); i < M; i += THREADS) ltw[i] = twN[2 * i];
  long row0 = long(blockIdx.x) * ROWS;
  cx<T> *rows = reinterpret_cast<cx<T> *>(data) + row0 * (M + 1);
  // rows come from the padded field itself, or from a dense array (row stride N reals)
  const cx<T> *srows = src ? reinterpret_cast<const cx<T> *>(src) + row0 * M : rows;
  const long srs = src ? M : M + 1;
  auto gld = [&](int line, int pt) -> cx<T> { return srows[long(line) * srs + pt]; };
  auto lidx = [](int line, int pt) { return ZRow<M>::idx(line, pt); };
  auto lst = [&](int line, int pt, cx<T> v) { lds[lidx(line, pt)] = v; };
  if constexpr (M == 1) {
    if (threadIdx.x < ROWS) lds[lidx(threadIdx.x, 0)] = gld(threadIdx.x, 0);
  } else {
    run_stages<T, M, ROWS, THREADS, -1, false, false, true, 0, false>(lds, ltw, 1, gld, lst, lidx);
  }
  __syncthreads();
  constexpr int H = M / 2 + 1;
  for (int q = threadIdx.x; q < ROWS * H; q += THREADS) {
    int line = q / H, k = q - line * H;
    cx<T> zk = lds[lidx(line, k % M)];
    cx<T> zm = lds[lidx(line, (M - k) % M)];
    cx<T> F{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
    // G = -i (Zk - conj Zm)/2
    cx<T> D{(zk.x - zm.x) * T(0.5), (zk.y + zm.y) * T(0.5)};
    cx<T> G{D.y, -D.x};
    cx<T> w = twN[k];
    cx<T> wg = cmul(w, G);
    cx<T> Xk{(F.x + wg.x) * scale, (F.y + wg.y) * scale};
    cx<T> Xm{(F.x - wg.x) * scale, -(F.y - wg.y) * scale};
    cx<T> *out = rows + long(line) * (M + 1);
    out[k] = Xk;
    if (M - k != k) out[M - k] = Xm;
  }
}

// ---- z pass, c2r --------------------------------------------------------------
// EPI selects what the pass does with its real output v (z fastest, padded
// rows): 0 stores it (in place, or to dst as a dense array); for the 2LPT
// source, with S and A padded fields of the same layout:
//   1: S = A v, A = A + v     2: S = S + A v     3: S = S - v^2
template <typename T, int N, int ROWS, int THREADS, int EPI = 0>
__global__ __launch_bounds__(THREADS) void k_fft_z_c2r(T *__restrict__ data, long nrows, T scale,
                                                       const cx<T> *__restrict__ twN, T *__restrict__ dst,
                                                       T *__restrict__ S = nullptr, T *__restrict__ A = nullptr) {
  constexpr int M = N / 2;
  extern __shared__ char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  // the M-point transform's twiddles (every other entry of the N-point
  // table) in LDS after the rows, for the stages after the first (which
  // needs none; the barriers before the second stage publish them)
  cx<T> *ltw = lds + ROWS * ZRow<M>::len;
  for (int i = int(threadIdx.x); i < M; i += THREADS) ltw[i] = twN[2 * i];
  long row0 = long(blockIdx.x) * ROWS;
  cx<T> *rows = reinterpret_cast<cx<T> *>(data) + row0 * (M + 1);
  auto lidx = [](int line, int pt) { return ZRow<M>::idx(line, pt); };
  constexpr int H = M / 2 + 1;
  for (int q = threadIdx.x; q < ROWS * H; q += THREADS) {
    int line = q / H, k = q - line * H;
    const cx<T> *in = rows + long(line) * (M + 1);
    cx<T> xk = in[k], xm = in[M - k];
    if (k == 0) {
      // c2r ignores the imaginary parts of the kz=0 and kz=N/2 entries
      xk.y = 0;
      xm.y = 0;
    }
    cx<T> E{xk.x + xm.x, xk.y - xm.y};
    cx<T> D{xk.x - xm.x, xk.y + xm.y};
    cx<T> w = twN[k];
    w.y = -w.y;                       // e^{+2 pi i k/N}
    cx<T> O = cmul(D, w);
    cx<T> Zk{E.x - O.y, E.y + O.x};   // E + i O
    cx<T> Zm{E.x + O.y, -E.y + O.x};  // conj(E) + i conj(O)
    if (k == 0) {
      lds[lidx(line, 0)] = Zk;
    } else {
      lds[lidx(line, k)] = Zk;
      if (M - k != k) lds[lidx(line, M - k)] = Zm;
    }
  }
  __syncthreads();
  auto lld = [&](int line, int pt) -> cx<T> { return lds[lidx(line, pt)]; };
  // results go back into the padded rows, or to a dense array (row stride N reals)
  cx<T> *drows = dst ? reinterpret_cast<cx<T> *>(dst) + row0 * M : rows;
  const long drs = dst ? M : M + 1;
  auto gst = [&](int line, int pt, cx<T> v) {
    cx<T> r{v.x * scale, v.y * scale};
    if constexpr (EPI == 0) {
      drows[long(line) * drs + pt] = r;
    } else {
      long q = (row0 + line) * (M + 1) + pt;
      cx<T> *Sc = reinterpret_cast<cx<T> *>(S), *Ac = reinterpret_cast<cx<T> *>(A);
      if constexpr (EPI == 1) {
        cx<T> a = Ac[q];
        Sc[q] = cx<T>{a.x * r.x, a.y * r.y};
        Ac[q] = cx<T>{a.x + r.x, a.y + r.y};
      } else if constexpr (EPI == 2) {
        cx<T> a = Ac[q], sv = Sc[q];
        Sc[q] = cx<T>{sv.x + a.x * r.x, sv.y + a.y * r.y};
      } else {
        cx<T> sv = Sc[q];
        Sc[q] = cx<T>{sv.x - r.x * r.x, sv.y - r.y * r.y};
      }
    }
  };
  if constexpr (M == 1) {
    if (threadIdx.x < ROWS) gst(threadIdx.x, 0, lld(threadIdx.x, 0));
  } else {
    run_stages<T, M, ROWS, THREADS, +1, false, true, false, 0, false>(lds, ltw, 1, lld, gst, lidx);
  }
}

// ---- z pass, c2r into a window ------------------------------------------------
// Same transform as k_fft_z_c2r, reading complex rows from `in` (row stride
// M+1 complex) and writing only z in [z0, z0+wz) (mod N) of each real row to
// `out` (row stride wz).
template <typename T, int N, int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_fft_z_c2r_window(const T *__restrict__ in, T *__restrict__ out,
                                                              long nrows, int z0, int wz, T scale,
                                                              const cx<T> *__restrict__ twN) {
  constexpr int M = N / 2;
  extern __shared__ char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  // the M-point transform's twiddles (every other entry of the N-point
  // table) in LDS after the rows, for the stages after the first (which
  // needs none; the barriers before the second stage publish them)
  cx<T> *ltw = lds + ROWS * ZRow<M>::len;
  for (int i = int(threadIdx.x); i < M; i += THREADS) ltw[i] = twN[2 * i];
  long row0 = long(blockIdx.x) * ROWS;
  const cx<T> *rows = reinterpret_cast<const cx<T> *>(in) + row0 * (M + 1);
  auto lidx = [](int line, int pt) { return ZRow<M>::idx(line, pt); };
  constexpr int H = M / 2 + 1;
  for (int q = threadIdx.x; q < ROWS * H; q += THREADS) {
    int line = q / H, k = q - line * H;
    if (row0 + line >= nrows) continue;
    const cx<T> *rin = rows + long(line) * (M + 1);
    cx<T> xk = rin[k], xm = rin[M - k];
    if (k == 0) {
      xk.y = 0;
      xm.y = 0;
    }
    cx<T> E{xk.x + xm.x, xk.y - xm.y};
    cx<T> D{xk.x - xm.x, xk.y + xm.y};
    cx<T> w = twN[k];
    w.y = -w.y;
    cx<T> O = cmul(D, w);
    cx<T> Zk{E.x - O.y, E.y + O.x};
    cx<T> Zm{E.x + O.y, -E.y + O.x};
    if (k == 0) {
      lds[lidx(line, 0)] = Zk;
    } else {
      lds[lidx(line, k)] = Zk;
      if (M - k != k) lds[lidx(line, M - k)] = Zm;
    }
  }
  __syncthreads();
  auto lld = [&](int line, int pt) -> cx<T> { return lds[lidx(line, pt)]; };
  auto gst = [&](int line, int pt, cx<T> v) {
    if (row0 + line >= nrows) return;
    T *orow = out + (row0 + line) * wz;
    // complex point pt holds real samples z = 2pt, 2pt+1
    for (int h = 0; h < 2; ++h) {
      int z = 2 * pt + h - z0;
      if (z < 0) z += N;
      if (z < wz) orow[z] = (h ? v.y : v.x) * scale;
    }
  };
  if constexpr (M == 1) {
    if (threadIdx.x < ROWS) gst(threadIdx.x, 0, lld(threadIdx.x, 0));
  } else {
    run_stages<T, M, ROWS, THREADS, +1, false, true, false, 0, false>(lds, ltw, 1, lld, gst, lidx);
  }
}

// ---- z pass, r2c from a window ------------------------------------------------
// Same transform as k_fft_z_r2c for rows whose real samples are zero except
// z in [z0, z0+wz) (mod N), read from `in` (row stride wz); the spectra go to
// `out` (row stride M+1 complex).
template <typename T, int N, int ROWS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_fft_z_r2c_window(const T *__restrict__ in, T *__restrict__ out,
                                                              long nrows, int z0, int wz, T scale,
                                                              const cx<T> *__restrict__ twN) {
  constexpr int M = N / 2;
  extern __shared__ char smem_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
  // the M-point transform's twiddles (every other entry of the N-point
  // table) in LDS after the rows, for the stages after the first (which
  // needs none; the barriers before the second stage publish them)
  cx<T> *ltw = lds + ROWS * ZRow<M>::len;
  for (int i = int(threadIdx.x); i < M; i += THREADS) ltw[i] = twN[2 * i];
  long row0 = long(blockIdx.x) * ROWS;
  auto gld = [&](int line, int pt) -> cx<T> {
    if (row0 + line >= nrows) return cx<T>{0, 0};
    const T *irow = in + (row0 + line) * wz;
    T v[2];
    for (int h = 0; h < 2; ++h) {
      int z = 2 * pt + h - z0;
      if (z < 0) z += N;
      v[h] = z < wz ? irow[z] : T(0);
    }
    return cx<T>{v[0], v[1]};
  };
  auto lidx = [](int line, int pt) { return ZRow<M>::idx(line, pt); };
  auto lst = [&](int line, int pt, cx<T> v) { lds[lidx(line, pt)] = v; };
  if constexpr (M == 1) {
    if (threadIdx.x < ROWS) lds[lidx(threadIdx.x, 0)] = gld(threadIdx.x, 0);
  } else {
    run_stages<T, M, ROWS, THREADS, -1, false, false, true, 0, false>(lds, ltw, 1, gld, lst, lidx);
  }
  __syncthreads();
  constexpr int H = M / 2 + 1;
  for (int q = threadIdx.x; q < ROWS * H; q += THREADS) {
    int line = q / H, k = q - line * H;
    if (row0 + line >= nrows) continue;
    cx<T> zk = lds[lidx(line, k % M)];
    cx<T> zm = lds[lidx(line, (M - k) % M)];
    cx<T> F{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
    cx<T> D{(zk.x - zm.x) * T(0.5), (zk.y + zm.y) * T(0.5)};
    cx<T> G{D.y, -D.x};
    cx<T> w = twN[k];
    cx<T> wg = cmul(w, G);
    cx<T> Xk{(F.x + wg.x) * scale, (F.y + wg.y) * scale};
    cx<T> Xm{(F.x - wg.x) * scale, -(F.y - wg.y) * scale};
    cx<T> *orow = reinterpret_cast<cx<T> *>(out) + (row0 + line) * (M + 1);
    orow[k] = Xk;
    if (M - k != k) orow[M - k] = Xm;
  }
}

// ---- Hermitian mirror of the kz=0 and kz=N/2 planes ------------------------------
// ensureFourierModesAreMirrored (fourier.hpp:331-357): for kx in [0,N/2] and
// ky from N/2 down to -N/2+1 the conjugate of mode (kx,ky) is copied into
// (-kx,-ky).  Equivalent closed form: a destination (kx',ky') with kx' > N/2,
// or kx' in {0,N/2} and ky' > N/2 receives conj(mode(-kx',-ky')); the
// self-conjugate modes get their imaginary part negated.
template <typename T>
__global__ void k_mirror_planes(cx<T> *data, int n) {
  int nz = n / 2 + 1;
  long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= long(n) * n) return;
  int kx = int(t / n), ky = int(t % n);
  int h = n / 2;
  int sx = (n - kx) % n, sy = (n - ky) % n;
  bool xedge = (kx == 0 || kx == h);
  bool dest = kx > h || (xedge && ky > h);
  bool self = xedge && (ky == 0 || ky == h);
  for (int kz : {0, h}) {
    if (dest) {
      cx<T> s = data[(long(sx) * n + sy) * nz + kz];
      data[(long(kx) * n + ky) * nz + kz] = cx<T>{s.x, -s.y};
    } else if (self) {
      cx<T> &v = data[(long(kx) * n + ky) * nz + kz];
      v.y = -v.y;
    }
  }
}

// ----------------------------------------------------------------------------------
template <typename T> struct TwTable {
  std::map<int, cx<T> *> tabs;
  std::mutex mu;
  const cx<T> *get(int n) {
    std::lock_guard<std::mutex> g(mu);
    auto it = tabs.find(n);
    if (it != tabs.end()) return it->second;
    std::vector<cx<T>> h(n);
    for (int k = 0; k < n; ++k) {
      // FP64-generated twiddles, rounded to T
      double a = -2.0 * M_PI * double(k) / double(n);
      h[k] = cx<T>{T(cos(a)), T(sin(a))};
    }
    cx<T> *d = nullptr;
    GIC_HIP(hipMalloc(&d, sizeof(cx<T>) * n));
    GIC_HIP(hipMemcpy(d, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    tabs[n] = d;
    return d;
  }
};
template <typename T> TwTable<T> &twtable() {
  static TwTable<T> t;
  return t;
}

// Tile shape of the strided passes: the tuned table, or one shape for every
// size when a tuning build names it (GIC_TUNE_THREADS, GIC_TUNE_COLS).
template <typename T, int N> constexpr int strided_threads() {
#ifdef GIC_TUNE_THREADS
  return GIC_TUNE_THREADS;
#else
  return TileShape<T, N>::threads;
#endif
}
template <typename T, int N> constexpr int strided_cols() {
#ifdef GIC_TUNE_COLS
  return sizeof(T) == 4 ? GIC_TUNE_COLS : (GIC_TUNE_COLS > 8 ? 8 : GIC_TUNE_COLS);
#else
  return TileShape<T, N>::cols;
#endif
}

// y passes' tile shape: YTileShape where fft_tiles.h gives one, else the
// x passes'
template <typename T, int N, class = void> struct YShape {
  static constexpr int threads = strided_threads<T, N>();
  static constexpr int cols = strided_cols<T, N>();
};
#ifndef GIC_TUNE_THREADS   // the tuner times one shape for every pass
template <typename T, int N> struct YShape<T, N, std::void_t<decltype(YTileShape<T, N>::cols)>> {
  static constexpr int threads = YTileShape<T, N>::threads;
  static constexpr int cols = YTileShape<T, N>::cols;
};
#endif

template <typename T, int N> struct FftN {
  static constexpr int THREADS = 256;
  static constexpr int C = strided_cols<T, N>();
  static constexpr int M = N / 2;
  static constexpr int NZ = N / 2 + 1;
  // z-pass rows per block: 8192 / N points' worth; at N = 1024 half that, so
  // that the rows' LDS (about 21 KB) admits seven blocks per CU instead of four
  static constexpr int ROWS0 = N >= 1024 ? 4096 / N : (8192 / N) < 1 ? 1 : (8192 / N);
  static constexpr int ROWS = ROWS0 > N * N ? N * N : ROWS0;

  static constexpr int STHREADS = strided_threads<T, N>();
  template <int DIR, int PRE, int POST, int PASS = 0>
  static void strided(const StridedArgs<T> &a0, int nouter, hipStream_t st, TimingHook *hook, const char *name) {
    constexpr bool YP = PASS == 1 || PASS == 3 || PASS == 5;
    constexpr int CP = YP ? YShape<T, N>::cols : C;
    constexpr int TP = YP ? YShape<T, N>::threads : STHREADS;
    StridedArgs<T> a = a0;
    a.ff.nf = -1;
    if (sizeof(T) == 4) {
      if (a.fil.nf >= 0) {
        if (a.fil.nf > FILTER_INLINE_MAX) throw std::runtime_error("FilterInline: at most 4 factors");
        a.ff.nf = a.fil.nf;
        a.ff.zero = a.fil.zero;
        a.ff.over_k2 = a.fil.over_k2;
        for (int i = 0; i < a.fil.nf; ++i) {
          a.ff.kind[i] = a.fil.kind[i];
          a.ff.kc[i] = float(a.fil.kc[i]);
          a.ff.inv_w[i] = 1.f / (float(a.fil.kc[i]) * 0.1f);
        }
        a.ff.scale = float(a.fil.scale);
        a.ff.kmin = float(a.kmin_d);
        a.ff.inv_kmin2 = a.kmin_d > 0 ? float(1.0 / (a.kmin_d * a.kmin_d)) : 0.f;
      }
      a.inv_k2_f = a.kmin_d > 0 ? float(a.inv_k2 / (a.kmin_d * a.kmin_d)) : 0.f;
    }
    a.ncol = NZ;
    a.nkzb = (NZ + CP - 1) / CP;
    a.n = N;
    a.nouter = nouter;
    int nblk = nouter * a.nkzb * (a.nbatch > 1 ? a.nbatch : 1);
    size_t shm = sizeof(cx<T>) * (CP + 1) * N;   // the tile, then the line's twiddles
    const cx<T> *tw = twtable<T>().get(N);
    auto kern = k_fft_strided<T, N, CP, DIR, TP, PRE, POST, PASS>;
    if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    std::string tag = std::string(name) + "." + std::to_string(N);
    if (hook) {
      hook->begin(tag.c_str(), st);
      // lines of the pass; a pruned input or output moves only its window
      double lines = double(nouter) * NZ * (a.nbatch > 1 ? a.nbatch : 1);
      double rd = lines * ((a.in_wlen > 0 && a.in_wlen < N) ? a.in_wlen : N);
      double wr = lines * (a.wlen < N ? a.wlen : N);
      bool reads_out = a.accum || POST == POST_AUX;   // accumulated output, or the spectrum it adds
      hook->bytes(tag.c_str(), sizeof(cx<T>) * (rd + wr + (reads_out ? wr : 0.0)));
    }
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(TP), shm, st, a, tw);
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
  }
  // in-place plain pass over a full field: y (line stride NZ) or x (line stride N*NZ)
  static StridedArgs<T> plain(cx<T> *d, bool xpass, T scale) {
    StridedArgs<T> a{};
    a.in = d;
    a.out = d;
    a.in_ls = a.out_ls = xpass ? long(N) * NZ : NZ;
    a.in_os = a.out_os = xpass ? NZ : long(N) * NZ;
    a.tab = nullptr;
    a.comp = 0;
    a.kmin = 0;
    a.win0 = 0;
    a.wlen = N;
    a.in_win0 = 0;
    a.in_wlen = N;
    a.scale = scale;
    return a;
  }
  // z rows of the padded field d; ext (if non-null) is the dense N^3 array the
  // r2c reads its rows from, or the c2r writes its rows to
  static void zpass(T *d, T scale, int dir, hipStream_t st, TimingHook *hook, const char *name, T *ext = nullptr,
                    int nfields = 1) {
    long nrows = long(N) * N * nfields;
    int nblk = int(nrows / ROWS);
    size_t shm = sizeof(cx<T>) * (ROWS * ZRow<M>::len + M);   // rows, then the M-point twiddles
    const cx<T> *tw = twtable<T>().get(N);
    std::string tag = std::string(name) + "." + std::to_string(N);
    if (hook) hook->bytes(tag.c_str(), double(nrows) * (2 * N + 2) * sizeof(T));   // N reals one way, N+2 the other
    if (dir < 0) {
      auto kern = k_fft_z_r2c<T, N, ROWS, THREADS>;
      if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      if (hook) hook->begin(tag.c_str(), st);
      hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, d, nrows, scale, tw, (const T *)ext);
    } else {
      auto kern = k_fft_z_c2r<T, N, ROWS, THREADS>;
      if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      if (hook) hook->begin(tag.c_str(), st);
      hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, d, nrows, scale, tw, ext, (T *)nullptr,
                         (T *)nullptr);
    }
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
  }
  static T norm() { return T(1.0 / pow(double(N), 1.5)); }
  static void mirror(cx<T> *c, hipStream_t st, TimingHook *hook) {
    int nt = 256;
    long nn = long(N) * N;
    std::string tag = "fft_mirror." + std::to_string(N);
    if (hook) {
      hook->begin(tag.c_str(), st);
      hook->bytes(tag.c_str(), 2.0 * nn * sizeof(cx<T>));   // half of two planes read, half written
    }
    hipLaunchKernelGGL(k_mirror_planes<T>, dim3((nn + nt - 1) / nt), dim3(nt), 0, st, c, N);
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
  }
  static void r2c(T *d, const T *post, hipStream_t st, TimingHook *hook, const T *src = nullptr) {
    cx<T> *c = reinterpret_cast<cx<T> *>(d);
    zpass(d, T(1), -1, st, hook, "fft_z_r2c", const_cast<T *>(src));
    strided<-1, PRE_NONE, POST_NONE, 1>(plain(c, false, T(1)), N, st, hook, "fft_y_fwd");
    StridedArgs<T> ax = plain(c, true, norm());
    if (post) {
      ax.tab = post;
      strided<-1, PRE_NONE, POST_TABLE>(ax, N, st, hook, "fft_x_post");
    } else {
      strided<-1, PRE_NONE, POST_NONE>(ax, N, st, hook, "fft_x_fwd");
    }
  }
  // r2c of nb padded fields stored back to back (no dense source), the
  // post table applied by the last pass of each
  static void r2c_batch(T *d, int nb, const T *post, hipStream_t st, TimingHook *hook) {
    cx<T> *c = reinterpret_cast<cx<T> *>(d);
    long fs = long(N) * N * NZ;
    zpass(d, T(1), -1, st, hook, "fft_z_r2c_batch", nullptr, nb);
    StridedArgs<T> ay = plain(c, false, T(1));
    ay.nbatch = nb;
    ay.in_bs = ay.out_bs = fs;
    strided<-1, PRE_NONE, POST_NONE, 1>(ay, N, st, hook, "fft_y_fwd_batch");
    StridedArgs<T> ax = plain(c, true, norm());
    ax.nbatch = nb;
    ax.in_bs = ax.out_bs = fs;
    if (post) {
      ax.tab = post;
      strided<-1, PRE_NONE, POST_TABLE>(ax, N, st, hook, "fft_x_post_batch");
    } else {
      strided<-1, PRE_NONE, POST_NONE>(ax, N, st, hook, "fft_x_fwd_batch");
    }
  }
  // r2c of nb fields that are zero outside a box of wl x wl cells in (x, y)
  // (all z), box b starting at (winx[b], winy[b]) (mod N; device arrays).
  // rows: the boxes' rows, nrows (a multiple of ROWS) dense rows of N reals,
  // row (b wl + i) wl + j holding x = winx[b]+i, y = winy[b]+j; zbuf (nrows
  // rows of NZ complex) and ybuf (nb wl N NZ complex) are scratch.  Field b's
  // spectrum, times post[|k|^2], goes to spec + b * spec_stride (complex).
  static void r2c_boxes(const T *rows, long nrows, int nb, int wl, const int *winx, const int *winy, T *zbuf,
                        T *ybuf, T *spec, long spec_stride, const T *post, hipStream_t st, TimingHook *hook) {
    int nblk = int(nrows / ROWS);
    size_t shm = sizeof(cx<T>) * (ROWS * ZRow<M>::len + M);   // rows, then the M-point twiddles
    const cx<T> *tw = twtable<T>().get(N);
    auto kern = k_fft_z_r2c<T, N, ROWS, THREADS>;
    if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    std::string tag = "fft_z_r2c_boxes." + std::to_string(N);
    if (hook) {
      hook->begin(tag.c_str(), st);
      hook->bytes(tag.c_str(), double(nrows) * (2 * N + 2) * sizeof(T));
    }
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, zbuf, nrows, T(1), tw, rows);
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
    // y pass: per field, tiles (i, kz block), lines over ky from the wl rows j
    StridedArgs<T> ay{};
    ay.in = reinterpret_cast<const cx<T> *>(zbuf);
    ay.out = reinterpret_cast<cx<T> *>(ybuf);
    ay.in_ls = NZ;
    ay.in_os = long(wl) * NZ;
    ay.out_ls = NZ;
    ay.out_os = long(N) * NZ;
    ay.win0 = 0;
    ay.wlen = N;
    ay.in_win0 = 0;
    ay.in_wlen = wl;
    ay.scale = T(1);
    ay.nbatch = nb;
    ay.in_bs = long(wl) * wl * NZ;
    ay.out_bs = long(wl) * N * NZ;
    ay.in_win0_b = winy;
    strided<-1, PRE_NONE, POST_NONE, 5>(ay, wl, st, hook, "fft_y_fwd_boxes");
    // x pass: per field, tiles (ky, kz block), lines over kx from the wl planes i
    StridedArgs<T> ax{};
    ax.in = reinterpret_cast<const cx<T> *>(ybuf);
    ax.out = reinterpret_cast<cx<T> *>(spec);
    ax.in_ls = long(N) * NZ;
    ax.in_os = NZ;
    ax.out_ls = long(N) * NZ;
    ax.out_os = NZ;
    ax.win0 = 0;
    ax.wlen = N;
    ax.in_win0 = 0;
    ax.in_wlen = wl;
    ax.scale = norm();
    ax.tab = post;
    ax.nbatch = nb;
    ax.in_bs = long(wl) * N * NZ;
    ax.out_bs = spec_stride;
    ax.in_win0_b = winx;
    if (post) strided<-1, PRE_NONE, POST_TABLE, 4>(ax, N, st, hook, "fft_x_fwd_boxes");
    else strided<-1, PRE_NONE, POST_NONE, 4>(ax, N, st, hook, "fft_x_fwd_boxes");
  }
  static void c2r(T *d, hipStream_t st, TimingHook *hook, T *dst = nullptr) {
    cx<T> *c = reinterpret_cast<cx<T> *>(d);
    mirror(c, st, hook);
    strided<+1, PRE_NONE, POST_NONE>(plain(c, true, T(1)), N, st, hook, "fft_x");
    strided<+1, PRE_NONE, POST_NONE, 1>(plain(c, false, T(1)), N, st, hook, "fft_y");
    zpass(d, norm(), +1, st, hook, "fft_z_c2r", dst);
  }
  // adjoint of one Zel'dovich component: r2c whose x pass adds
  // conj(i k_c tab[|k|^2]) times each mode into acc (or stores it when !accum)
  static void r2c_za_adjoint(const T *src, T *d, T *acc, const T *tab, int comp, T kmin, bool accum, hipStream_t st,
                             TimingHook *hook) {
    cx<T> *c = reinterpret_cast<cx<T> *>(d);
    zpass(d, T(1), -1, st, hook, "fft_z_r2c", const_cast<T *>(src));
    strided<-1, PRE_NONE, POST_NONE, 1>(plain(c, false, T(1)), N, st, hook, "fft_y_fwd");
    StridedArgs<T> ax = plain(c, true, norm());
    ax.out = reinterpret_cast<cx<T> *>(acc);
    ax.tab = tab;
    ax.comp = comp;
    ax.kmin = kmin;
    ax.accum = accum ? 1 : 0;
    strided<-1, PRE_NONE, POST_ZA_ADJ>(ax, N, st, hook, "fft_x_za_adj");
  }
  // out <- c2r(op(spec)) with op applied while the x pass reads the spectrum
  static void c2r_from(const T *spec, T *out, int pre, const T *tab, int comp, double kmin, hipStream_t st,
                       TimingHook *hook, T *dst = nullptr, double inv_k2 = 0, const FilterInline *fil = nullptr) {
    const cx<T> *ci = reinterpret_cast<const cx<T> *>(spec);
    cx<T> *co = reinterpret_cast<cx<T> *>(out);
    StridedArgs<T> ax = plain(co, true, T(1));
    ax.in = ci;
    ax.tab = tab;
    ax.comp = comp;
    ax.kmin = T(kmin);
    ax.kmin_d = kmin;
    ax.inv_k2 = inv_k2;
    if (fil) ax.fil = *fil;
    if (!tab && pre != PRE_NONE && inv_k2 == 0) throw std::runtime_error("c2r_from: operator without a table");
    if (pre == PRE_ZA) strided<+1, PRE_ZA, POST_NONE>(ax, N, st, hook, "fft_x_za");
    else if (pre == PRE_KK) strided<+1, PRE_KK, POST_NONE>(ax, N, st, hook, "fft_x_kk");
    else if (pre == PRE_TABLE) strided<+1, PRE_TABLE, POST_NONE>(ax, N, st, hook, "fft_x_pre");
    else strided<+1, PRE_NONE, POST_NONE>(ax, N, st, hook, "fft_x");
    strided<+1, PRE_NONE, POST_NONE, 1>(plain(co, false, T(1)), N, st, hook, "fft_y");
    zpass(out, norm(), +1, st, hook, "fft_z_c2r", dst);
  }
  // the x pass of c2r_from alone, into half
  static void xpass_from(const T *spec, T *half, int pre, const T *tab, int comp, double kmin, double inv_k2,
                         hipStream_t st, TimingHook *hook) {
    StridedArgs<T> ax = plain(reinterpret_cast<cx<T> *>(half), true, T(1));
    ax.in = reinterpret_cast<const cx<T> *>(spec);
    ax.tab = tab;
    ax.comp = comp;
    ax.kmin = T(kmin);
    ax.kmin_d = kmin;
    ax.inv_k2 = inv_k2;
    if (!tab && pre != PRE_NONE && inv_k2 == 0) throw std::runtime_error("c2r_xpass: operator without a table");
    if (pre == PRE_ZA) strided<+1, PRE_ZA, POST_NONE>(ax, N, st, hook, "fft_x_za");
    else if (pre == PRE_KK) strided<+1, PRE_KK, POST_NONE>(ax, N, st, hook, "fft_x_kk");
    else if (pre == PRE_TABLE) strided<+1, PRE_TABLE, POST_NONE>(ax, N, st, hook, "fft_x_pre");
    else strided<+1, PRE_NONE, POST_NONE>(ax, N, st, hook, "fft_x");
  }
  // the y pass from half into out, times the FFT_YZ_* factor yop
  static void ypass_from(const T *half, T *out, int yop, double kmin, hipStream_t st, TimingHook *hook) {
    if (yop < FFT_YZ_ONE || yop > FFT_YZ_KZKZ) throw std::runtime_error("c2r_yz: bad y factor");
    StridedArgs<T> ay = plain(reinterpret_cast<cx<T> *>(out), false, T(1));
    ay.in = reinterpret_cast<const cx<T> *>(half);
    ay.comp = yop;
    ay.kmin = T(kmin);
    if (yop == FFT_YZ_ONE) strided<+1, PRE_NONE, POST_NONE, 1>(ay, N, st, hook, "fft_y");
    else strided<+1, PRE_YZ, POST_NONE, 1>(ay, N, st, hook, "fft_y_op");
  }
  static void yz_from(const T *half, T *out, int yop, double kmin, hipStream_t st, TimingHook *hook, T *dst) {
    ypass_from(half, out, yop, kmin, st, hook);
    zpass(out, norm(), +1, st, hook, "fft_z_c2r", dst);
  }
  static void yz_from_epi(const T *half, T *out, int yop, double kmin, int epi, T *S, T *A, hipStream_t st,
                          TimingHook *hook) {
    ypass_from(half, out, yop, kmin, st, hook);
    zpass_epi(out, epi, S, A, st, hook);
  }
  // c2r(op(spec)) whose last pass folds its output into the 2LPT source
  // accumulators S, A (EPI 1..3 of k_fft_z_c2r); out is scratch
  static void c2r_from_epi(const T *spec, T *out, int pre, const T *tab, int comp, double kmin, int epi, T *S, T *A,
                           hipStream_t st, TimingHook *hook, double inv_k2) {
    const cx<T> *ci = reinterpret_cast<const cx<T> *>(spec);
    cx<T> *co = reinterpret_cast<cx<T> *>(out);
    StridedArgs<T> ax = plain(co, true, T(1));
    ax.in = ci;
    ax.tab = tab;
    ax.comp = comp;
    ax.kmin = T(kmin);
    ax.kmin_d = kmin;
    ax.inv_k2 = inv_k2;
    if (!tab && inv_k2 == 0) throw std::runtime_error("c2r_from_epi: operator without a table");
    if (pre == PRE_KK) strided<+1, PRE_KK, POST_NONE>(ax, N, st, hook, "fft_x_kk");
    else throw std::runtime_error("c2r_from_epi: only the k_a k_b operator");
    strided<+1, PRE_NONE, POST_NONE, 1>(plain(co, false, T(1)), N, st, hook, "fft_y");
    zpass_epi(out, epi, S, A, st, hook);
  }
  // the z pass (c2r) of out folded into S, A (EPI 1..3 of k_fft_z_c2r)
  static void zpass_epi(T *out, int epi, T *S, T *A, hipStream_t st, TimingHook *hook) {
    long nrows = long(N) * N;
    int nblk = int(nrows / ROWS);
    size_t shm = sizeof(cx<T>) * (ROWS * ZRow<M>::len + M);   // rows, then the M-point twiddles
    const cx<T> *tw = twtable<T>().get(N);
    std::string tag = "fft_z_c2r_lpt." + std::to_string(N);
    auto launch = [&](auto kern) {
      if (shm > 65536)
        GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      if (hook) {
        hook->begin(tag.c_str(), st);
        // the spectrum rows in; S and A: read A, write S and A (1), read A and
        // S, write S (2), read and write S (3)
        hook->bytes(tag.c_str(), double(nrows) * ((N + 2) + (epi == 3 ? 2 : 3) * N) * sizeof(T));
      }
      hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, out, nrows, norm(), tw, (T *)nullptr, S, A);
      GIC_HIP(hipGetLastError());
      if (hook) hook->end(tag.c_str(), st);
    };
    if (epi == 1) launch(k_fft_z_c2r<T, N, ROWS, THREADS, 1>);
    else if (epi == 2) launch(k_fft_z_c2r<T, N, ROWS, THREADS, 2>);
    else if (epi == 3) launch(k_fft_z_c2r<T, N, ROWS, THREADS, 3>);
    else throw std::runtime_error("c2r_from_epi: bad epilogue");
  }
  // real-space window [w0, w0+wl) (per axis, mod N) of c2r(op(spec)):
  // bufA holds wl[0] x-planes after the x pass, bufB wl[0]*wl[1] rows after
  // the y pass; out receives wl[0]*wl[1]*wl[2] reals.
  static void c2r_window(const T *spec, int pre, const T *tab, int comp, T kmin, const int *w0, const int *wl, T *bufA,
                         T *bufB, T *out, hipStream_t st, TimingHook *hook, const FilterInline *fil = nullptr,
                         int yop = FFT_YZ_ONE, bool xpass = true) {
    const cx<T> *ci = reinterpret_cast<const cx<T> *>(spec);
    cx<T> *ca = reinterpret_cast<cx<T> *>(bufA);
    cx<T> *cb = reinterpret_cast<cx<T> *>(bufB);
    // x pass: tiles (ky, kz block), lines over kx, keep x in window; A = [wl0][N][NZ]
    StridedArgs<T> ax{};
    ax.in = ci;
    ax.out = ca;
    ax.in_ls = long(N) * NZ;
    ax.in_os = NZ;
    ax.out_ls = long(N) * NZ;
    ax.out_os = NZ;
    ax.tab = tab;
    ax.comp = comp;
    ax.kmin = kmin;
    ax.kmin_d = double(kmin);
    if (fil) ax.fil = *fil;
    ax.win0 = w0[0];
    ax.wlen = wl[0];
    ax.scale = T(1);
    if (!xpass) {
      // bufA already holds this x pass (an earlier call on the same spectrum,
      // operator and window)
    } else if (pre == PRE_ZA) {
      strided<+1, PRE_ZA, POST_NONE, 2>(ax, N, st, hook, "fft_x_za_win");
    } else if (pre == PRE_TABLE) {
      strided<+1, PRE_TABLE, POST_NONE, 2>(ax, N, st, hook, "fft_x_pre_win");
    } else {
      strided<+1, PRE_NONE, POST_NONE, 2>(ax, N, st, hook, "fft_x_win");
    }
    // y pass: tiles (x_local, kz block), lines over ky, keep y in window; B = [wl0][wl1][NZ]
    StridedArgs<T> ay{};
    ay.in = ca;
    ay.out = cb;
    ay.in_ls = NZ;
    ay.in_os = long(N) * NZ;
    ay.out_ls = NZ;
    ay.out_os = long(wl[1]) * NZ;
    ay.tab = nullptr;
    ay.win0 = w0[1];
    ay.wlen = wl[1];
    ay.scale = T(1);
    ay.comp = yop;
    ay.kmin = kmin;
    if (yop < FFT_YZ_ONE || yop > FFT_YZ_KZKZ) throw std::runtime_error("c2r_window: bad y factor");
    if (yop == FFT_YZ_ONE) strided<+1, PRE_NONE, POST_NONE, 3>(ay, wl[0], st, hook, "fft_y_win");
    else strided<+1, PRE_YZ, POST_NONE, 3>(ay, wl[0], st, hook, "fft_y_op_win");
    // z pass: rows (x_local, y_local), keep z in window
    long nrows = long(wl[0]) * wl[1];
    int nblk = int((nrows + ROWS - 1) / ROWS);
    size_t shm = sizeof(cx<T>) * (ROWS * ZRow<M>::len + M);   // rows, then the M-point twiddles
    const cx<T> *tw = twtable<T>().get(N);
    auto kern = k_fft_z_c2r_window<T, N, ROWS, THREADS>;
    if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    std::string tag = "fft_z_c2r_win." + std::to_string(N);
    if (hook) {
      hook->begin(tag.c_str(), st);
      hook->bytes(tag.c_str(), double(nrows) * ((N + 2) + wl[2]) * sizeof(T));
    }
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, bufB, out, nrows, w0[2], wl[2], norm(), tw);
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
  }
  // spectrum of a field that is zero outside the real-space window
  // [w0, w0+wl) (per axis, mod N), given as the dense window `win`
  // (wl[0]*wl[1]*wl[2] reals): pruned z and y passes, a full x pass whose
  // output is scaled by post[|k|^2] (if given) and stored or added into spec.
  // bufB holds wl[0]*wl[1] padded rows, bufA wl[0]*N*(N/2+1) complex.
  static void r2c_window(const T *win, const int *w0, const int *wl, T *bufB, T *bufA, T *spec, const T *post,
                         bool accum, hipStream_t st, TimingHook *hook, const T *aux = nullptr, const FilterInline *fil = nullptr,
                         double kmin = 0) {
    long nrows = long(wl[0]) * wl[1];
    int nblk = int((nrows + ROWS - 1) / ROWS);
    size_t shm = sizeof(cx<T>) * (ROWS * ZRow<M>::len + M);   // rows, then the M-point twiddles
    const cx<T> *tw = twtable<T>().get(N);
    auto kern = k_fft_z_r2c_window<T, N, ROWS, THREADS>;
    if (shm > 65536) GIC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    std::string tag = "fft_z_r2c_win." + std::to_string(N);
    if (hook) {
      hook->begin(tag.c_str(), st);
      hook->bytes(tag.c_str(), double(nrows) * ((N + 2) + wl[2]) * sizeof(T));
    }
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(THREADS), shm, st, win, bufB, nrows, w0[2], wl[2], T(1), tw);
    GIC_HIP(hipGetLastError());
    if (hook) hook->end(tag.c_str(), st);
    // y pass: tiles (x_local, kz block), lines over ky from the wl[1] window rows
    StridedArgs<T> ay{};
    ay.in = reinterpret_cast<const cx<T> *>(bufB);
    ay.out = reinterpret_cast<cx<T> *>(bufA);
    ay.in_ls = NZ;
    ay.in_os = long(wl[1]) * NZ;
    ay.out_ls = NZ;
    ay.out_os = long(N) * NZ;
    ay.win0 = 0;
    ay.wlen = N;
    ay.in_win0 = w0[1];
    ay.in_wlen = wl[1];
    ay.scale = T(1);
    strided<-1, PRE_NONE, POST_NONE, 5>(ay, wl[0], st, hook, "fft_y_fwd_win");
    // x pass: tiles (ky, kz block), lines over kx from the wl[0] window planes
    StridedArgs<T> ax{};
    ax.in = reinterpret_cast<const cx<T> *>(bufA);
    ax.out = reinterpret_cast<cx<T> *>(spec);
    ax.in_ls = long(N) * NZ;
    ax.in_os = NZ;
    ax.out_ls = long(N) * NZ;
    ax.out_os = NZ;
    ax.win0 = 0;
    ax.wlen = N;
    ax.in_win0 = w0[0];
    ax.in_wlen = wl[0];
    ax.scale = norm();
    ax.tab = post;
    ax.accum = accum ? 1 : 0;
    ax.aux = reinterpret_cast<const cx<T> *>(aux);
    if (fil) ax.fil = *fil;
    ax.kmin_d = kmin;
    if (aux) strided<-1, PRE_NONE, POST_AUX, 4>(ax, N, st, hook, "fft_x_fwd_win");
    else if (post) strided<-1, PRE_NONE, POST_TABLE, 4>(ax, N, st, hook, "fft_x_fwd_win");
    else strided<-1, PRE_NONE, POST_NONE, 4>(ax, N, st, hook, "fft_x_fwd_win");
  }
};

#define GIC_FFT_SWITCH(n, CALL)                                                                     \
  switch (n) {                                                                                      \
    case 4: { using F = FftN<T, 4>; CALL; } break;                                                  \
    case 8: { using F = FftN<T, 8>; CALL; } break;                                                  \
    case 16: { using F = FftN<T, 16>; CALL; } break;                                                \
    case 32: { using F = FftN<T, 32>; CALL; } break;                                                \
    case 64: { using F = FftN<T, 64>; CALL; } break;                                                \
    case 128: { using F = FftN<T, 128>; CALL; } break;                                              \
    case 256: { using F = FftN<T, 256>; CALL; } break;                                              \
    case 512: { using F = FftN<T, 512>; CALL; } break;                                              \
    case 1024: { using F = FftN<T, 1024>; CALL; } break;                                            \
    default:                                                                                        \
      throw std::runtime_error("FFT size " + std::to_string(n) + " is not supported (powers of two 4..1024)"); \
  }

template <typename T> void fft3d(T *d, int n, int dir, hipStream_t st, TimingHook *hook) {
  if (dir < 0) {
    GIC_FFT_SWITCH(n, F::r2c(d, nullptr, st, hook))
  } else {
    GIC_FFT_SWITCH(n, F::c2r(d, st, hook))
  }
}
template <typename T> void fft_r2c_post(T *d, int n, const T *post, hipStream_t st, TimingHook *hook, const T *src) {
  GIC_FFT_SWITCH(n, F::r2c(d, post, st, hook, src))
}
template <typename T> void fft_r2c_batch(T *d, int n, int nb, const T *post, hipStream_t st, TimingHook *hook) {
  if (nb < 1) throw std::runtime_error("fft_r2c_batch: no fields");
  GIC_FFT_SWITCH(n, F::r2c_batch(d, nb, post, st, hook))
}
template <typename T>
void fft_r2c_boxes(const T *rows, int n, long nrows, int nb, int wl, const int *winx, const int *winy, T *zbuf, T *ybuf,
                   T *spec, long spec_stride, const T *post, hipStream_t st, TimingHook *hook) {
  if (nb < 1 || wl < 1 || wl >= n) throw std::runtime_error("fft_r2c_boxes: bad box count or size");
  if (nrows < long(nb) * wl * wl || nrows % fft_box_rows_multiple<T>(n)) throw std::runtime_error("fft_r2c_boxes: bad row count");
  GIC_FFT_SWITCH(n, F::r2c_boxes(rows, nrows, nb, wl, winx, winy, zbuf, ybuf, spec, spec_stride, post, st, hook))
}
template <typename T> int fft_box_rows_multiple(int n) {
  int r = 0;
  GIC_FFT_SWITCH(n, r = F::ROWS)
  return r;
}
template <typename T> void fft_prepare(int n) {
  if (n < 4 || n > 1024 || (n & (n - 1))) throw std::runtime_error("FFT size " + std::to_string(n) + " is not supported");
  twtable<T>().get(n);
}
template <typename T> void fft_c2r_to(T *d, T *dst, int n, hipStream_t st, TimingHook *hook) {
  GIC_FFT_SWITCH(n, F::c2r(d, st, hook, dst))
}
template <typename T>
void fft_r2c_za_adjoint(const T *src, T *d, T *acc, int n, const T *tab, int comp, double kmin, bool accumulate,
                        hipStream_t st, TimingHook *hook) {
  GIC_FFT_SWITCH(n, F::r2c_za_adjoint(src, d, acc, tab, comp, T(kmin), accumulate, st, hook))
}
template <typename T>
void fft_c2r_from(const T *spec, T *out, int n, int pre, const T *tab, int comp, double kmin, hipStream_t st,
                  TimingHook *hook, T *dst, double inv_k2, const FilterInline *fil) {
  GIC_FFT_SWITCH(n, F::c2r_from(spec, out, pre, tab, comp, kmin, st, hook, dst, inv_k2, fil))
}
template <typename T>
void fft_c2r_lpt(const T *spec, T *scratch, int n, const T *tab, int comp, double kmin, int epi, T *S, T *A,
                 hipStream_t st, TimingHook *hook, double inv_k2) {
  GIC_FFT_SWITCH(n, F::c2r_from_epi(spec, scratch, FFT_PRE_KK, tab, comp, kmin, epi, S, A, st, hook, inv_k2))
}
template <typename T>
void fft_c2r_xpass(const T *spec, T *half, int n, int pre, const T *tab, int comp, double kmin, hipStream_t st,
                   TimingHook *hook, double inv_k2) {
  GIC_FFT_SWITCH(n, F::xpass_from(spec, half, pre, tab, comp, kmin, inv_k2, st, hook))
}
template <typename T>
void fft_c2r_yz(const T *half, T *out, int n, int yop, double kmin, hipStream_t st, TimingHook *hook, T *dst) {
  GIC_FFT_SWITCH(n, F::yz_from(half, out, yop, kmin, st, hook, dst))
}
template <typename T>
void fft_c2r_yz_lpt(const T *half, T *scratch, int n, int yop, double kmin, int epi, T *S, T *A, hipStream_t st,
                    TimingHook *hook) {
  GIC_FFT_SWITCH(n, F::yz_from_epi(half, scratch, yop, kmin, epi, S, A, st, hook))
}
template <typename T>
void fft_r2c_window(const T *win, int n, const int *w0, const int *wl, T *bufB, T *bufA, T *spec, const T *post,
                    bool accumulate, hipStream_t st, TimingHook *hook, const T *aux, const FilterInline *fil,
                    double kmin) {
  for (int a = 0; a < 3; ++a)
    if (wl[a] < 1 || wl[a] > n || w0[a] < 0 || w0[a] >= n) throw std::runtime_error("fft_r2c_window: bad window");
  if (aux && (!post || accumulate)) throw std::runtime_error("fft_r2c_window: aux needs a table and no accumulation");
  GIC_FFT_SWITCH(n, F::r2c_window(win, w0, wl, bufB, bufA, spec, post, accumulate, st, hook, aux, fil, kmin))
}
template <typename T>
void fft_c2r_window(const T *spec, int n, int pre, const T *tab, int comp, double kmin, const int *w0, const int *wl,
                    T *bufA, T *bufB, T *out, hipStream_t st, TimingHook *hook, const FilterInline *fil, int yop,
                    bool xpass) {
  for (int a = 0; a < 3; ++a)
    if (wl[a] < 1 || wl[a] > n || w0[a] < 0 || w0[a] >= n) throw std::runtime_error("fft_c2r_window: bad window");
  GIC_FFT_SWITCH(n, F::c2r_window(spec, pre, tab, comp, T(kmin), w0, wl, bufA, bufB, out, st, hook, fil, yop, xpass))
}

#define GIC_FFT_INST(T)                                                                                     \
  template void fft_r2c_post<T>(T *, int, const T *, hipStream_t, TimingHook *, const T *);                 \
  template void fft_c2r_to<T>(T *, T *, int, hipStream_t, TimingHook *);                                    \
  template void fft_r2c_batch<T>(T *, int, int, const T *, hipStream_t, TimingHook *);                      \
  template void fft_prepare<T>(int);                                                                        \
  template int fft_box_rows_multiple<T>(int);                                                               \
  template void fft_r2c_boxes<T>(const T *, int, long, int, int, const int *, const int *, T *, T *, T *, long,  \
                                 const T *, hipStream_t, TimingHook *);                                       \
  template void fft_r2c_za_adjoint<T>(const T *, T *, T *, int, const T *, int, double, bool, hipStream_t,   \
                                      TimingHook *);                                                          \
  template void fft_c2r_from<T>(const T *, T *, int, int, const T *, int, double, hipStream_t, TimingHook *, T *, \
                                double, const FilterInline *);                                               \
  template void fft_c2r_lpt<T>(const T *, T *, int, const T *, int, double, int, T *, T *, hipStream_t,       \
                               TimingHook *, double);                                                        \
  template void fft_c2r_xpass<T>(const T *, T *, int, int, const T *, int, double, hipStream_t, TimingHook *,  \
                                 double);                                                                    \
  template void fft_c2r_yz<T>(const T *, T *, int, int, double, hipStream_t, TimingHook *, T *);              \
  template void fft_c2r_yz_lpt<T>(const T *, T *, int, int, double, int, T *, T *, hipStream_t, TimingHook *); \
  template void fft_r2c_window<T>(const T *, int, const int *, const int *, T *, T *, T *, const T *, bool,      \
                                  hipStream_t, TimingHook *, const T *, const FilterInline *, double);         \
  template void fft_c2r_window<T>(const T *, int, int, const T *, int, double, const int *, const int *, T *, T *, \
                                  T *, hipStream_t, TimingHook *, const FilterInline *, int, bool);
GIC_FFT_INST(float)
GIC_FFT_INST(double)

template void fft3d<float>(float *, int, int, hipStream_t, TimingHook *);
template void fft3d<double>(double *, int, int, hipStream_t, TimingHook *);

}  // namespace gic
