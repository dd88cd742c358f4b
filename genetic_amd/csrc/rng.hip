// Parallel, bit-exact ranlxs2 + ziggurat Gaussians on gfx950.
//
// Replaces the serial draw loop of RandomFieldGenerator::drawRandomForSpecifiedGrid
// (genetIC/src/simulation/field/randomfieldgenerator.hpp:181-197), which calls
// gsl_ran_gaussian_ziggurat(ranlxs2, 1) once per cell in index order and
// continues the same stream across levels.
//
// ranlxs2 (Luescher's RANLUX at luxury p=397, GSL's 48-bit implementation)
// runs the subtract-with-borrow recurrence x_n = x_{n-5} - x_{n-12} - c_{n-1}
// (mod b = 2^48) and emits the 12 newest words after every 397 steps as 24
// 24-bit numbers (low half first, oldest word first).  The recurrence is an
// LCG in disguise: Z_n = sum_j a_j x_{n-12+j} + b c_{n-1} (mod m, m =
// b^12 - b^5 + 1, a_j = b^{j+1} - [j>=7] b^{j-6}) obeys Z_{n+1} = b^{-1} Z_n.
// So chunk t of the stream starts from a 576-bit modular product, is decoded
// to the 12-word state (the carry ambiguity of the decode disappears after
// 12 recurrence steps), and is then generated independently.
//
// gsl_ran_gaussian_ziggurat consumes 2 numbers per accepted iteration and
// 3 or 4 on the rare edge/tail tests, so the Gaussian index of a chunk is
// data dependent.  Pass 1 generates every chunk and parses it from the 4
// possible entry offsets (the parses merge after a few iterations), pass 2
// scans the entry->(exit, count) maps, pass 3 parses each chunk again from
// its true entry and writes the Gaussians to their cells.  Pass 1 stores the
// chunk's raw 24-bit numbers (4 bytes each, about 2.05 per Gaussian) for
// pass 3 to read back, so the generator runs once; within a pass the parse
// reads them from a per-thread LDS ring.
#include "rng.h"
#include "fft.h"
#include <cstring>
#include <cmath>
#include <mutex>
#include <string>

namespace gic {

#include "zig_tables.inc"

__constant__ double c_ytab[128];
__constant__ double c_wtab[128];
__constant__ unsigned int c_ktab[128];

typedef unsigned __int128 u128;
static constexpr uint64_t MASK48 = (uint64_t(1) << 48) - 1;

// ---- 576-bit arithmetic mod m = 2^576 - 2^240 + 1 -------------------------
__host__ __device__ inline bool u576_geq_m(const uint64_t *a) {
  // m limbs: w0 = 1, w3 = 0xFFFF000000000000 (bits 240..255 clear... ) see below
  // m = 2^576 - 2^240 + 1: limbs 0..2: [1, 0, 0], limb 3: 2^64 - 2^48 (bits 48..63 set),
  // limbs 4..8: all ones.
  const uint64_t m[9] = {1, 0, 0, 0xFFFF000000000000ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
  #pragma unroll
  for (int i = 8; i >= 0; --i) {
    if (a[i] != m[i]) return a[i] > m[i];
  }
  return true;
}
__host__ __device__ inline void u576_sub_m(uint64_t *a) {
  const uint64_t m[9] = {1, 0, 0, 0xFFFF000000000000ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
  uint64_t borrow = 0;
  #pragma unroll
  for (int i = 0; i < 9; ++i) {
    u128 d = (u128)a[i] - m[i] - borrow;
    a[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
}

// s (nlimbs) <- L + H * 2^240 - H  with H given as hn limbs; result limbs
// written to out (length 14).  Requires H*2^240 >= H, true for any H >= 0.
template <int hn>
__host__ __device__ inline void fold(const uint64_t *L, const uint64_t *H, uint64_t *out) {
  uint64_t t[14];
  #pragma unroll
  for (int i = 0; i < 14; ++i) t[i] = 0;
  #pragma unroll
  for (int i = 0; i < 9; ++i) t[i] = L[i];
  // add H << 240 (= 3 limbs + 48 bits)
  u128 carry = 0;
  #pragma unroll
  for (int i = 0; i < 11; ++i) {
    uint64_t sh = 0;
    if (i < hn) sh |= H[i] << 48;
    if (i >= 1 && i - 1 < hn) sh |= H[i - 1] >> 16;
    u128 s = (u128)t[i + 3] + sh + carry;
    t[i + 3] = (uint64_t)s;
    carry = s >> 64;
  }
  // subtract H
  uint64_t borrow = 0;
  #pragma unroll
  for (int i = 0; i < 14; ++i) {
    uint64_t hv = i < hn ? H[i] : 0;
    u128 d = (u128)t[i] - hv - borrow;
    t[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
  #pragma unroll
  for (int i = 0; i < 14; ++i) out[i] = t[i];
}

__host__ __device__ inline void u576_mulmod(const uint64_t *a, const uint64_t *b, uint64_t *r) {
  uint64_t p[18];
  #pragma unroll
  for (int i = 0; i < 18; ++i) p[i] = 0;
  #pragma unroll
  for (int i = 0; i < 9; ++i) {
    uint64_t carry = 0;
    #pragma unroll
    for (int j = 0; j < 9; ++j) {
      u128 cur = (u128)a[i] * b[j] + p[i + j] + carry;
      p[i + j] = (uint64_t)cur;
      carry = (uint64_t)(cur >> 64);
    }
    p[i + 9] = carry;
  }
  uint64_t s[14];
  fold<9>(p, p + 9, s);          // < 2^576 + 2^816
  uint64_t s2[14];
  fold<5>(s, s + 9, s2);         // < 2^576 + 2^561
  #pragma unroll
  for (int i = 0; i < 9; ++i) r[i] = s2[i];
  // s2 fits in 9 limbs plus a possible carry bit in limb 9
  if (s2[9]) {
    // value >= 2^576 > m: subtract m once (2^576 - m = 2^240 - 1 < m)
    u576_sub_m(r);  // wraps limb 9 away
  }
  while (u576_geq_m(r)) u576_sub_m(r);
}

void u576_mulmod_host(const U576 &a, const U576 &b, U576 &r) { u576_mulmod(a.w, b.w, r.w); }

// 48-bit field at bit offset off of a little-endian limb array
__host__ __device__ inline uint64_t get48(const uint64_t *v, int off) {
  int li = off >> 6, bi = off & 63;
  uint64_t x = v[li] >> bi;
  if (bi > 16) x |= v[li + 1] << (64 - bi);
  return x & MASK48;
}
__host__ __device__ inline void put48(uint64_t *v, int off, uint64_t x) {
  int li = off >> 6, bi = off & 63;
  v[li] |= x << bi;
  if (bi > 16) v[li + 1] |= x >> (64 - bi);
}

// Decode V (the LCG state one step past the state S_n) into the 12 words of
// S_n with carry 0: V = A_low + (b^7 - 1) B + c, A_low the 7 oldest words,
// B the 5 newest.
__host__ __device__ inline void decode_state(const uint64_t *V, uint64_t *x) {
  // v0 = V mod 2^336, v1 = V >> 336
  uint64_t v0[7], v1[7];
  for (int i = 0; i < 7; ++i) v0[i] = v1[i] = 0;
  for (int i = 0; i < 5; ++i) v0[i] = V[i];
  v0[5] = V[5] & 0xFFFF;   // bits 320..335
  for (int i = 0; i < 4; ++i) v1[i] = (V[5 + i] >> 16) | (i + 6 < 9 ? V[6 + i] << 48 : 0);
  // v1 has at most 240 bits (4 limbs, top limb < 2^48)
  uint64_t s[7];
  u128 carry = 0;
  for (int i = 0; i < 7; ++i) {
    u128 t = (u128)v0[i] + v1[i] + carry;
    s[i] = (uint64_t)t;
    carry = t >> 64;
  }
  // compare s with D = 2^336 - 1 (limbs 0..4 all ones, limb 5 = 0xFFFF)
  bool ge = s[6] > 0 || s[5] > 0xFFFF ||
            (s[5] == 0xFFFF && s[4] == ~0ull && s[3] == ~0ull && s[2] == ~0ull && s[1] == ~0ull && s[0] == ~0ull);
  uint64_t B[7], R[7];
  for (int i = 0; i < 7; ++i) {
    B[i] = v1[i];
    R[i] = s[i];
  }
  if (ge) {
    // B += 1 ; R -= 2^336 - 1  (i.e. R = R - 2^336 + 1)
    u128 c = 1;
    for (int i = 0; i < 7; ++i) {
      u128 t = (u128)B[i] + c;
      B[i] = (uint64_t)t;
      c = t >> 64;
    }
    // R - 2^336 + 1: add 1 then clear bit 336 (R >= 2^336 - 1 so R+1 >= 2^336)
    c = 1;
    for (int i = 0; i < 7; ++i) {
      u128 t = (u128)R[i] + c;
      R[i] = (uint64_t)t;
      c = t >> 64;
    }
    // subtract 2^336 = limb 5 bit 16
    u128 d = (u128)R[5] - (uint64_t(1) << 16);
    R[5] = (uint64_t)d;
    if ((uint64_t)(d >> 64) & 1) R[6] -= 1;
  }
  // B >= 2^240 (only for V = m - 1): move one unit back into R
  if (B[3] >> 48) {
    uint64_t borrow = 1;
    for (int i = 0; i < 7; ++i) {
      u128 t = (u128)B[i] - borrow;
      B[i] = (uint64_t)t;
      borrow = (uint64_t)(t >> 64) & 1;
    }
    // R += 2^336 - 1
    u128 c = 0;
    const uint64_t D[7] = {~0ull, ~0ull, ~0ull, ~0ull, ~0ull, 0xFFFF, 0};
    for (int i = 0; i < 7; ++i) {
      u128 t = (u128)R[i] + D[i] + c;
      R[i] = (uint64_t)t;
      c = t >> 64;
    }
  }
  for (int j = 0; j < 7; ++j) x[j] = get48(R, 48 * j);
  for (int j = 0; j < 5; ++j) x[7 + j] = get48(B, 48 * j);
}

// one subtract-with-borrow step on the ring: x[i] <- x[j] - x[i] - c
#define SWB(i, j)                              \
  {                                            \
    int64_t d = (int64_t)x[j] - (int64_t)x[i] - c; \
    c = (int64_t)((uint64_t)d >> 63);          \
    x[i] = (uint64_t)d & MASK48;               \
  }

// 397 recurrence steps (33 rounds of 12 plus one), keeping x[0] as the
// oldest word afterwards by rotating the registers once.
__device__ __forceinline__ void ranlux_block(uint64_t *x, int64_t &c) {
  for (int r = 0; r < 33; ++r) {
    SWB(0, 7) SWB(1, 8) SWB(2, 9) SWB(3, 10) SWB(4, 11) SWB(5, 0)
    SWB(6, 1) SWB(7, 2) SWB(8, 3) SWB(9, 4) SWB(10, 5) SWB(11, 6)
  }
  SWB(0, 7)
  uint64_t t0 = x[0];
#pragma unroll
  for (int i = 0; i < 11; ++i) x[i] = x[i + 1];
  x[11] = t0;
}

struct RngDev {
  uint64_t seed_words[12];
  U576 y0;
  U576 jump[RngPlan::MAXBITS];
  int kc;   // ranlxs2 blocks per chunk
};

__device__ void chunk_start_state(const RngDev &p, long t, uint64_t *x, int64_t &c) {
  c = 0;
  if (t == 0) {
    for (int i = 0; i < 12; ++i) x[i] = p.seed_words[i];
    return;
  }
  uint64_t acc[9];
  for (int i = 0; i < 9; ++i) acc[i] = p.y0.w[i];
  for (int b = 0; b < RngPlan::MAXBITS; ++b) {
    if ((t >> b) & 1) {
      uint64_t r[9];
      u576_mulmod(acc, p.jump[b].w, r);
      for (int i = 0; i < 9; ++i) acc[i] = r[i];
    }
  }
  decode_state(acc, x);
  // 12 steps resolve the carry ambiguity and land on the chunk's start
  SWB(0, 7) SWB(1, 8) SWB(2, 9) SWB(3, 10) SWB(4, 11) SWB(5, 0)
  SWB(6, 1) SWB(7, 2) SWB(8, 3) SWB(9, 4) SWB(10, 5) SWB(11, 6)
}

// ---- ziggurat iteration ---------------------------------------------------------
// The ziggurat tables in LDS: the lookups are indexed by random bytes, so
// each lane reads a different entry (constant memory serves one address per
// wavefront)
struct ZigTabs {
  double y[128];
  double w[128];
  uint32_t k[128];
};
// all threads of the block call this before any of them returns
__device__ __forceinline__ void zig_tabs_load(ZigTabs &z) {
  for (int i = int(threadIdx.x); i < 128; i += int(blockDim.x)) {
    z.y[i] = c_ytab[i];
    z.w[i] = c_wtab[i];
    z.k[i] = c_ktab[i];
  }
  __syncthreads();
}

// One loop iteration of gsl_ran_gaussian_ziggurat starting at stream position
// pos: returns the number of numbers consumed (2, 3 or 4) and whether it
// accepted; *val receives the Gaussian when accepted.
template <class Get>
__device__ __forceinline__ int zig_iter(Get get, long pos, bool &accept, double &val, const ZigTabs &zt) {
  const double PARAM_R = 3.44428647676;
  uint32_t k1 = get(pos), k2 = get(pos + 1);
  uint32_t i = k1 & 0xFF, j = k2 & 0xFFFFFF;
  double sign = (i & 0x80) ? 1.0 : -1.0;
  i &= 0x7f;
  double x = j * zt.w[i];
  if (j < zt.k[i]) {
    accept = true;
    val = sign * x;
    return 2;
  }
  double y;
  int used;
  if (i < 127) {
    double y0 = zt.y[i], y1 = zt.y[i + 1];
    double U1 = get(pos + 2) / 16777216.0;
    y = y1 + (y0 - y1) * U1;
    used = 3;
  } else {
    double U1 = 1.0 - get(pos + 2) / 16777216.0;
    double U2 = get(pos + 3) / 16777216.0;
    x = PARAM_R - log(U1) / PARAM_R;
    y = exp(-PARAM_R * (x - 0.5 * PARAM_R)) * U2;
    used = 4;
  }
  accept = y < exp(-0.5 * x * x);
  val = sign * x;
  return used;
}

// Chunks are parsed straight out of the generator: each thread keeps the
// newest 48 numbers of its chunk in an LDS ring (slot-major, so a wave's
// threads touch consecutive banks) and advances its ziggurat cursors while
// an iteration is guaranteed to fit in the numbers generated so far.
constexpr int RING = 48;

template <int SLOTS> struct RingT {
  uint32_t *base;
  int stride;
  // positions within a chunk are non-negative and below 2^31
  __device__ __forceinline__ uint32_t get(long pos) const { return base[int(uint32_t(pos) % SLOTS) * stride]; }
  __device__ __forceinline__ void put(long pos, uint32_t v) const { base[int(uint32_t(pos) % SLOTS) * stride] = v; }
};
using Ring = RingT<RING>;

template <class R> __device__ __forceinline__ void push_block(const R &r, long at, const uint64_t *x) {
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    r.put(at + 2 * k, uint32_t(x[k] & 0xFFFFFF));
    r.put(at + 2 * k + 1, uint32_t(x[k] >> 24));
  }
}

// Raw numbers of the summary pass, kept for the emit pass: block q, number k
// of chunk t at ((t/64)*rawb + q*24 + k)*64 + t%64 with rawb = (kc+1)*24, so a
// wavefront's 64 chunks store and load each number as one contiguous 256-byte
// segment.
__device__ __forceinline__ size_t raw_at(long t, int q, int k, int rawb) {
  return (size_t(t >> 6) * rawb + q * 24 + k) * 64 + (t & 63);
}

// Pass 1: per chunk, the exit offset and Gaussian count for each of the four
// possible entry offsets.  Once the four cursors sit on the same position
// they share the rest of the parse, so only one is advanced and its later
// acceptances are added to all four.
__global__ __launch_bounds__(256) void k_rng_summary(const RngDev *__restrict__ pp, long nchunks,
                                                    uint32_t *__restrict__ summary,
                                                    uint32_t *__restrict__ raw) {
  extern __shared__ uint32_t ring_mem[];
  __shared__ ZigTabs zt;
  zig_tabs_load(zt);
  long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nchunks) return;
  Ring ring{ring_mem + threadIdx.x, int(blockDim.x)};
  const RngDev &p = *pp;
  const int kc = p.kc, ch = kc * 24, rawb = (kc + 1) * 24;
  uint64_t x[12];
  int64_t c;
  chunk_start_state(p, t, x, c);
  auto get = [&](long pos) { return ring.get(pos); };
  long cur[4] = {0, 1, 2, 3};
  uint32_t cnt[4] = {0, 0, 0, 0};
  uint32_t shared_cnt = 0;   // acceptances after the cursors merged
  bool merged = false;
  long avail = 0;
  for (int q = 0; q <= kc; ++q) {
    ranlux_block(x, c);
    push_block(ring, avail, x);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      raw[raw_at(t, q, 2 * k, rawb)] = uint32_t(x[k] & 0xFFFFFF);
      raw[raw_at(t, q, 2 * k + 1, rawb)] = uint32_t(x[k] >> 24);
    }
    avail += 24;
    if (merged) {
      long pos = cur[0];
      while (pos < ch && pos + 4 <= avail) {
        bool acc;
        double v;
        pos += zig_iter(get, pos, acc, v, zt);
        shared_cnt += acc;
      }
      cur[0] = pos;
      if (pos >= ch) break;
    } else {
      bool done = true;
      for (int e = 0; e < 4; ++e) {
        long pos = cur[e];
        while (pos < ch && pos + 4 <= avail) {
          bool acc;
          double v;
          pos += zig_iter(get, pos, acc, v, zt);
          cnt[e] += acc;
        }
        cur[e] = pos;
        done &= pos >= ch;
      }
      merged = cur[0] == cur[1] && cur[0] == cur[2] && cur[0] == cur[3];
      if (done) break;
    }
  }
  uint32_t ex = 0;
  for (int e = 0; e < 4; ++e) {
    long end = merged ? cur[0] : cur[e];
    ex |= uint32_t(end - ch) << (2 * e);
    summary[t * 5 + 1 + e] = cnt[e] + (merged ? shared_cnt : 0);
  }
  summary[t * 5] = ex;
}

// Pass 2a: compose the maps of SEG consecutive chunks.
struct Map4 {
  uint8_t ex[4];
  uint32_t cnt[4];
};
__device__ __forceinline__ Map4 compose(const Map4 &a, const Map4 &b) {
  Map4 r;
  for (int e = 0; e < 4; ++e) {
    r.ex[e] = b.ex[a.ex[e]];
    r.cnt[e] = a.cnt[e] + b.cnt[a.ex[e]];
  }
  return r;
}
__device__ __forceinline__ Map4 load_map(const uint32_t *s, long t) {
  Map4 m;
  uint32_t ex = s[t * 5];
  for (int e = 0; e < 4; ++e) {
    m.ex[e] = (ex >> (2 * e)) & 3;
    m.cnt[e] = s[t * 5 + 1 + e];
  }
  return m;
}
__device__ __forceinline__ Map4 identity_map() {
  Map4 m;
  for (int e = 0; e < 4; ++e) {
    m.ex[e] = e;
    m.cnt[e] = 0;
  }
  return m;
}
constexpr int SEG = 64;

__global__ __launch_bounds__(256) void k_rng_seg_compose(const uint32_t *__restrict__ summary, long nchunks, long nseg,
                                                         uint32_t *__restrict__ segmaps) {
  long s = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  Map4 acc = identity_map();
  long b0 = s * SEG, b1 = min(nchunks, b0 + SEG);
  for (long t = b0; t < b1; ++t) acc = compose(acc, load_map(summary, t));
  uint32_t ex = 0;
  for (int e = 0; e < 4; ++e) {
    ex |= uint32_t(acc.ex[e]) << (2 * e);
    segmaps[s * 5 + 1 + e] = acc.cnt[e];
  }
  segmaps[s * 5] = ex;
}

// Pass 2b: one workgroup scans the segment maps; segstart[s] = (base << 2) | entry.
__global__ __launch_bounds__(1024) void k_rng_seg_scan(const uint32_t *__restrict__ segmaps, long nseg,
                                                      uint64_t *__restrict__ segstart) {
  __shared__ Map4 sm[1024];
  int tid = threadIdx.x;
  long per = (nseg + 1023) / 1024;
  long b0 = tid * per, b1 = min(nseg, b0 + per);
  Map4 acc = identity_map();
  for (long t = b0; t < b1; ++t) acc = compose(acc, load_map(segmaps, t));
  sm[tid] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    Map4 v = sm[tid];
    if (tid >= off) v = compose(sm[tid - off], v);
    __syncthreads();
    sm[tid] = v;
    __syncthreads();
  }
  int e = 0;
  uint64_t base = 0;
  if (tid > 0) {
    e = sm[tid - 1].ex[0];
    base = sm[tid - 1].cnt[0];
  }
  for (long t = b0; t < b1; ++t) {
    segstart[t] = (base << 2) | uint64_t(e);
    Map4 m = load_map(segmaps, t);
    base += m.cnt[e];
    e = m.ex[e];
  }
}

// Pass 3: regenerate each chunk, parse from its true entry, write Gaussians.
struct LevelTargets {
  void *f[4];
  int n[4];        // grid size of a padded real field; 0 for a dense array
  size_t cum[4];   // cumulative Gaussian counts at the end of each target
  double scale;    // every value is stored as T(v * scale)
};

// Pass 3 works one wavefront per workgroup.  Each lane parses its chunk from
// the stored raw numbers (read through an LDS ring, one coalesced block at a
// time) and stages its Gaussians in LDS; lanes holding 32 or more flush
// them cooperatively, two lanes' runs of 32 consecutive cells per store
// instruction, so every store writes whole 128-byte lines.
constexpr int EMIT_RING = 32;
// staged Gaussians per lane: a flush leaves fewer than EMIT_FLUSH and one
// block of 24 numbers (plus up to 3 left over) adds at most 13
constexpr int EMIT_FLUSH = 32;   // flushing 16 at a time ran slower at 1024^3 (more barriers than occupancy gained)
constexpr int EMIT_STAGE = 48;
static_assert(EMIT_FLUSH - 1 + 13 <= EMIT_STAGE, "staging overflow");

template <typename T>
__device__ __forceinline__ T *cell_ptr(const LevelTargets &lt, uint64_t g) {
  int L = g < lt.cum[0] ? 0 : g < lt.cum[1] ? 1 : g < lt.cum[2] ? 2 : 3;
  size_t i = g - (L ? lt.cum[L - 1] : 0);
  int n = lt.n[L];
  if (n == 0) return reinterpret_cast<T *>(lt.f[L]) + i;
  size_t nn = size_t(n) * n;
  size_t xx = i / nn, r = i - xx * nn, y = r / n, z = r - y * n;
  return reinterpret_cast<T *>(lt.f[L]) + (xx * n + y) * (n + 2) + z;
}

// Level and (x, y, z) of global Gaussian index g; n[L] == 0 (a dense target)
// leaves the coordinates unused
__device__ __forceinline__ void cell_coords(const LevelTargets &lt, uint64_t g, int &L, uint32_t &x, uint32_t &y,
                                            uint32_t &z) {
  L = g < lt.cum[0] ? 0 : g < lt.cum[1] ? 1 : g < lt.cum[2] ? 2 : 3;
  size_t i = g - (L ? lt.cum[L - 1] : 0);
  int n = lt.n[L];
  x = y = z = 0;
  if (n == 0) return;
  size_t nn = size_t(n) * n;
  size_t xx = i / nn, r = i - xx * nn, yy = r / n;
  x = uint32_t(xx);
  y = uint32_t(yy);
  z = uint32_t(r - yy * n);
}

template <typename T>
__global__ __launch_bounds__(64) void k_rng_emit(const uint32_t *__restrict__ raw, long nchunks,
                                                const uint32_t *__restrict__ summary,
                                                const uint64_t *__restrict__ segstart, LevelTargets lt,
                                                size_t total, int kc) {
  __shared__ uint32_t ring_mem[EMIT_RING * 64];
  __shared__ T stage_mem[EMIT_STAGE * 64];
  __shared__ ZigTabs zt;
  zig_tabs_load(zt);
  const int lane = threadIdx.x;
  long t = long(blockIdx.x) * 64 + lane;
  bool active = t < nchunks;
  uint64_t g = 0;
  int e = 0;
  if (active) {
    long s = t / SEG;
    uint64_t ss = segstart[s];
    g = ss >> 2;
    e = int(ss & 3);
    for (long u = s * SEG; u < t; ++u) {
      // the map's entry e only (no per-thread array indexed by e)
      uint32_t exw = summary[u * 5];
      g += summary[u * 5 + 1 + e];
      e = int((exw >> (2 * e)) & 3u);
    }
    if (g >= total) active = false;
  }
  RingT<EMIT_RING> ring{ring_mem + lane, 64};
  auto get = [&](long pos) { return ring.get(pos); };
  long pos = e;
  long avail = 0;
  int fill = 0;
  uint64_t gstage = g;   // global index of staged slot 0
  auto flush = [&](bool final) {
    __syncthreads();
    int want = final ? fill : (fill >= EMIT_FLUSH ? EMIT_FLUSH : 0);
    unsigned long long mask = __ballot(want > 0);
    // each lane locates its own run's first cell once; the 32 lanes storing
    // a run step from there (no divisions per stored value)
    int myL = 0;
    uint32_t mx = 0, my = 0, mz = 0;
    if (want > 0) cell_coords(lt, gstage, myL, mx, my, mz);
    while (mask) {
      int s0 = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      int s1 = -1;
      if (mask) {
        s1 = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
      }
      int src = lane < 32 ? s0 : s1;
      int j = lane & 31;
      int sl = src < 0 ? 0 : src;
      int w = __shfl(want, sl);
      uint64_t gs = __shfl(gstage, sl);
      int L = __shfl(myL, sl);
      uint32_t x = __shfl(mx, sl), y = __shfl(my, sl), z = __shfl(mz, sl);
      if (src >= 0 && j < w) {
        T *ptr;
        int n = lt.n[L];
        if (n != 0 && gs + j < lt.cum[L]) {
          uint32_t zz = z + j, yy = y, xx = x;
          while (zz >= uint32_t(n)) {   // at most once when n >= 32
            zz -= n;
            if (++yy == uint32_t(n)) {
              yy = 0;
              ++xx;
            }
          }
          ptr = reinterpret_cast<T *>(lt.f[L]) + (size_t(xx) * n + yy) * (n + 2) + zz;
        } else {
          ptr = cell_ptr<T>(lt, gs + j);   // a dense target, or a run that crosses into the next level
        }
        *ptr = stage_mem[j * 64 + src];
      }
    }
    __syncthreads();
    if (want > 0) {
      for (int j = 0; j + want < fill; ++j) stage_mem[j * 64 + lane] = stage_mem[(j + want) * 64 + lane];
      fill -= want;
      gstage += want;
    }
    __syncthreads();
  };
  const int ch = kc * 24, rawb = (kc + 1) * 24;
  // the next block's numbers are loaded while this block is parsed (every
  // block q <= kc of a chunk lies inside the raw buffer)
  uint32_t nxt[24];
  if (active) {
#pragma unroll
    for (int k = 0; k < 24; ++k) nxt[k] = raw[raw_at(t, 0, k, rawb)];
  }
  for (int q = 0; q <= kc; ++q) {
    bool busy = active && pos < ch && g < total;
    if (!__any(busy)) break;
    if (busy) {
#pragma unroll
      for (int k = 0; k < 24; ++k) ring.put(avail + k, nxt[k]);
      avail += 24;
      if (q < kc) {
#pragma unroll
        for (int k = 0; k < 24; ++k) nxt[k] = raw[raw_at(t, q + 1, k, rawb)];
      }
      while (pos < ch && pos + 4 <= avail && g < total) {
        bool acc;
        double v;
        pos += zig_iter(get, pos, acc, v, zt);
        if (acc) {
          stage_mem[fill * 64 + lane] = T(v * lt.scale);
          ++fill;
          ++g;
        }
      }
    }
    flush(false);
  }
  flush(true);
}

// Raw 24-bit numbers in stream order (for bit-exactness checks).
__global__ __launch_bounds__(256) void k_rng_raw(RngDev p, long nchunks, uint32_t *__restrict__ out, size_t n) {
  long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nchunks) return;
  uint64_t x[12];
  int64_t c;
  chunk_start_state(p, t, x, c);
  size_t base = size_t(t) * (p.kc * 24);
  for (int q = 0; q < p.kc; ++q) {
    ranlux_block(x, c);
    for (int k = 0; k < 12; ++k) {
      size_t i = base + q * 24 + 2 * k;
      if (i < n) out[i] = uint32_t(x[k] & 0xFFFFFF);
      if (i + 1 < n) out[i + 1] = uint32_t(x[k] >> 24);
    }
  }
}

// ---- host set-up ------------------------------------------------------------------
static void seed_words(unsigned long s, uint64_t *xw) {
  // GSL ranlxs_set_lux / Luescher rlxs_init: bits of the seed drive a
  // 31-bit shift register clocked 48 times per word.
  if (s == 0) s = 1;
  long seed = (long)s;
  int i = (int)(seed & 0xFFFFFFFFUL);
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
    xw[k] = (uint64_t)xv & MASK48;
  }
}

static void u576_from_words(const uint64_t *x, int nw, uint64_t *out) {
  for (int i = 0; i < 9; ++i) out[i] = 0;
  for (int j = 0; j < nw; ++j) put48(out, 48 * j, x[j]);
}

static void u576_powmod(const uint64_t *base, u128 e_lo, uint64_t *r) {
  uint64_t acc[9] = {1, 0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t b[9];
  for (int i = 0; i < 9; ++i) b[i] = base[i];
  while (e_lo) {
    if (e_lo & 1) {
      uint64_t t[9];
      u576_mulmod(acc, b, t);
      memcpy(acc, t, sizeof acc);
    }
    uint64_t t[9];
    u576_mulmod(b, b, t);
    memcpy(b, t, sizeof b);
    e_lo >>= 1;
  }
  memcpy(r, acc, sizeof acc);
}

void rng_plan(unsigned long seed, size_t n_gaussians, RngPlan &plan) {
  seed_words(seed, plan.seed_words);
  // A = sum_j x_j b^j (12 words), B = sum_j x_{7+j} b^j (5 words)
  uint64_t A[9], B[9];
  u576_from_words(plan.seed_words, 12, A);
  u576_from_words(plan.seed_words + 7, 5, B);
  // reduce A mod m (A < 2^576 but may exceed m)
  while (u576_geq_m(A)) u576_sub_m(A);
  // V_0 = A - B (mod m) is the LCG state one step past the seed (carry 0)
  uint64_t V[9];
  {
    uint64_t borrow = 0;
    for (int i = 0; i < 9; ++i) {
      u128 d = (u128)A[i] - B[i] - borrow;
      V[i] = (uint64_t)d;
      borrow = (uint64_t)(d >> 64) & 1;
    }
    if (borrow) {
      // add m back
      const uint64_t m[9] = {1, 0, 0, 0xFFFF000000000000ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
      u128 c = 0;
      for (int i = 0; i < 9; ++i) {
        u128 t = (u128)V[i] + m[i] + c;
        V[i] = (uint64_t)t;
        c = t >> 64;
      }
    }
  }
  // b^{-1} = m - b^11 + b^4 ; b^12 = b^5 - 1 (mod m)
  uint64_t binv[9];
  {
    const uint64_t m[9] = {1, 0, 0, 0xFFFF000000000000ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
    uint64_t b11[9] = {0}, b4[9] = {0};
    put48(b11, 48 * 11, 1);
    put48(b4, 48 * 4, 1);
    uint64_t borrow = 0;
    for (int i = 0; i < 9; ++i) {
      u128 d = (u128)m[i] - b11[i] - borrow;
      binv[i] = (uint64_t)d;
      borrow = (uint64_t)(d >> 64) & 1;
    }
    u128 c = 0;
    for (int i = 0; i < 9; ++i) {
      u128 t = (u128)binv[i] + b4[i] + c;
      binv[i] = (uint64_t)t;
      c = t >> 64;
    }
  }
  // raw numbers needed: 2 per Gaussian plus rejections (~1.2% of iterations
  // take 3-4 numbers and may repeat); 2.2x plus slack is never exceeded in
  // practice and is checked after the scan.  Chunks of kc blocks: long
  // chunks amortise the jump to each chunk's start, and a small draw takes
  // shorter ones (down to four blocks) so that it still has about 64k
  // chunks (a wavefront per SIMD) to run on; shorter chunks than that spend
  // more on the jumps than they gain in parallelism.
  double need = 2.2 * double(n_gaussians) + 4096.0;
  plan.kc = 64;
  while (plan.kc > 4 && need / (plan.kc * 24.0) < 65536.0) plan.kc /= 2;
  // The chunk-t start state is step T_t = 397*kc*t; the device decodes the
  // LCG state V_{T-12} = Z_{T-11} = b^{-(T-12)} V_0 and runs 12 steps.
  // y0 = b^{12} V_0 ;  jump[i] = b^{-397*kc*2^i}.
  uint64_t b12[9];
  {
    uint64_t one_b[9] = {0};
    put48(one_b, 48, 1);        // b
    u576_powmod(one_b, 12, b12);
  }
  u576_mulmod(b12, V, plan.y0.w);
  uint64_t step[9];
  u576_powmod(binv, (u128)397 * plan.kc, step);
  for (int i = 0; i < RngPlan::MAXBITS; ++i) {
    memcpy(plan.jump[i].w, step, sizeof step);
    uint64_t t[9];
    u576_mulmod(step, step, t);
    memcpy(step, t, sizeof t);
  }
  plan.nchunks = long(need / plan.ch()) + 2;
}

static size_t raw_bytes(long nc, int kc) { return size_t((nc + 63) / 64) * 64 * size_t(kc + 1) * 24 * 4; }

size_t rng_workspace_bytes(const RngPlan &plan) {
  long nc = plan.nchunks;
  long nseg = (nc + SEG - 1) / SEG;
  return (size_t(nc) * 5 * 4 + 255) / 256 * 256 + (size_t(nseg) * 5 * 4 + 255) / 256 * 256 +
         (size_t(nseg) * 8 + 255) / 256 * 256 + raw_bytes(nc, plan.kc) + (sizeof(RngDev) + 255) / 256 * 256 + 256;
}

struct RngBufs {
  RngDev *dev;   // the plan, read by the summary pass
  uint32_t *summary, *segmaps, *raw;
  uint64_t *segstart;
  long nseg;
};
static RngBufs carve(const RngPlan &plan, void *ws, size_t bytes) {
  long nc = plan.nchunks;
  long nseg = (nc + SEG - 1) / SEG;
  char *p = (char *)ws;
  auto take = [&](size_t b) {
    char *r = p;
    p += (b + 255) & ~size_t(255);
    return r;
  };
  RngBufs r;
  r.summary = (uint32_t *)take(size_t(nc) * 5 * 4);
  r.segmaps = (uint32_t *)take(size_t(nseg) * 5 * 4);
  r.segstart = (uint64_t *)take(size_t(nseg) * 8);
  r.raw = (uint32_t *)take(raw_bytes(nc, plan.kc));
  r.dev = (RngDev *)take(sizeof(RngDev));
  r.nseg = nseg;
  if ((size_t)(p - (char *)ws) > bytes) throw std::runtime_error("RNG workspace too small");
  return r;
}

constexpr int RNG_MAX_DEVICES = 64;
static std::mutex g_tables_mu;
static bool g_tables_done[RNG_MAX_DEVICES];

// the ziggurat tables into the constant memory of the current device (each
// device has its own copy), once per device
static void upload_tables() {
  int dev = 0;
  GIC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= RNG_MAX_DEVICES) throw std::runtime_error("RNG: device index out of range");
  std::lock_guard<std::mutex> g(g_tables_mu);
  if (g_tables_done[dev]) return;
  GIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_ytab), ZIG_YTAB, sizeof ZIG_YTAB));
  GIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_wtab), ZIG_WTAB, sizeof ZIG_WTAB));
  GIC_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_ktab), ZIG_KTAB, sizeof ZIG_KTAB));
  g_tables_done[dev] = true;
}

void rng_prepare() { upload_tables(); }

static RngDev to_dev(const RngPlan &plan) {
  RngDev d;
  memcpy(d.seed_words, plan.seed_words, sizeof d.seed_words);
  d.y0 = plan.y0;
  memcpy(d.jump, plan.jump, sizeof d.jump);
  d.kc = plan.kc;
  return d;
}

// the device-side form of the coverage check below
__global__ void k_rng_cover(const uint64_t *__restrict__ last_start, const uint32_t *__restrict__ last_map,
                            uint64_t total, uint32_t *__restrict__ flag) {
  if (threadIdx.x != 0) return;
  uint64_t last = *last_start;
  uint64_t produced = (last >> 2) + last_map[1 + (last & 3)];
  if (produced < total) *flag = 1u;
}

// the plan's device copy, written by a kernel from its arguments
__global__ void k_rng_put_plan(RngDev d, RngDev *__restrict__ dst) {
  const uint64_t *s = reinterpret_cast<const uint64_t *>(&d);
  uint64_t *o = reinterpret_cast<uint64_t *>(dst);
  for (int i = threadIdx.x; i < int(sizeof(RngDev) / 8); i += blockDim.x) o[i] = s[i];
}
static_assert(sizeof(RngDev) % 8 == 0, "RngDev is copied as 64-bit words");

size_t rng_dev_bytes() { return sizeof(RngDev); }
void rng_dev_pack(const RngPlan &plan, void *host_dst) {
  RngDev d = to_dev(plan);
  memcpy(host_dst, &d, sizeof d);
}

template <typename T>
static void draw_real_list(const RngPlan &plan, const RngLevelOut *levels, size_t nlev, void *ws, size_t ws_bytes,
                           hipStream_t st, TimingHook *hook, double scale, uint32_t *cover_flag, const void *dev_plan) {
  if (nlev < 1 || nlev > 4) throw std::runtime_error("real-space draw supports 1 to 4 levels");
  upload_tables();
  RngBufs b = carve(plan, ws, ws_bytes);
  long nc = plan.nchunks;
  const int TB = 256;
  size_t ring_bytes = size_t(RING) * TB * 4;
  const RngDev *pd = static_cast<const RngDev *>(dev_plan);
  if (!pd) {
    hipLaunchKernelGGL(k_rng_put_plan, dim3(1), dim3(64), 0, st, to_dev(plan), b.dev);
    GIC_HIP(hipGetLastError());
    pd = b.dev;
  }
  if (hook) hook->begin("rng_summary", st);
  hipLaunchKernelGGL(k_rng_summary, dim3((nc + TB - 1) / TB), dim3(TB), ring_bytes, st, pd, nc, b.summary, b.raw);
  GIC_HIP(hipGetLastError());
  if (hook) hook->end("rng_summary", st);
  if (hook) hook->begin("rng_scan", st);
  hipLaunchKernelGGL(k_rng_seg_compose, dim3((b.nseg + 255) / 256), dim3(256), 0, st, b.summary, nc, b.nseg,
                     b.segmaps);
  GIC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rng_seg_scan, dim3(1), dim3(1024), 0, st, b.segmaps, b.nseg, b.segstart);
  GIC_HIP(hipGetLastError());
  if (hook) hook->end("rng_scan", st);
  LevelTargets lt;
  size_t acc = 0;
  for (size_t i = 0; i < 4; ++i) {
    lt.f[i] = i < nlev ? levels[i].field : nullptr;
    lt.n[i] = i < nlev ? levels[i].n : 0;
    if (i < nlev) acc += levels[i].count;
    lt.cum[i] = acc;
  }
  lt.scale = scale;
  size_t total = acc;
  // coverage: the stream must hold at least `total` Gaussians
  if (cover_flag) {
    hipLaunchKernelGGL(k_rng_cover, dim3(1), dim3(64), 0, st, (const uint64_t *)(b.segstart + (b.nseg - 1)),
                       (const uint32_t *)(b.segmaps + 5 * (b.nseg - 1)), uint64_t(total), cover_flag);
    GIC_HIP(hipGetLastError());
  } else {
    uint64_t last = 0;
    uint32_t lastmap[5];
    GIC_HIP(hipMemcpyAsync(&last, b.segstart + (b.nseg - 1), 8, hipMemcpyDeviceToHost, st));
    GIC_HIP(hipMemcpyAsync(lastmap, b.segmaps + 5 * (b.nseg - 1), 20, hipMemcpyDeviceToHost, st));
    GIC_HIP(hipStreamSynchronize(st));
    uint64_t produced = (last >> 2) + lastmap[1 + (last & 3)];
    if (produced < total) throw std::runtime_error("internal error: RNG plan produced too few Gaussians");
  }
  if (hook) hook->begin("rng_emit", st);
  hipLaunchKernelGGL(k_rng_emit<T>, dim3((nc + 63) / 64), dim3(64), 0, st, b.raw, nc, b.summary, b.segstart, lt,
                     total, plan.kc);
  GIC_HIP(hipGetLastError());
  if (hook) hook->end("rng_emit", st);
}

template <typename T>
void rng_draw_real(const RngPlan &plan, const std::vector<RngLevelOut> &levels, void *ws, size_t ws_bytes,
                   hipStream_t st, TimingHook *hook, double scale, uint32_t *cover_flag, const void *dev_plan) {
  draw_real_list<T>(plan, levels.data(), levels.size(), ws, ws_bytes, st, hook, scale, cover_flag, dev_plan);
}

template void rng_draw_real<float>(const RngPlan &, const std::vector<RngLevelOut> &, void *, size_t, hipStream_t,
                                   TimingHook *, double, uint32_t *, const void *);
template void rng_draw_real<double>(const RngPlan &, const std::vector<RngLevelOut> &, void *, size_t, hipStream_t,
                                    TimingHook *, double, uint32_t *, const void *);

// ---- draws with a fixed workspace ------------------------------------------------
// For callers that cannot allocate device memory: the draw's workspace (chunk
// summaries, segment maps and starts, raw numbers, the plan) lives in this
// file-scope device array, one copy per GPU, held until the process ends.
// Its address on each device is looked up once (rng_scratch_setup, called by
// gic_create with the device current) and only that address is passed to
// kernels: the array's name in host code is a host-side stand-in.  The
// workspace needs no clearing (every pass writes what the next one reads).
// Calls that overlap on one GPU share the array and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(13) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[RNG_MAX_DEVICES];

void rng_scratch_setup(int device) {
  if (device < 0 || device >= RNG_MAX_DEVICES) throw std::runtime_error("RNG: device index out of range");
  upload_tables();
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  GIC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw std::runtime_error("RNG: no device address for the fixed workspace");
  g_scratch_ptr[device] = p;
}

size_t rng_noalloc_max_gaussians() { return size_t(1) << 20; }

void rng_gaussians_noalloc(int device, unsigned long seed, double *dev_out, size_t n, hipStream_t st) {
  if (n > rng_noalloc_max_gaussians())
    throw std::runtime_error("gic_gaussian_stream_noalloc: at most " + std::to_string(rng_noalloc_max_gaussians()) +
                             " Gaussians per call (the fixed workspace holds " +
                             std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                             " MiB); gic_gaussian_stream draws longer streams");
  void *ws = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device >= 0 && device < RNG_MAX_DEVICES) ws = g_scratch_ptr[device];
  }
  if (!ws) throw std::runtime_error("gic_gaussian_stream_noalloc: the fixed workspace of this device is not set up");
  RngPlan plan;
  rng_plan(seed, n, plan);
  if (rng_workspace_bytes(plan) > NOALLOC_SCRATCH_BYTES)
    throw std::runtime_error("internal error: the fixed RNG workspace is too small for this draw");
  RngLevelOut out{dev_out, 0, n};
  draw_real_list<double>(plan, &out, 1, ws, NOALLOC_SCRATCH_BYTES, st, nullptr, 1.0, nullptr, nullptr);
}

void rng_raw(const RngPlan &plan, uint32_t *d_out, size_t n, void *, size_t, hipStream_t st) {
  if (n > size_t(plan.nchunks) * plan.ch()) throw std::runtime_error("rng_raw: plan too small");
  RngDev d = to_dev(plan);
  long nc = long((n + plan.ch() - 1) / plan.ch());
  hipLaunchKernelGGL(k_rng_raw, dim3((nc + 255) / 256), dim3(256), 0, st, d, nc, d_out, n);
  GIC_HIP(hipGetLastError());
}

}  // namespace gic
