// Batches of independent constrained realisations of one single-level grid.
//
// The reference makes one realisation per run: the draw
// (randomfieldgenerator.hpp:122-139, 181-197), linear modifications
// (linearmodification.hpp, modificationmanager.hpp:160-312), the transfer
// function (field.hpp:394-412) and the Zel'dovich displacement
// (zeldovich.hpp:152-253).  Here many seeds, each with its own set of top-hat
// overdensity constraints, run as one sequence of device work with no host
// synchronisation inside, which is recorded into one HIP graph per batch:
//   w      = draw(seed)                          (the stream of gic_draw)
//   c_i    = sqrtP_dm r2c(1_{S_i})               (the realisation's constraints in one batched r2c)
//   G_ij   = <c_i, c_j>,  d_i = <c_i, w>         (FP64 sums, one pass over all vectors)
//   x      = (G_ij / N_i N_j)^-1 (t_i - d_i / N_i)  (FP64 Cholesky on the device, N_i = |S_i|)
//   delta  = sqrtP (w + sum_i x_i c_i / N_i)
//   out    = c2r(delta), c2r(i k_a delta / k^2)   (dense n^3 arrays)
// S_i is ICGenerator::select with a sphere (ic.hpp:1458-1500): the cells whose
// centroids lie closer than the radius (as a float) to the centre, with
// box-wrapped offsets.  The covector of the mean overdensity in S_i is
// 1_{S_i} / N_i (linearmodification.hpp, OverdensityModification), carried to
// white-noise space by sqrtP_dm; the Gram solve is the Hoffman-Ribak
// correction of modificationmanager.hpp:160-172,266-283.
#include "context.h"

namespace gic {

namespace {

constexpr int MAXCONS = 32;        // constraints per realisation
constexpr int GRAM_THREADS = 256;
constexpr int GRAM_BLOCKS = 512;
constexpr int GRAM_VMAX = 36;      // MAXCONS + 1 vectors rounded up to tiles of 4
constexpr int GRAM_NWMAX = 45;     // tile pairs of 9 tiles

__device__ __forceinline__ double stored_mode_weight(int x, int y, int z, int h) {
  // each conjugate pair once: the kz = 0 and kz = N/2 planes hold both halves
  if (z > 0 && z < h) return 2;
  if (x > 0 && x < h) return 2;
  if (x > h) return 0;
  if (y > 0 && y < h) return 2;
  if (y > h) return 0;
  return 1;
}

// 1 on the cells of each constraint's sphere, 0 elsewhere (padding included);
// cnt[f] += the number of cells.  The whole-field form, for spheres too large
// for a box.
template <typename T>
__global__ __launch_bounds__(256) void k_batch_tophat(T *__restrict__ cov, long fs, int n, double dx, double ox,
                                                      double oy, double oz, double periodic,
                                                      const double *__restrict__ cons, uint32_t *__restrict__ cnt) {
#pragma clang fp contract(off)
  int f = blockIdx.y;
  // 32-bit index arithmetic: a field holds n^2 (n+2) < 2^31 values for n <= 1024
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned tot = unsigned(n) * unsigned(n) * unsigned(n + 2);
  bool inside = false;
  bool live = i < tot;
  if (live) {
    unsigned row = i / unsigned(n + 2);
    int z = int(i - row * unsigned(n + 2));
    int x = int(row / unsigned(n)), y = int(row - unsigned(x) * unsigned(n));
    if (z < n) {
      const double *q = cons + 5 * f;
      double half = periodic / 2;
      auto off = [&](int idx, double o, double c) {
        double r = double(idx) * dx;
        r = r + o;
        r = r + dx / 2;
        r = r - c;
        if (r > half) r -= periodic;
        if (r < -half) r += periodic;
        return r;
      };
      double rx = off(x, ox, q[0]), ry = off(y, oy, q[1]), rz = off(z, oz, q[2]);
      float rf = float(q[3]);
      float rr = rf * rf;
      double d2 = rx * rx + ry * ry;
      d2 = d2 + rz * rz;
      inside = d2 < double(rr);
    }
  }
  unsigned long long m = __ballot(inside);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(cnt + f, uint32_t(__popcll(m)));
  if (live) cov[size_t(f) * fs + i] = inside ? T(1) : T(0);
}

// The rows of each constraint's box (wl x wl cells in x, y from (winx[f],
// winy[f]) mod n, all z): 1 on the cells of the sphere, 0 elsewhere; row
// (f wl + i) wl + j is x = winx[f]+i, y = winy[f]+j.  Rows past nm wl^2 (up
// to nrows) are zero.  cnt[f] += the number of cells of constraint f.  The
// box holds every selected cell: a cell is selected only if its centroid
// lies within the radius of the centre, and the box spans the centre's
// +-radius with a cell to spare.
template <typename T>
__global__ __launch_bounds__(256) void k_batch_box_rows(T *__restrict__ rows, unsigned nrows, int nm, int wl, int n,
                                                        double dx, double ox, double oy, double oz, double periodic,
                                                        const double *__restrict__ cons, const int *__restrict__ winx,
                                                        const int *__restrict__ winy, uint32_t *__restrict__ cnt) {
#pragma clang fp contract(off)
  unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = e < nrows * unsigned(n);
  unsigned row = live ? e / unsigned(n) : 0u;
  int z = live ? int(e - row * unsigned(n)) : 0;
  unsigned ww = unsigned(wl) * unsigned(wl);
  int f = int(row / ww);
  bool inside = false;
  if (live && f < nm) {
    unsigned r2 = row - unsigned(f) * ww;
    int i = int(r2 / unsigned(wl)), j = int(r2 - unsigned(i) * unsigned(wl));
    int x = (winx[f] + i) % n, y = (winy[f] + j) % n;
    const double *q = cons + 5 * f;
    double half = periodic / 2;
    auto off = [&](int idx, double o, double c) {
      double r = double(idx) * dx;
      r = r + o;
      r = r + dx / 2;
      r = r - c;
      if (r > half) r -= periodic;
      if (r < -half) r += periodic;
      return r;
    };
    double rx = off(x, ox, q[0]), ry = off(y, oy, q[1]), rz = off(z, oz, q[2]);
    float rf = float(q[3]);
    float rr = rf * rf;
    double d2 = rx * rx + ry * ry;
    d2 = d2 + rz * rz;
    inside = d2 < double(rr);
  }
  // one atomic per wavefront when its lanes share a constraint
  int f0 = __shfl(f, 0);
  unsigned long long m = __ballot(inside);
  if (__all(!live || f == f0)) {
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(cnt + f0, uint32_t(__popcll(m)));
  } else if (inside) {
    atomicAdd(cnt + f, 1u);
  }
  if (live) rows[e] = inside ? T(1) : T(0);
}

// Partial sums of the products of the vectors v_0..v_nm (v_i = c_i for i <
// nm, v_nm = w), weighted by the stored-mode weights: chunks of K modes of
// every vector are staged through LDS, and each thread forms a 4 x 4 tile of
// products (vectors 4ti+u against 4tj+v, tj <= ti) over every G-th mode of
// the chunk, so that each LDS value feeds four products.  The G mode-groups
// of a tile are summed in order of the group at the end; block b writes the
// 16 values of tile pair p = ti(ti+1)/2 + tj at part[(b nw + p) 16 + 4u + v].
template <typename T>
__global__ __launch_bounds__(GRAM_THREADS) void k_batch_gram(const cx<T> *__restrict__ cov, long fs,
                                                             const cx<T> *__restrict__ w, int nm, int n,
                                                             double *__restrict__ part) {
  constexpr int K = sizeof(T) == 4 ? 128 : 64;
  constexpr int S = K + 1;   // LDS row stride: rows of different vectors start in different banks
  __shared__ cx<T> sv[GRAM_VMAX * S];
  __shared__ double sw[K];
  __shared__ double red[GRAM_NWMAX * 16];
  const int nv = nm + 1;
  const int nt = (nv + 3) / 4;
  const int nw = nt * (nt + 1) / 2;
  const int G = GRAM_THREADS / nw;
  const int h = n / 2;
  const size_t nmodes = size_t(n) * n * (h + 1);
  const size_t nchunks = (nmodes + K - 1) / K;
  const int t = threadIdx.x;
  const int p = t % nw, g = t / nw;
  const bool live = g < G;
  int ti = 0;
  while ((ti + 1) * (ti + 2) / 2 <= p) ++ti;
  const int tj = p - ti * (ti + 1) / 2;
  double acc[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0;
  for (size_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    size_t m0 = ch * K;
    for (int e = t; e < nt * 4 * K; e += GRAM_THREADS) {
      int v = e / K, q = e - v * K;
      size_t m = m0 + q;
      cx<T> val{T(0), T(0)};
      if (m < nmodes && v < nv) val = v < nm ? cov[size_t(v) * fs + m] : w[m];
      sv[v * S + q] = val;
    }
    if (t < K) {
      size_t m = m0 + t;
      double wt = 0;
      if (m < nmodes) {
        int z = int(m % size_t(h + 1));
        size_t r = m / size_t(h + 1);
        wt = stored_mode_weight(int(r / size_t(n)), int(r % size_t(n)), z, h);
      }
      sw[t] = wt;
    }
    __syncthreads();
    if (live) {
      const cx<T> *A = sv + 4 * ti * S;
      const cx<T> *B = sv + 4 * tj * S;
      for (int q = g; q < K; q += G) {
        double wt = sw[q];
        double ax[4], ay[4], bx[4], by[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          cx<T> va = A[u * S + q], vb = B[u * S + q];
          ax[u] = wt * double(va.x);
          ay[u] = wt * double(va.y);
          bx[u] = double(vb.x);
          by[u] = double(vb.y);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u * 4 + v] += ax[u] * bx[v] + ay[u] * by[v];
      }
    }
    __syncthreads();
  }
  for (int e = t; e < nw * 16; e += GRAM_THREADS) red[e] = 0;
  __syncthreads();
  for (int gg = 0; gg < G; ++gg) {
    if (live && g == gg) {
#pragma unroll
      for (int e = 0; e < 16; ++e) red[p * 16 + e] += acc[e];
    }
    __syncthreads();
  }
  for (int e = t; e < nw * 16; e += GRAM_THREADS) part[size_t(blockIdx.x) * nw * 16 + e] = red[e];
}

// where the product of v_i and v_j (i >= j) sits among the tile sums
__device__ __forceinline__ int gram_slot(int i, int j) {
  int ti = i >> 2, tj = j >> 2;
  return ((ti * (ti + 1) / 2 + tj) << 4) + ((i & 3) << 2) + (j & 3);
}

// sums[p] = sum over blocks of part[b * np + p], one block per pair, fixed order
__global__ __launch_bounds__(256) void k_batch_reduce(const double *__restrict__ part, int nblk, int np,
                                                      double *__restrict__ sums) {
  __shared__ double sh[256];
  int p = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += 256) s += part[size_t(b) * np + p];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[p] = sh[0];
}

// The FP64 solve of one realisation on one wavefront, lane i holding row i
// (nm <= MAXCONS): G = L L^T by right-looking Cholesky with the column
// broadcast by lane shuffles, then L y = t - d / N and L^T x = y;
// coef_i = x_i / N_i.  *flag gets 2 if a constraint selects no cell, 4 if
// the constraints are not linearly independent (the RNG coverage check uses
// 1); either leaves the coefficients at 0.
__global__ __launch_bounds__(64) void k_batch_solve(const double *__restrict__ sums, const uint32_t *__restrict__ cnt,
                                                    const double *__restrict__ cons, int nm,
                                                    double *__restrict__ coef, uint32_t *__restrict__ flag) {
  __shared__ double Ls[MAXCONS * MAXCONS];
  const int i = threadIdx.x;
  const bool row = i < nm;
  const uint32_t ni = row ? cnt[i] : 1u;
  if (row) coef[i] = 0;
  if (__ballot(ni == 0u)) {
    if (i == 0) *flag |= 2u;
    return;
  }
  const double inv = 1.0 / double(ni);
  double a[MAXCONS];
#pragma unroll
  for (int j = 0; j < MAXCONS; ++j) {
    double invj = __shfl(inv, j);
    a[j] = (row && j <= i) ? sums[gram_slot(i, j)] * inv * invj : 0.0;
  }
  // constant trip counts (k < nm as a predicate) keep a[] in registers
#pragma unroll
  for (int k = 0; k < MAXCONS; ++k) {
    const bool act = k < nm;
    double dkk = __shfl(a[k], k);
    if (act && !(dkk > 0)) {
      if (i == 0) *flag |= 4u;
      if (row) coef[i] = 0;
      return;
    }
    double lkk = act ? sqrt(dkk) : 1.0;
    if (act && i == k) a[k] = lkk;
    if (act && i > k) a[k] /= lkk;
#pragma unroll
    for (int j = k + 1; j < MAXCONS; ++j) {
      double ljk = __shfl(a[k], j);
      if (act && i >= j) a[j] -= a[k] * ljk;
    }
  }
  // forward substitution, column by column
  double y = row ? cons[5 * i + 4] - sums[gram_slot(nm, i)] * inv : 0.0;
#pragma unroll
  for (int k = 0; k < MAXCONS; ++k) {
    const bool act = k < nm;
    double yk = __shfl(y / a[k], k);
    if (act && i == k) y = yk;
    if (act && i > k) y -= a[k] * yk;
  }
  if (row) {
#pragma unroll
    for (int j = 0; j < MAXCONS; ++j)
      if (j <= i) Ls[i * MAXCONS + j] = a[j];
  }
  __syncthreads();
  // back substitution with L^T: x_k final once rows above it are folded in
  for (int k = nm - 1; k >= 0; --k) {
    double xk = __shfl(y / Ls[k * MAXCONS + k], k);
    if (i == k) y = xk;
    if (i < k) y -= Ls[k * MAXCONS + i] * xk;
  }
  if (row) coef[i] = y * inv;
}

// w <- sqrtP (w + sum_i coef_i c_i), mode by mode
template <typename T>
__global__ __launch_bounds__(256) void k_batch_update(cx<T> *__restrict__ w, const cx<T> *__restrict__ cov, long fs,
                                                      int nm, const double *__restrict__ coef,
                                                      const T *__restrict__ sqrtp, int n) {
  __shared__ T s[MAXCONS];
  if (threadIdx.x < nm) s[threadIdx.x] = T(coef[threadIdx.x]);
  __syncthreads();
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t nmodes = size_t(n) * n * (n / 2 + 1);
  if (i >= nmodes) return;
  cx<T> v = w[i];
  for (int j = 0; j < nm; ++j) {
    cx<T> c = cov[size_t(j) * fs + i];
    v.x += s[j] * c.x;
    v.y += s[j] * c.y;
  }
  int h = n / 2;
  int iz = int(i % size_t(h + 1));
  size_t r = i / size_t(h + 1);
  int iy = int(r % size_t(n)), ix = int(r / size_t(n));
  if (ix > h) ix -= n;
  if (iy > h) iy -= n;
  T f = sqrtp[ix * ix + iy * iy + iz * iz];
  w[i] = cx<T>{v.x * f, v.y * f};
}

}  // namespace

// Buffers of one chain of realisations; chains run concurrently on their own
// streams, each realisation of a chain after the previous one.
struct BatchLane {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  DevBuf w, cov, scratch, ws, part, sums, cnt, coef;
  DevBuf rows, zbuf, ybuf;   // covectors from the constraints' boxes
};

struct BatchState {
  std::vector<std::unique_ptr<BatchLane>> lanes;   // lanes never move once made
  hipEvent_t fork = nullptr;
  DevBuf cons, flags, win, plans;
  hipGraphExec_t exec = nullptr;   // the last recorded batch, replayed while its shape and buffers stay
  std::vector<uint64_t> key;
  ~BatchState() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (auto &l : lanes) {
      if (l->done) (void)hipEventDestroy(l->done);
      if (l->st) (void)hipStreamDestroy(l->st);
    }
    if (fork) (void)hipEventDestroy(fork);
  }
};

template <typename T>
void do_batch(gic_context *c, int nreal, const uint64_t *seeds, int ncons, const double *cons, T *out, int nlanes,
              bool graph, double *ms) {
  if (c->levels.size() != 1) throw std::runtime_error("batched realisations need exactly one grid level");
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
  if (nreal < 1 || !seeds || !out) throw std::runtime_error("batch: no realisations, seeds or output");
  if (ncons < 1 || ncons > MAXCONS || !cons)
    throw std::runtime_error("batch: between 1 and " + std::to_string(MAXCONS) + " constraints per realisation");
  if (nlanes < 1 || nlanes > 16) throw std::runtime_error("batch: 1 to 16 concurrent chains");
  nlanes = std::min(nlanes, nreal);
  for (int i = 0; i < nreal * ncons; ++i)
    if (!(cons[5 * i + 3] > 0) || !std::isfinite(cons[5 * i + 4]))
      throw std::runtime_error("batch: each constraint needs a positive radius and a finite target");
  Level &L = *c->levels[0];
  const LevelGeom &g = L.g;
  const int n = g.n;
  const long fs_reals = long(g.padded_reals());
  const long fs = long(g.nmodes());   // complex elements per padded field
  const size_t n3 = g.n3();
  const int ntile = (ncons + 1 + 3) / 4;
  const int np = ntile * (ntile + 1) / 2 * 16;   // tile sums of the Gram kernel

  // everything a recorded graph needs exists before the first launch
  fft_prepare<T>(n);
  rng_prepare();
  build_sqrtp<T>(c, L);
  build_sqrtp_dm<T>(c, L);
  if (!c->batch) c->batch = std::make_shared<BatchState>();
  BatchState &B = *c->batch;
  RngPlan plan0;
  rng_plan(seeds[0], n3, plan0);
  size_t wsb = rng_workspace_bytes(plan0);
  while (int(B.lanes.size()) < nlanes) {
    B.lanes.push_back(std::make_unique<BatchLane>());
    BatchLane &ln = *B.lanes.back();
    GIC_HIP(hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking));
    GIC_HIP(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
  }
  if (!B.fork) GIC_HIP(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
  // boxes around the spheres: x, y windows of wl cells from
  // floor((c - o - R)/dx - 1/2) (mod n), which hold every selected cell
  int wl = 0;
  for (int q = 0; q < nreal * ncons; ++q) {
    double span = 2.0 * std::ceil(cons[5 * q + 3] / g.dx) + 3.0;
    wl = std::max(wl, span >= double(n) ? n : int(span));
  }
  // (the row kernel indexes rows x n values in 32 bits)
  const bool boxes = wl < n && (double(ncons) * wl * wl + 1024.0) * n < 2e9;
  const int rows_mult = fft_box_rows_multiple<T>(n);
  const long nrows = boxes ? (long(ncons) * wl * wl + rows_mult - 1) / rows_mult * rows_mult : 0;
  std::vector<int> win(boxes ? size_t(nreal) * ncons * 2 : 0);
  if (boxes) {
    const double org[2] = {g.ox, g.oy};
    for (int q = 0; q < nreal * ncons; ++q)
      for (int a = 0; a < 2; ++a) {
        long w0 = long(std::floor((cons[5 * q + a] - org[a] - cons[5 * q + 3]) / g.dx - 0.5));
        w0 %= n;
        if (w0 < 0) w0 += n;
        win[size_t(q) * 2 + a] = int(w0);
      }
  }
  for (int k = 0; k < nlanes; ++k) {
    BatchLane &ln = *B.lanes[k];
    if (boxes) {
      ln.rows.ensure(size_t(nrows) * n * sizeof(T));
      ln.zbuf.ensure(size_t(nrows) * (n + 2) * sizeof(T));
      ln.ybuf.ensure(size_t(ncons) * wl * n * (n + 2) * sizeof(T));
    }
    ln.w.ensure(fs_reals * sizeof(T));
    ln.cov.ensure(size_t(ncons) * fs_reals * sizeof(T));
    ln.scratch.ensure(fs_reals * sizeof(T));
    ln.ws.ensure(wsb);
    ln.part.ensure(size_t(GRAM_BLOCKS) * np * sizeof(double));
    ln.sums.ensure(size_t(np) * sizeof(double));
    ln.cnt.ensure(size_t(ncons) * sizeof(uint32_t));
    ln.coef.ensure(size_t(ncons) * sizeof(double));
  }
  B.cons.ensure(size_t(nreal) * ncons * 5 * sizeof(double));
  B.flags.ensure(size_t(nreal) * sizeof(uint32_t));
  GIC_HIP(hipMemcpyAsync(B.cons.p, cons, size_t(nreal) * ncons * 5 * sizeof(double), hipMemcpyHostToDevice, c->st));
  if (boxes) {
    // per realisation: the x origins of its boxes, then the y origins
    std::vector<int> wxy(size_t(nreal) * ncons * 2);
    for (int r = 0; r < nreal; ++r)
      for (int f = 0; f < ncons; ++f) {
        wxy[(size_t(r) * 2 + 0) * ncons + f] = win[(size_t(r) * ncons + f) * 2 + 0];
        wxy[(size_t(r) * 2 + 1) * ncons + f] = win[(size_t(r) * ncons + f) * 2 + 1];
      }
    B.win.ensure(wxy.size() * sizeof(int));
    GIC_HIP(hipMemcpy(B.win.p, wxy.data(), wxy.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  fill<uint32_t>((uint32_t *)B.flags.p, size_t(nreal), 0u, c->st);
  // every realisation's RNG plan in device memory: the recorded work holds
  // no seed, so one instantiated graph serves every batch of this shape
  std::vector<RngPlan> plans(nreal);
  const size_t pb = (rng_dev_bytes() + 255) / 256 * 256;
  std::vector<char> hplans(size_t(nreal) * pb, 0);
  for (int r = 0; r < nreal; ++r) {
    rng_plan(seeds[r], n3, plans[r]);
    rng_dev_pack(plans[r], hplans.data() + size_t(r) * pb);
  }
  B.plans.ensure(hplans.size());
  GIC_HIP(hipMemcpy(B.plans.p, hplans.data(), hplans.size(), hipMemcpyHostToDevice));
  GIC_HIP(hipStreamSynchronize(c->st));
  std::vector<uint64_t> key;
  {
    auto u = [&](double v) {
      uint64_t b;
      memcpy(&b, &v, 8);
      key.push_back(b);
    };
    auto q = [&](const void *ptr) { key.push_back(uint64_t(uintptr_t(ptr))); };
    for (long v : {long(sizeof(T)), long(n), long(nreal), long(ncons), long(nlanes), long(boxes), long(wl), nrows,
                   long(np), plans[0].nchunks, long(plans[0].kc)})
      key.push_back(uint64_t(v));
    for (double v : {g.dx, g.ox, g.oy, g.oz, g.periodic, g.kmin()}) u(v);
    for (const void *ptr : std::initializer_list<const void *>{out, B.cons.p, B.flags.p, B.win.p, B.plans.p, L.sqrtp.p,
                                                               L.sqrtp_dm.p})
      q(ptr);
    for (int k = 0; k < nlanes; ++k) {
      BatchLane &ln = *B.lanes[k];
      for (const void *ptr : std::initializer_list<const void *>{ln.st, ln.w.p, ln.cov.p, ln.scratch.p, ln.ws.p,
                                                                 ln.part.p, ln.sums.p, ln.cnt.p, ln.coef.p, ln.rows.p,
                                                                 ln.zbuf.p, ln.ybuf.p})
        q(ptr);
      key.push_back(uint64_t(ln.ws.bytes));
    }
  }

  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (ms) {
    GIC_HIP(hipEventCreate(&t0));
    GIC_HIP(hipEventCreate(&t1));
    GIC_HIP(hipEventRecord(t0, c->st));
  }
  auto record = [&]() {
    GIC_HIP(hipEventRecord(B.fork, c->st));
    for (int k = 0; k < nlanes; ++k) GIC_HIP(hipStreamWaitEvent(B.lanes[k]->st, B.fork, 0));
    const double *dcons = (const double *)B.cons.p;
    uint32_t *dflags = (uint32_t *)B.flags.p;
    for (int r = 0; r < nreal; ++r) {
      BatchLane &ln = *B.lanes[r % nlanes];
      hipStream_t st = ln.st;
      T *w = (T *)ln.w.p;
      T *cv = (T *)ln.cov.p;
      const double *rc = dcons + size_t(r) * ncons * 5;
      // the white noise, then its spectrum with Re(w_0) = 0 (randomfieldgenerator.hpp:194-195)
      std::vector<RngLevelOut> outs{{ln.w.p, n, n3}};
      rng_draw_real<T>(plans[r], outs, ln.ws.p, ln.ws.bytes, st, nullptr, 1.0, dflags + r,
                       (const char *)B.plans.p + size_t(r) * pb);
      fft_r2c_post<T>(w, n, nullptr, st, nullptr);
      fill<T>(w, 1, T(0), st);
      // the constraints' covectors in white-noise space
      fill<uint32_t>((uint32_t *)ln.cnt.p, size_t(ncons), 0u, st);
      if (boxes) {
        const int *wx = (const int *)B.win.p + size_t(r) * 2 * ncons, *wy = wx + ncons;
        unsigned ne = unsigned(nrows) * unsigned(n);
        hipLaunchKernelGGL(k_batch_box_rows<T>, dim3((ne + 255) / 256), dim3(256), 0, st, (T *)ln.rows.p,
                           unsigned(nrows), ncons, wl, n, g.dx, g.ox, g.oy, g.oz, g.periodic, rc, wx, wy,
                           (uint32_t *)ln.cnt.p);
        GIC_HIP(hipGetLastError());
        fft_r2c_boxes<T>((const T *)ln.rows.p, n, nrows, ncons, wl, wx, wy, (T *)ln.zbuf.p, (T *)ln.ybuf.p, cv, fs,
                         (const T *)L.sqrtp_dm.p, st, nullptr);
      } else {
        hipLaunchKernelGGL(k_batch_tophat<T>, dim3(unsigned((fs_reals + 255) / 256), ncons), dim3(256), 0, st, cv,
                           fs_reals, n, g.dx, g.ox, g.oy, g.oz, g.periodic, rc, (uint32_t *)ln.cnt.p);
        GIC_HIP(hipGetLastError());
        fft_r2c_batch<T>(cv, n, ncons, (const T *)L.sqrtp_dm.p, st, nullptr);
      }
      // Gram matrix and current values, FP64 solve, update with the transfer function
      hipLaunchKernelGGL(k_batch_gram<T>, dim3(GRAM_BLOCKS), dim3(GRAM_THREADS), 0, st,
                         reinterpret_cast<const cx<T> *>(cv), fs, reinterpret_cast<const cx<T> *>(w), ncons, n,
                         (double *)ln.part.p);
      GIC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_batch_reduce, dim3(np), dim3(256), 0, st, (const double *)ln.part.p, GRAM_BLOCKS, np,
                         (double *)ln.sums.p);
      GIC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_batch_solve, dim3(1), dim3(64), 0, st, (const double *)ln.sums.p,
                         (const uint32_t *)ln.cnt.p, rc, ncons, (double *)ln.coef.p, dflags + r);
      GIC_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_batch_update<T>, dim3(unsigned((fs + 255) / 256)), dim3(256), 0, st,
                         reinterpret_cast<cx<T> *>(w), reinterpret_cast<const cx<T> *>(cv), fs, ncons,
                         (const double *)ln.coef.p, (const T *)L.sqrtp.p, n);
      GIC_HIP(hipGetLastError());
      // delta and the Zel'dovich displacements, dense, into the caller's slots
      T *o = out + size_t(r) * 4 * n3;
      fft_c2r_from<T>(w, (T *)ln.scratch.p, n, FFT_PRE_NONE, nullptr, 0, g.kmin(), st, nullptr, o);
      for (int a = 0; a < 3; ++a)
        fft_c2r_from<T>(w, (T *)ln.scratch.p, n, FFT_PRE_ZA, nullptr, a, g.kmin(), st, nullptr, o + (a + 1) * n3,
                        1.0);
    }
    for (int k = 0; k < nlanes; ++k) {
      GIC_HIP(hipEventRecord(B.lanes[k]->done, B.lanes[k]->st));
      GIC_HIP(hipStreamWaitEvent(c->st, B.lanes[k]->done, 0));
    }
  };
  if (!graph) {
    record();
  } else {
    if (B.exec && B.key != key) {
      GIC_HIP(hipGraphExecDestroy(B.exec));
      B.exec = nullptr;
    }
    if (!B.exec) {
      hipGraph_t gr = nullptr;
      GIC_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
      try {
        record();
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(c->st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
      }
      GIC_HIP(hipStreamEndCapture(c->st, &gr));
      hipError_t e = hipGraphInstantiate(&B.exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        B.exec = nullptr;
        GIC_HIP(e);
      }
      B.key = key;
    }
    GIC_HIP(hipGraphLaunch(B.exec, c->st));
  }
  if (ms) GIC_HIP(hipEventRecord(t1, c->st));
  GIC_HIP(hipStreamSynchronize(c->st));
  if (ms) {
    float e = 0;
    GIC_HIP(hipEventElapsedTime(&e, t0, t1));
    *ms = e;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
  }
  std::vector<uint32_t> fl(nreal);
  GIC_HIP(hipMemcpy(fl.data(), B.flags.p, size_t(nreal) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int r = 0; r < nreal; ++r) {
    if (fl[r] & 1u) throw std::runtime_error("internal error: RNG plan produced too few Gaussians");
    if (fl[r] & 2u)
      throw std::runtime_error("batch realisation " + std::to_string(r) + ": a constraint selects no cells");
    if (fl[r] & 4u)
      throw std::runtime_error("batch realisation " + std::to_string(r) +
                               ": Modifications are not linearly independent");
  }
}

template void do_batch<float>(gic_context *, int, const uint64_t *, int, const double *, float *, int, bool, double *);
template void do_batch<double>(gic_context *, int, const uint64_t *, int, const double *, double *, int, bool,
                               double *);

}  // namespace gic
