// Splicing (ICGenerator::splice, ic.hpp:1657-1686; spliceOneLevel,
// splice.hpp:37-118; conjugateGradient, cg.hpp:13-59): a new white-noise
// field a is drawn from a second seed, and each level's noise b becomes the
// field f with f = a on the flagged cells (in the overdensity basis) that
// minimises (f - b) C^-1 (f - b) elsewhere.  With P the covariance whose
// k = 0 entry is set to 1, M the flagged cells and M' the others, all in
// real space (r2c / c2r between the steps):
//   d  = P^1/2 (b - a)
//   z  = P^1/2 M' P^-1 M d
//   X v = P^1/2 M' P^-1 M' P^1/2 v,   X alpha = z by conjugate gradients
//   f  = P^-1/2 ( M' P^1/2 alpha + P^1/2 b - M d )
// The CG keeps the reference's steps: Q(x) recomputed for the residual each
// iteration, rtol 1e-6, atol 1e-12, FP64 inner products over the cells.
#include "context.h"

namespace gic {

namespace {

// the three powers of the preconditioner P = sqrtp^2 (P(0) = 1): 1/2, -1, -1/2
template <typename T>
__global__ void k_splice_tables(const T *__restrict__ sqrtp, long nent, T *__restrict__ half, T *__restrict__ minus1,
                                T *__restrict__ minus_half) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nent) return;
  T s = i == 0 ? T(1) : sqrtp[i];
  if (s == T(0)) s = T(1);   // applyTransferFunction leaves modes with a zero spectrum alone
  half[i] = s;
  minus1[i] = T(1) / (s * s);
  minus_half[i] = T(1) / s;
}

// out = (b - a) tab[|k|^2] on the stored modes
template <typename T>
__global__ void k_splice_diff(cx<T> *__restrict__ out, const cx<T> *__restrict__ b, const cx<T> *__restrict__ a,
                              const T *__restrict__ tab, int n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int h = n / 2;
  size_t nm = size_t(n) * n * (h + 1);
  if (i >= nm) return;
  int iz = int(i % size_t(h + 1));
  size_t r = i / size_t(h + 1);
  int iy = int(r % size_t(n)), ix = int(r / size_t(n));
  if (ix > h) ix -= n;
  if (iy > h) iy -= n;
  T f = tab[ix * ix + iy * iy + iz * iz];
  out[i] = cx<T>{(b[i].x - a[i].x) * f, (b[i].y - a[i].y) * f};
}

// v *= M (keep == 1) or M' (keep == 0) on the cells of a padded real field
template <typename T>
__global__ void k_splice_mask(T *__restrict__ v, const uint8_t *__restrict__ mask, int n, int keep) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  size_t row = i / size_t(n), z = i - row * size_t(n);
  if ((mask[i] != 0) != (keep != 0)) v[row * size_t(n + 2) + z] = T(0);
}

// partial sums of a.b over the cells of padded real fields
template <typename T>
__global__ __launch_bounds__(256) void k_splice_dot(const T *__restrict__ a, const T *__restrict__ b, int n,
                                                    double *__restrict__ part) {
  __shared__ double sh[256];
  size_t n3 = size_t(n) * n * n;
  double s = 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n3; i += size_t(gridDim.x) * blockDim.x) {
    size_t row = i / size_t(n), z = i - row * size_t(n);
    size_t q = row * size_t(n + 2) + z;
    s += double(a[q]) * double(b[q]);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// y = p x + q y over a padded field (padding included)
template <typename T> __global__ void k_splice_axpby(T *__restrict__ y, const T *__restrict__ x, T p, T q, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = p * x[i] + q * y[i];
}

constexpr int DOT_BLOCKS = 512;

template <typename T> struct Splicer {
  gic_context *c;
  int n;
  size_t reals;
  const uint8_t *mask;
  const T *half, *minus1, *minus_half;
  DevBuf part;

  unsigned blocks(size_t m) const { return unsigned((m + 255) / 256); }
  // v (real) <- c2r(tab r2c(v))
  void filt(T *v, const T *tab) {
    fft_r2c_post<T>(v, n, tab, c->st, nullptr);
    fft3d<T>(v, n, +1, c->st, nullptr);
  }
  void mul_mask(T *v, int keep) {
    hipLaunchKernelGGL(k_splice_mask<T>, dim3(blocks(size_t(n) * n * n)), dim3(256), 0, c->st, v, mask, n, keep);
    GIC_HIP(hipGetLastError());
  }
  // out <- X v, v unchanged
  void apply(T *out, const T *v) {
    GIC_HIP(hipMemcpyAsync(out, v, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    filt(out, half);
    mul_mask(out, 0);
    filt(out, minus1);
    mul_mask(out, 0);
    filt(out, half);
  }
  double dot(const T *a, const T *b) {
    hipLaunchKernelGGL(k_splice_dot<T>, dim3(DOT_BLOCKS), dim3(256), 0, c->st, a, b, n, (double *)part.p);
    GIC_HIP(hipGetLastError());
    std::vector<double> h(DOT_BLOCKS);
    GIC_HIP(hipMemcpyAsync(h.data(), part.p, DOT_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, c->st));
    GIC_HIP(hipStreamSynchronize(c->st));
    double s = 0;
    for (double v : h) s += v;
    return s;
  }
  void axpby(T *y, const T *x, double p, double q) {
    hipLaunchKernelGGL(k_splice_axpby<T>, dim3(blocks(reals)), dim3(256), 0, c->st, y, x, T(p), T(q), reals);
    GIC_HIP(hipGetLastError());
  }
};

}  // namespace

template <typename T> void do_splice(gic_context *c, unsigned long new_seed, int *cg_iterations) {
  if (!c->drawn) do_draw<T>(c);
  if (c->power_applied)
    throw std::runtime_error(
        "It is too late in the IC generation process to perform splicing; try moving the splice command earlier");
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
  size_t nl = c->levels.size();
  // the new field: a real-space draw from the new seed, one stream across the levels
  std::vector<DevBuf> fresh(nl);
  std::vector<RngLevelOut> outs;
  size_t total = 0;
  for (size_t l = 0; l < nl; ++l) {
    alloc_field(c, fresh[l], c->levels[l]->g);
    outs.push_back({fresh[l].p, c->levels[l]->g.n, c->levels[l]->g.n3()});
    total += c->levels[l]->g.n3();
  }
  {
    RngPlan plan;
    rng_plan(new_seed, total, plan);
    DevBuf ws;
    ws.ensure(rng_workspace_bytes(plan));
    rng_draw_real<T>(plan, outs, ws.p, ws.bytes, c->st, nullptr);
    GIC_HIP(hipStreamSynchronize(c->st));
  }
  int iters_total = 0;
  for (size_t l = 0; l < nl; ++l) {
    Level &L = *c->levels[l];
    const LevelGeom &g = L.g;
    int n = g.n;
    size_t reals = g.padded_reals();
    T *a = (T *)fresh[l].p;
    fft_r2c_post<T>(a, n, nullptr, c->st, nullptr);
    fill<T>(a, 1, T(0), c->st);   // Re(w_0) = 0, as every draw
    ensure_noise_fourier<T>(c, L);
    T *b = (T *)L.noise.p;
    // the flagged cells of this level as a byte mask
    std::vector<uint8_t> hm(g.n3(), 0);
    if (l < c->flags.size())
      for (uint64_t q : c->flags[l])
        if (q < g.n3()) hm[q] = 1;
    DevBuf dmask;
    dmask.ensure(g.n3());
    GIC_HIP(hipMemcpy(dmask.p, hm.data(), g.n3(), hipMemcpyHostToDevice));
    build_sqrtp<T>(c, L);
    long h = n / 2, nent = 3 * h * h + 1;
    DevBuf tabs;
    tabs.ensure(3 * nent * sizeof(T));
    T *half = (T *)tabs.p, *minus1 = half + nent, *minus_half = minus1 + nent;
    hipLaunchKernelGGL(k_splice_tables<T>, dim3(unsigned((nent + 255) / 256)), dim3(256), 0, c->st,
                       (const T *)L.sqrtp.p, nent, half, minus1, minus_half);
    GIC_HIP(hipGetLastError());
    Splicer<T> S{c, n, reals, (const uint8_t *)dmask.p, half, minus1, minus_half, {}};
    S.part.ensure(DOT_BLOCKS * sizeof(double));
    DevBuf bd, bz, bx, br, bdir, bq;
    for (DevBuf *p : {&bd, &bz, &bx, &br, &bdir, &bq}) alloc_field(c, *p, g);
    T *d = (T *)bd.p, *z = (T *)bz.p, *x = (T *)bx.p, *r = (T *)br.p, *dir = (T *)bdir.p, *qd = (T *)bq.p;
    // d = P^1/2 (b - a), real
    size_t nm = g.nmodes();
    hipLaunchKernelGGL(k_splice_diff<T>, dim3(S.blocks(nm)), dim3(256), 0, c->st, reinterpret_cast<cx<T> *>(d),
                       reinterpret_cast<const cx<T> *>(b), reinterpret_cast<const cx<T> *>(a), (const T *)half, n);
    GIC_HIP(hipGetLastError());
    fft3d<T>(d, n, +1, c->st, nullptr);
    // z = P^1/2 M' P^-1 M d
    GIC_HIP(hipMemcpyAsync(z, d, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    S.mul_mask(z, 1);
    S.filt(z, minus1);
    S.mul_mask(z, 0);
    S.filt(z, half);
    // conjugate gradients for X alpha = z (cg.hpp:13-59)
    fill<T>(x, reals, T(0), c->st);
    GIC_HIP(hipMemcpyAsync(r, z, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    GIC_HIP(hipMemcpyAsync(dir, z, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    S.axpby(dir, z, -1.0, 0.0);   // direction = -residual
    double scale = std::sqrt(S.dot(z, z));
    size_t it = 0;
    if (scale != 0.0) {
      size_t dim = g.n3();
      for (it = 0; it < dim + 1; ++it) {
        S.apply(qd, dir);
        double dqd = S.dot(dir, qd);
        double alpha = -S.dot(r, dir) / dqd;
        S.axpby(x, dir, alpha, 1.0);
        S.apply(r, x);
        S.axpby(r, z, -1.0, 1.0);   // residual = Q(x) - b
        double norm = std::sqrt(S.dot(r, r));
        if (norm < 1e-6 * scale || norm < 1e-12) break;
        double beta = S.dot(r, qd) / dqd;
        S.axpby(dir, r, -1.0, beta);   // direction = beta direction - residual
      }
    }
    iters_total += int(it);
    // f = P^-1/2 ( M' P^1/2 alpha + P^1/2 b - M d ); x holds alpha
    S.filt(x, half);
    S.mul_mask(x, 0);
    T *bdelta = qd;
    GIC_HIP(hipMemcpyAsync(bdelta, b, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    mul_table<T>(bdelta, n, half, c->st);
    fft3d<T>(bdelta, n, +1, c->st, nullptr);
    S.axpby(x, bdelta, 1.0, 1.0);
    S.mul_mask(d, 1);
    S.axpby(x, d, -1.0, 1.0);
    fft_r2c_post<T>(x, n, minus_half, c->st, nullptr);
    GIC_HIP(hipMemcpyAsync(b, x, reals * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    L.noise_fourier = true;
    L.dc_pending = false;
    GIC_HIP(hipStreamSynchronize(c->st));
  }
  if (cg_iterations) *cg_iterations = iters_total;
}

template void do_splice<float>(gic_context *, unsigned long, int *);
template void do_splice<double>(gic_context *, unsigned long, int *);

}  // namespace gic
