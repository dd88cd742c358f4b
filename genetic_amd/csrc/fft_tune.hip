// Timing harness of the FFT tile tuner (genetic_amd/tune_fft.py): linked
// with fft.hip built for one tile shape (GIC_TUNE_THREADS, GIC_TUNE_COLS)
// into a library of its own.  gic_tune_fft checks that the forward + inverse
// 3D transform of a padded n^3 field gives the field back, then times the
// plain transforms and the operator and windowed forms the path uses.
#include "gic_common.h"
#include "fft.h"
#include <cstring>

namespace {

// a hash of the index in [-1, 1) on the n^3 cells, 0 in the padding
template <typename T> __global__ void k_tune_fill(T *__restrict__ d, int n, size_t reals) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= reals) return;
  size_t z = i % size_t(n + 2);
  uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  d[i] = z < size_t(n) ? T(double(h >> 11) * (2.0 / 9007199254740992.0) - 1.0) : T(0);
}

// max |a - b| over the cells (not the padding), as the bits of a
// non-negative float, by atomicMax
template <typename T>
__global__ void k_tune_maxdiff(const T *__restrict__ a, const T *__restrict__ b, int n, size_t reals,
                               unsigned *__restrict__ out) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= reals) return;
  if (i % size_t(n + 2) >= size_t(n)) return;
  float d = float(fabs(double(a[i]) - double(b[i])));
  atomicMax(out, __float_as_uint(d));
}

template <typename T> __global__ void k_tune_ones(T *__restrict__ d, long n) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] = T(1);
}

// ms[0]: forward + inverse transform; ms[1]: the Zel'dovich inverse
// transform (i k_x / k^2, dense output); ms[2]: a windowed inverse transform
// with a low-pass filter; ms[3]: a windowed forward transform that adds a
// filtered spectrum (the metric's self term).  The window is n/16 + 8 cells
// a side, as the zoom windows of a base grid are.
template <typename T> void run(int n, int reps, double *ms, double *err) {
  size_t reals = size_t(n) * n * (n + 2);
  int wlen = n / 16 + 8 < n ? n / 16 + 8 : n / 2;
  int w0[3] = {n / 8, n / 8, n / 8}, wl[3] = {wlen, wlen, wlen};
  size_t nz = size_t(n / 2 + 1);
  size_t bufa = size_t(wl[0]) * n * nz * 2, bufb = size_t(wl[0]) * (wl[1] > n ? wl[1] : n) * nz * 2;
  size_t box = size_t(wl[0]) * wl[1] * wl[2];
  long h = n / 2, nent = 3 * h * h + 1;
  hipStream_t st;
  GIC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  T *a = nullptr, *b = nullptr, *c = nullptr, *A = nullptr, *B = nullptr, *o = nullptr, *tab = nullptr;
  unsigned *md = nullptr;
  GIC_HIP(hipMalloc(&a, reals * sizeof(T)));
  GIC_HIP(hipMalloc(&b, reals * sizeof(T)));
  GIC_HIP(hipMalloc(&c, reals * sizeof(T)));
  GIC_HIP(hipMalloc(&A, bufa * sizeof(T)));
  GIC_HIP(hipMalloc(&B, bufb * sizeof(T)));
  GIC_HIP(hipMalloc(&o, box * sizeof(T)));
  GIC_HIP(hipMalloc(&tab, nent * sizeof(T)));
  GIC_HIP(hipMalloc(&md, sizeof(unsigned)));
  unsigned blocks = unsigned((reals + 255) / 256);
  hipLaunchKernelGGL(k_tune_fill<T>, dim3(blocks), dim3(256), 0, st, a, n, reals);
  GIC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_tune_ones<T>, dim3(unsigned((nent + 255) / 256)), dim3(256), 0, st, tab, nent);
  GIC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_tune_fill<T>, dim3(unsigned((box + 255) / 256)), dim3(256), 0, st, o, 1, box);
  GIC_HIP(hipGetLastError());
  GIC_HIP(hipMemcpyAsync(b, a, reals * sizeof(T), hipMemcpyDeviceToDevice, st));
  gic::fft_prepare<T>(n);
  // one round trip checked against the field
  gic::fft3d<T>(a, n, -1, st, nullptr);
  gic::fft3d<T>(a, n, +1, st, nullptr);
  GIC_HIP(hipMemsetAsync(md, 0, sizeof(unsigned), st));
  hipLaunchKernelGGL(k_tune_maxdiff<T>, dim3(blocks), dim3(256), 0, st, (const T *)a, (const T *)b, n, reals, md);
  GIC_HIP(hipGetLastError());
  unsigned bits = 0;
  GIC_HIP(hipMemcpyAsync(&bits, md, sizeof bits, hipMemcpyDeviceToHost, st));
  GIC_HIP(hipStreamSynchronize(st));
  float e;
  memcpy(&e, &bits, sizeof e);
  *err = e;
  gic::fft3d<T>(a, n, -1, st, nullptr);   // a holds a spectrum from here on
  double kmin = 2 * M_PI / 100.0;
  gic::FilterInline fil;
  fil.nf = 1;
  fil.kind[0] = 0;
  fil.kc[0] = kmin * n / 8;
  fil.scale = 1.0;
  auto timed = [&](auto &&work) {
    hipEvent_t t0, t1;
    GIC_HIP(hipEventCreate(&t0));
    GIC_HIP(hipEventCreate(&t1));
    work();   // warm-up
    GIC_HIP(hipEventRecord(t0, st));
    for (int r = 0; r < reps; ++r) work();
    GIC_HIP(hipEventRecord(t1, st));
    GIC_HIP(hipEventSynchronize(t1));
    float m = 0;
    GIC_HIP(hipEventElapsedTime(&m, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return double(m) / reps;
  };
  ms[0] = timed([&] {
    gic::fft3d<T>(c, n, -1, st, nullptr);
    gic::fft3d<T>(c, n, +1, st, nullptr);
  });
  ms[1] = timed([&] {
    gic::fft_c2r_from<T>(a, c, n, gic::FFT_PRE_ZA, nullptr, 0, kmin, st, nullptr, b, 1.0, nullptr);
  });
  ms[2] = timed([&] {
    gic::fft_c2r_window<T>(a, n, gic::FFT_PRE_TABLE, tab, 0, kmin, w0, wl, A, B, o, st, nullptr, &fil);
  });
  ms[3] = timed([&] {
    gic::fft_r2c_window<T>(o, n, w0, wl, B, A, c, tab, false, st, nullptr, a, &fil, kmin);
  });
  for (T *p : {a, b, c, A, B, o, tab}) (void)hipFree(p);
  (void)hipFree(md);
  (void)hipStreamDestroy(st);
}

}  // namespace

// ms: four timings (milliseconds per call), as run() lists them
extern "C" int gic_tune_fft(int n, int prec, int reps, double *ms, double *max_err) {
  try {
    if (reps < 1 || !ms || !max_err || n < 32) return 2;
    if (prec == 32) run<float>(n, reps, ms, max_err);
    else run<double>(n, reps, ms, max_err);
    return 0;
  } catch (...) {
    return 1;
  }
}
