// Shared definitions for the genetIC field-generation kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace gic {

template <typename T> struct Cx;
template <> struct Cx<float> { using type = float2; };
template <> struct Cx<double> { using type = double2; };
template <typename T> using cx = typename Cx<T>::type;

template <typename C> __device__ __host__ inline C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __host__ inline C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __host__ inline C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __host__ inline C cconj(C a) { return C{a.x, -a.y}; }

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define GIC_HIP(x) ::gic::hip_check((x), #x)

// Grid geometry of one level.  Real-space fields are stored padded: element
// (x,y,z) at (x*N + y)*(N+2) + z; Fourier-space fields reuse the same memory
// as [N][N][N/2+1] interleaved complex (FFTW's in-place r2c layout).
struct LevelGeom {
  int n = 0;
  double box = 0;        // thisGridSize (Mpc/h)
  double periodic = 0;   // periodicDomainSize
  double dx = 0;
  double ox = 0, oy = 0, oz = 0;
  size_t padded_reals() const { return size_t(n) * n * (n + 2); }
  size_t n3() const { return size_t(n) * n * n; }
  size_t nmodes() const { return size_t(n) * n * (n / 2 + 1); }
  double kmin() const { return 2. * M_PI / box; }
  bool covers_full() const { return periodic == box; }
};

}  // namespace gic
