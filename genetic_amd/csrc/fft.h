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
};

// In-place 3D transform of a padded field (N*N*(N+2) reals of type T).
// dir < 0: r2c (real -> FFTW layout), dir > 0: c2r (with Hermitian mirroring
// of the kz = 0, N/2 planes).  Both directions scale by N^-1.5, as
// FieldFourierManager::performTransform does.
template <typename T> void fft3d(T *d, int n, int dir, hipStream_t st, TimingHook *hook);

}  // namespace gic
