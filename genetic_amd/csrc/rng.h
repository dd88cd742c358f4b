// Bit-exact parallel ranlxs2 stream and ziggurat Gaussians.
#pragma once
#include "gic_common.h"
#include <vector>

namespace gic {

// 576-bit numbers as 9 little-endian 64-bit limbs.
struct U576 {
  uint64_t w[9];
};

// Host-side set-up of one real-space draw: the seeded RANLUX state and the
// table of jump multipliers b^{-397*KC*2^i} mod m (b = 2^48, m = b^12-b^5+1).
struct RngPlan {
  static constexpr int MAXBITS = 32;
  int kc = 64;                             // ranlxs2 blocks (24 numbers) per chunk: 4..64, set by rng_plan
  int ch() const { return kc * 24; }       // raw numbers per chunk
  uint64_t seed_words[12];                 // the 12 seeded 48-bit words, oldest first
  U576 y0;                                 // LCG state 11 steps past the seeded state, times ... (see rng.hip)
  U576 jump[MAXBITS];
  long nchunks = 0;
};

void rng_plan(unsigned long seed, size_t n_gaussians, RngPlan &plan);

// Draw the stream's Gaussians into the listed targets in order, as the
// reference's serial loop does: a padded real field (n > 0, count n^3, cell
// order x, y, z) or a dense array (n == 0, count values).  Each value is
// stored as T(v * scale).
struct RngLevelOut {
  void *field;       // T*
  int n;
  size_t count;
};
template <typename T>
void rng_draw_real(const RngPlan &plan, const std::vector<RngLevelOut> &levels, void *workspace, size_t ws_bytes,
                   hipStream_t st, struct TimingHook *hook, double scale = 1.0, uint32_t *cover_flag = nullptr,
                   const void *dev_plan = nullptr);
// With cover_flag the draw has no host synchronisation: a device check sets
// *cover_flag to 1 if the stream held too few Gaussians (the cells past the
// end are then left as they were), instead of the host raising an error.
// dev_plan: the plan already in device memory (rng_dev_pack'd, rng_dev_bytes
// long, copied there by the caller); otherwise a kernel writes it into the
// workspace, so that the draw's launches carry no seed-dependent arguments
// past that one.
size_t rng_dev_bytes();
void rng_dev_pack(const RngPlan &plan, void *host_dst);
size_t rng_workspace_bytes(const RngPlan &plan);
// uploads the ziggurat tables (once, synchronously) ahead of draws that a
// graph records
void rng_prepare();

// Draws into a caller's device array with the fixed per-device workspace:
// rng_scratch_setup(device) once per device with that device current, then
// rng_gaussians_noalloc (no device allocation; at most
// rng_noalloc_max_gaussians() values; overlapping calls on one GPU race).
void rng_scratch_setup(int device);
size_t rng_noalloc_max_gaussians();
void rng_gaussians_noalloc(int device, unsigned long seed, double *dev_out, size_t n, hipStream_t st);

// Raw 24-bit stream (testing / bit-exactness checks): the first n numbers.
void rng_raw(const RngPlan &plan, uint32_t *d_out, size_t n, void *workspace, size_t ws_bytes, hipStream_t st);

// Host reference pieces used by tests of the jump arithmetic.
void u576_mulmod_host(const U576 &a, const U576 &b, U576 &r);

}  // namespace gic
