// Displacement map w -> psi as an explicit linear operator, and its transpose.
//
// The forward map is the Zel'dovich path of the context (ICGenerator::done ->
// MultiLevelField / ZeldovichParticleGenerator, ic.hpp:1633-1648,
// zeldovich.hpp:152-253, level combination field.hpp:816-849): per level l
//   delta_l = sqrtP_l * r2c(w_l),
//   psi_{l,a} = c2r(i k_a F_l / k^2 * delta_l)        (F_0 = 1, F_l = HP_l)
// and for a zoom level psi_l += I_l [ LP_{l-1} psi_{l-1} ] with I_l the
// tricubic interpolation onto the fine cells.  Every step is linear in w, so
// the gradient of a scalar loss of psi is the transposed chain (no
// counterpart in the reference):
//   g_l = psibar_l + LP_l c2r r2c [ I_{l+1}^T g_{l+1} ]         (finest first)
//   wbar_l = c2r( sum_a conj(i k_a F_l sqrtP_l / k^2) r2c(g_{l,a}) )
// With the unitary FFT (N^-1.5 both ways) r2c^T = c2r on real fields, so the
// transpose of c2r diag(D) r2c is c2r diag(conj D) r2c.
#include "context.h"

namespace gic {

namespace {

template <typename T> __global__ void k_tab_mul(T *__restrict__ out, const T *__restrict__ a, long n) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] *= a[i];
}

// g[padded] += src[dense]
template <typename T> __global__ void k_add_dense(T *__restrict__ g, const T *__restrict__ src, int n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t n3 = size_t(n) * n * n;
  if (i >= n3) return;
  size_t row = i / n, z = i - row * n;
  g[row * (n + 2) + z] += src[i];
}

long table_entries(const LevelGeom &g) {
  long h = g.n / 2;
  return 3 * h * h + 1;
}

// out[|k|^2] = sqrtP(k) filter(k) / k^2 with the filter of kind (0 low, 1
// high, 2 unity) at kc, on level L
template <typename T> void adjoint_table(gic_context *c, Level &L, int kind, double kc, DevBuf &out) {
  long nent = table_entries(L.g);
  out.ensure(nent * sizeof(T));
  filter_table<T>(kind, kc, L.g.kmin(), nent, (T *)out.p, c->st, true);
  hipLaunchKernelGGL(k_tab_mul<T>, dim3((nent + 255) / 256), dim3(256), 0, c->st, (T *)out.p, (const T *)L.sqrtp.p,
                     nent);
  GIC_HIP(hipGetLastError());
}

void check_ready(gic_context *c) {
  if (c->levels.empty()) throw std::runtime_error("No grid has yet been initialised");
  if (!c->have_camb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
}

}  // namespace

template <typename T> void do_displacement_forward(gic_context *c, const T *const *noise, T *const *psi) {
  check_ready(c);
  size_t nl = c->levels.size();
  for (size_t l = 0; l < nl; ++l)
    if (!noise[l] || !psi[3 * l] || !psi[3 * l + 1] || !psi[3 * l + 2])
      throw std::runtime_error("displacement forward: null device pointer");
  // delta_l = sqrtP r2c(w_l), read straight from the caller's dense arrays
  for (auto &Lp : c->levels) {
    Level &L = *Lp;
    alloc_field(c, L.noise, L.g);
    build_sqrtp<T>(c, L);
  }
  for (size_t l = 0; l < nl; ++l) {
    Level &L = *c->levels[l];
    fft_r2c_post<T>((T *)L.noise.p, L.g.n, (const T *)L.sqrtp.p, c->st, c->hook(), noise[l]);
    L.noise_fourier = true;
    L.dc_pending = false;
    L.have_psi = L.have_psi2 = false;
  }
  c->seeded = c->drawn = c->power_applied = true;
  c->zeldovich_done = false;
  if (nl == 1) {
    // one level: each component's c2r writes the caller's array directly
    Level &L = *c->levels[0];
    long nent = table_entries(L.g);
    c->tab_c.ensure(nent * sizeof(T));
    filter_table<T>(2, 0.0, L.g.kmin(), nent, (T *)c->tab_c.p, c->st, true);
    alloc_field(c, L.psi[0], L.g);
    for (int a = 0; a < 3; ++a)
      fft_c2r_from<T>((const T *)L.noise.p, (T *)L.psi[0].p, L.g.n, FFT_PRE_ZA, nullptr, a, L.g.kmin(), c->st,
                      c->hook(), psi[a], 1.0);
  } else {
    int order = c->lpt_order;
    c->lpt_order = 1;
    try {
      do_zeldovich<T>(c);
    } catch (...) {
      c->lpt_order = order;
      throw;
    }
    c->lpt_order = order;
    for (size_t l = 0; l < nl; ++l) {
      Level &L = *c->levels[l];
      for (int a = 0; a < 3; ++a) unpad<T, T>((const T *)L.psi[a].p, L.g.n, psi[3 * l + a], c->st);
    }
  }
  GIC_HIP(hipStreamSynchronize(c->st));
}

template <typename T> void do_displacement_adjoint(gic_context *c, const T *const *psi_bar, T *const *noise_bar) {
  check_ready(c);
  size_t nl = c->levels.size();
  for (size_t l = 0; l < nl; ++l)
    if (!noise_bar[l] || !psi_bar[3 * l] || !psi_bar[3 * l + 1] || !psi_bar[3 * l + 2])
      throw std::runtime_error("displacement adjoint: null device pointer");
  size_t maxreals = 0;
  for (auto &Lp : c->levels) {
    Level &L = *Lp;
    build_sqrtp<T>(c, L);
    alloc_field(c, L.adj_acc, L.g);
    maxreals = std::max(maxreals, L.g.padded_reals());
  }
  for (size_t l = 0; l + 1 < nl; ++l) alloc_field(c, c->levels[l]->adj_u, c->levels[l]->g);
  c->adj_scratch.ensure(maxreals * sizeof(T));
  T *scratch = (T *)c->adj_scratch.p;
  // per level: the Zel'dovich operator's table sqrtP F_l / k^2, and the low
  // pass towards the next finer level (for level 0 with sqrtP / k^2 folded in)
  for (size_t l = 0; l < nl; ++l) {
    Level &L = *c->levels[l];
    adjoint_table<T>(c, L, l >= 1 ? 1 : 2, l >= 1 ? kcut_for_level(c, int(l) - 1) : 0.0, L.adj_tab);
    if (l + 1 < nl) {
      if (l == 0) {
        adjoint_table<T>(c, L, 0, kcut_for_level(c, 0), L.adj_lp);
      } else {
        L.adj_lp.ensure(table_entries(L.g) * sizeof(T));
        filter_table<T>(0, kcut_for_level(c, int(l)), L.g.kmin(), table_entries(L.g), (T *)L.adj_lp.p, c->st);
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    for (size_t li = nl; li-- > 0;) {
      Level &L = *c->levels[li];
      int n = L.g.n;
      const T *pb = psi_bar[3 * li + a];
      // the level's gradient g: psibar alone on the finest level, else
      // psibar + LP c2r r2c (I^T g_finer) held in adj_u (padded)
      const T *g_src = pb;      // dense
      T *g_pad = nullptr;       // padded
      if (li + 1 < nl && li >= 1) {
        g_pad = (T *)L.adj_u.p;
        fft_r2c_post<T>(g_pad, n, (const T *)L.adj_lp.p, c->st, c->hook());
        fft3d<T>(g_pad, n, +1, c->st, c->hook());
        hipLaunchKernelGGL(k_add_dense<T>, dim3((L.g.n3() + 255) / 256), dim3(256), 0, c->st, g_pad, pb, n);
        GIC_HIP(hipGetLastError());
      }
      if (li >= 1) {
        // push g to the coarser level through the transposed interpolation
        Level &C = *c->levels[li - 1];
        T *u = (T *)C.adj_u.p;
        fill<T>(u, C.g.padded_reals(), T(0), c->st);
        InterpGeom ig = interp_geom(L.g, C.g);
        if (g_pad) interp_adjoint<T>(g_pad, n, n + 2, u, ig, c->st);
        else interp_adjoint<T>(g_src, n, n, u, ig, c->st);
      }
      T *acc = (T *)L.adj_acc.p;
      if (g_pad) {
        fft_r2c_za_adjoint<T>(nullptr, g_pad, acc, n, (const T *)L.adj_tab.p, a, L.g.kmin(), a > 0, c->st, c->hook());
      } else {
        fft_r2c_za_adjoint<T>(g_src, scratch, acc, n, (const T *)L.adj_tab.p, a, L.g.kmin(), a > 0, c->st, c->hook());
        if (li == 0 && nl >= 2) {
          // level 0 under level 1: its LP part enters through r2c(I^T g_1)
          fft_r2c_za_adjoint<T>(nullptr, (T *)L.adj_u.p, acc, n, (const T *)L.adj_lp.p, a, L.g.kmin(), true, c->st,
                                c->hook());
        }
      }
    }
  }
  for (size_t l = 0; l < nl; ++l) {
    Level &L = *c->levels[l];
    fft_c2r_to<T>((T *)L.adj_acc.p, noise_bar[l], L.g.n, c->st, c->hook());
  }
  GIC_HIP(hipStreamSynchronize(c->st));
}

template void do_displacement_forward<float>(gic_context *, const float *const *, float *const *);
template void do_displacement_forward<double>(gic_context *, const double *const *, double *const *);
template void do_displacement_adjoint<float>(gic_context *, const float *const *, float *const *);
template void do_displacement_adjoint<double>(gic_context *, const double *const *, double *const *);

}  // namespace gic
