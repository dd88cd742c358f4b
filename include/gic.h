/*
 * C-ABI of libgic_hip.so: genetIC's field-generation path on an MI355X.
 *
 * The reference (pynbody/genetIC) has no FFI; its hot path is reached
 * through the paramfile command dispatcher (genetIC/src/main.cpp:22-172,
 * genetIC/src/tools/parser.hpp:159-211) which calls ICGenerator methods
 * (genetIC/src/ic.hpp).  Each entry point below replaces one such method,
 * with plain pointers and sizes, so a paramfile front-end (C++ or ctypes)
 * can drive it.  Conventions:
 *   - every call returns 0 on success, non-zero on error; the message is
 *     available from gic_last_error() (the front-end rethrows it, keeping the
 *     reference's 'Error "<what>" on line N' contract of parser.hpp:202-207);
 *   - the context owns all device memory; host pointers are borrowed for the
 *     duration of the call; calls are stream-ordered and synchronous on
 *     return;
 *   - one host thread per context.
 */
#ifndef GIC_H
#define GIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gic_context gic_context;

enum { GIC_FLOAT32 = 32, GIC_FLOAT64 = 64 };
enum {
  GIC_SEED_REAL_SPACE = 0,       /* random_seed_real_space, seed */
  GIC_SEED_FOURIER_SERIAL = 1,   /* random_seed_serial, seedfourier */
  GIC_SEED_FOURIER_PARALLEL = 2, /* random_seed, seedfourier_parallel */
  GIC_SEED_FOURIER_REVERSE = 3   /* seedfourier_reverse */
};
enum {
  GIC_FIELD_OVERDENSITY = 0,
  GIC_FIELD_PSI_X = 1,
  GIC_FIELD_PSI_Y = 2,
  GIC_FIELD_PSI_Z = 3,
  GIC_FIELD_PSI2_X = 4,   /* second-order displacements (lpt_order 2) */
  GIC_FIELD_PSI2_Y = 5,
  GIC_FIELD_PSI2_Z = 6
};
enum { GIC_REAL_SPACE = 0, GIC_FOURIER_SPACE = 1 };
enum { GIC_DTYPE_F32 = 0, GIC_DTYPE_F64 = 1 };

/* Context life cycle.  precision is GIC_FLOAT32 (the reference's -f) or
 * GIC_FLOAT64 (its default).  Replaces ICGenerator construction,
 * genetIC/src/main.cpp:200-241 and ic.hpp:178-220. */
int gic_create(gic_context **ctx, int device, int precision);
void gic_destroy(gic_context *ctx);
const char *gic_last_error(const gic_context *ctx);

/* Cosmology commands Om, Ol, Ob, hubble, s8, ns, zin (ic.hpp:229-340). */
int gic_set_cosmology(gic_context *ctx, double omega_m, double omega_lambda, double omega_b, double hubble,
                      double sigma8, double ns, double redshift);

/* camb <file> (ic.hpp:725-728, cosmology/camb.hpp:173-332): the file's
 * numbers flattened row by row, and its column count (7 or 13).  Freezes
 * ns, sigma8 and the growth factor at the current cosmology. */
int gic_set_camb_table(gic_context *ctx, const double *values, size_t nvalues, int ncols);

/* powerlaw_amplitude <A> (setPowerLawAmplitude, ic.hpp:731-734;
 * PowerLawPowerSpectrum, camb.hpp:145-168): P(k) = A k^ns for every species,
 * ns as set now, no sigma8 normalisation; replaces a CAMB table. */
int gic_set_power_law(gic_context *ctx, double amplitude);

/* base_grid <box> <n> (ic.hpp:365-386). */
int gic_init_base_grid(gic_context *ctx, double box, int n);
/* zoom_grid_with_origin_at <x0> <y0> <z0> <zoomfac> <n> (ic.hpp:501-601). */
int gic_init_zoom_grid_with_origin_at(gic_context *ctx, int x0, int y0, int z0, int zoomfac, int n);
int gic_num_levels(gic_context *ctx, int *out);
/* n, box size (thisGridSize), offsetLower[3] and cell size of a level. */
int gic_level_geometry(gic_context *ctx, int level, int *n, double *box, double *offset3, double *dx);

/* random_seed* / seedfourier* (ic.hpp:620-666). */
int gic_seed(gic_context *ctx, unsigned long seed, int mode);
/* Start a new realisation on the same grids with a new seed: forgets the
 * drawn noise and everything derived from it but keeps the device memory.
 * No reference counterpart (the reference seeds once per process); used to
 * run many seeds through one context (batches, benchmarks). */
int gic_reseed(gic_context *ctx, unsigned long seed, int mode);
/* initialiseAllRandomComponents (ic.hpp:1334-1352): draws the white noise of
 * every level; the real-space stream continues across levels. */
int gic_draw(gic_context *ctx);
/* zero_level <level> (ic.hpp:787-802). */
int gic_zero_level(gic_context *ctx, int level);
/* fix_power (setExactPowerSpectrumEnforcement, ic.hpp:669-671): the draw is
 * followed by enforceUnitVariance (multilevelfield.hpp:395-407,498-510), every
 * mode set to unit modulus.  Must precede the draw. */
int gic_set_exact_power_spectrum(gic_context *ctx, int enable);
/* reverse (ic.hpp:1652-1656): the white noise of every level changes sign. */
int gic_reverse(gic_context *ctx);
/* reverse_small_k <kmax> (ic.hpp:1687-1706): modes with 0 < k^2 < kmax^2 change sign. */
int gic_reverse_small_k(gic_context *ctx, double kmax);
/* import_level <level> <file> (importLevel, ic.hpp:805-820): host holds the n^3
 * real-space dark-matter overdensity of the level (z fastest); it replaces the
 * level's white noise after the transfer ratio P_dm^-1/2. */
int gic_import_level(gic_context *ctx, int level, const double *host, size_t n);
/* Splicing (ICGenerator::splice, ic.hpp:1657-1686; spliceOneLevel,
 * splice.hpp:37-118): a new white-noise field is drawn in real space from
 * new_seed, and on every level the noise becomes the field equal to the new
 * one on the flagged cells (in the overdensity basis) that minimises its
 * chi^2 distance to the old one elsewhere, by conjugate gradients
 * (cg.hpp:13-59, rtol 1e-6).  Before the output is generated.
 * cg_iterations (optional) receives the CG iterations, summed over levels. */
int gic_splice(gic_context *ctx, unsigned long new_seed, int *cg_iterations);
/* applyPowerSpec (ic.hpp:825-843): delta_k = w_k sqrt(P(k)) on every level. */
int gic_apply_power_spectrum(gic_context *ctx);
/* Zel'dovich displacements of every level and the combination of levels
 * (multilevelgenerator.hpp:78-128, zeldovich.hpp:152-253). */
int gic_zeldovich(gic_context *ctx);

/* Copy a field of a level to host memory (dump_grid, ic.hpp:852-918).
 * space GIC_REAL_SPACE: n^3 values in index order (x*n+y)*n+z;
 * GIC_FOURIER_SPACE: FFTW layout, n*n*(n/2+1) complex values.  The field is
 * transformed in place first if needed, as Field::toReal/toFourier do. */
int gic_copy_field(gic_context *ctx, int field, int level, int space, void *host, int dtype);
/* Overwrite a field (import_level, ic.hpp:809-821, and tests). */
int gic_set_field(gic_context *ctx, int field, int level, int space, const void *host, int dtype);

/* Order of the displacements: 1 = Zel'dovich (the reference's only option,
 * zeldovich.hpp:152-253), 2 = Zel'dovich plus second-order LPT.  No reference
 * counterpart; the paramfile command lpt_order sets it. */
int gic_set_lpt_order(gic_context *ctx, int order);

/* ---- constrained realisations ------------------------------------------------
 * Flagged cells of a level (the selection commands' result, ic.hpp:1442-1523),
 * read when a modification or a calculation is defined. */
int gic_set_flags(gic_context *ctx, int level, const uint64_t *cells, size_t n);
/* filtering_scale (ic.hpp:356-358): scale of the variance's high-pass filter. */
int gic_set_variance_filter_scale(gic_context *ctx, double scale);
/* calculate <name> (ic.hpp:1600-1606, modificationmanager.hpp:37-44): the current
 * value of overdensity, potential, vx|vy|vz, Lx|Ly|Lz (linearmodification.hpp)
 * or variance (quadraticmodification.hpp:152-221) on the flagged cells. */
int gic_calculate(gic_context *ctx, const char *name, double *out);
/* modify <name> absolute|relative <target> (ic.hpp:1613-1619,
 * modificationmanager.hpp:52-77); init_steps and precision drive the quadratic
 * iterations (ic.hpp:155-161). */
int gic_add_modification(gic_context *ctx, const char *name, int relative, double target, int init_steps,
                         double precision);
/* apply_modifications (modificationmanager.hpp:80-115): Hoffman-Ribak on the white
 * noise; chi^2 before and after. */
int gic_apply_modifications(gic_context *ctx, double *pre_chi2, double *post_chi2);
/* clear_modifications (modificationmanager.hpp:117-121). */
int gic_clear_modifications(gic_context *ctx);
/* chi2 (ic.hpp:1592-1597, multilevelfield.hpp:461-487): <w|M|w> of the white noise and
 * the degrees of freedom (multilevelgrid.hpp:160-172). */
int gic_chi2(gic_context *ctx, double *chi2, uint64_t *dof);

/* ---- Differentiation (no counterpart in the reference, which has no adjoint) ----
 * Device pointers: arrays of the context's precision (float for 32, double
 * for 64), dense n_l^3 per level with z fastest, on the context's device.
 * Calls are ordered on the context's stream and synchronous on return; the
 * caller finishes its own work on the arrays first.
 *
 * Forward displacement map w -> psi: noise[l] is the white noise of level l,
 * psi[3l+a] receives component a of its combined Zel'dovich displacement --
 * the path of done (ic.hpp:1633-1648) through applyPowerSpectrum, the
 * Zel'dovich generator (zeldovich.hpp:152-253) and the level combination
 * (field.hpp:816-849).  Replaces the context's fields, like done. */
int gic_displacement_forward(gic_context *ctx, const void *const *noise, void *const *psi);
/* Transpose of that map: noise_bar[l] = sum over levels and components of
 * (d psi / d w_l)^T psi_bar, the gradient of a loss of psi with respect to the
 * white noise.  Independent of the context's fields. */
int gic_displacement_adjoint(gic_context *ctx, const void *const *psi_bar, void *const *noise_bar);
/* Filtered variance V(w) of the flagged region for the white noise noise[l]
 * (calculate variance: QuadraticModification::calculateCurrentValue,
 * quadraticmodification.hpp:50-62, at the current filtering_scale) and, if
 * grad is non-null, its gradient 2 Q w into grad[l]. */
int gic_variance_gradient(gic_context *ctx, const void *const *noise, double *value, void *const *grad);
/* Quadratic modification steps (ModificationManager::performIterations,
 * modificationmanager.hpp:246-283): 0 the reference's linearised steps and
 * step-count trial (default), 1 an exact line search along the same gradient
 * direction, so that each step lands on its intermediate target. */
int gic_set_quadratic_line_search(gic_context *ctx, int enable);

/* ---- Batched realisations (no counterpart in the reference, which makes one
 * realisation per run) ----
 * For a context with exactly one level: nreal realisations, realisation r
 * drawn in real space from seeds[r] (the stream of gic_draw), then
 * constrained by ncons (1..32) top-hat mean-overdensity modifications with
 * absolute targets -- cons[5 (r ncons + i) + {0,1,2,3,4}] = centre x, y, z
 * (Mpc/h), radius (Mpc/h), target -- as select_sphere + modify overdensity
 * absolute + apply_modifications would (ic.hpp:1458-1500,1613-1631,
 * modificationmanager.hpp:160-312), then the transfer function and the
 * Zel'dovich displacement.  out (device, the context's precision) receives
 * per realisation four dense n^3 arrays: delta, psi_x, psi_y, psi_z at
 * out + r*4*n^3.  lanes (1..16) chains run concurrently on streams of their
 * own; use_graph != 0 records the whole batch into one HIP graph and launches
 * it once.  elapsed_ms (optional) receives the time from the first recorded
 * work to the end of the batch, graph recording included.  The context's own
 * fields are untouched; no chi^2 is logged. */
int gic_batch_constrained(gic_context *ctx, int nreal, const uint64_t *seeds, int ncons, const double *cons,
                          void *out, int lanes, int use_graph, double *elapsed_ms);

/* Power-spectrum histogram for dump_ps (dumpPowerSpectrum,
 * cosmology/parameters.hpp:112-205): over every (ix,iy,iz) in [-n/2,n/2]^3,
 * sum_power[m] += |delta_k|^2 and count[m] += 1 with m = ix^2+iy^2+iz^2;
 * theory[m] = P(k_m) n^3/V, the covariance the reference bins beside it.
 * nm must be 3(n/2)^2+1.  The binning into log-k bins is host work. */
int gic_mode_power(gic_context *ctx, int level, double *sum_power, double *count, double *theory, size_t nm);

/* Particle evaluation (ZeldovichParticleEvaluator::getParticle,
 * zeldovich.hpp:87-109, generator.hpp:49-53): for cells[i] (or start+i when
 * cells is NULL) of a level, pos = centroid + psi wrapped into the box
 * (Mpc/h) and vel = psi * H(a) sqrt(a) (gadget units), as float32. */
int gic_particles(gic_context *ctx, int level, const uint64_t *cells, size_t start, size_t count, float *pos,
                  float *vel);
/* Particle mass (1e10 Msol/h) and softening (Mpc/h) of a level (zeldovich.hpp:53-119). */
int gic_particle_mass_eps(gic_context *ctx, int level, double *mass, double *eps);
/* zeldovichVelocityToOffsetRatio (cosmology/parameters.hpp:83-106). */
int gic_velocity_ratio(gic_context *ctx, double *out);

/* In-place 3D transform of a caller-owned device buffer holding a padded
 * n*n*(n+2) field of the context's precision: FieldFourierManager::
 * performTransform (tools/numerics/fourier.hpp:479-518).  direction -1:
 * real -> Fourier, +1: Fourier -> real; both scale by n^-1.5. */
int gic_fft_inplace(gic_context *ctx, void *device_field, int n, int direction);

/* First n 24-bit outputs of gsl_rng_ranlxs2 seeded with seed (the integers
 * gsl_rng_get returns; randomfieldgenerator.hpp:43-45,115). */
int gic_raw_stream(gic_context *ctx, unsigned long seed, uint32_t *host_out, size_t n);
/* First n gsl_ran_gaussian_ziggurat(ranlxs2, 1) draws (randomfieldgenerator.hpp:191). */
int gic_gaussian_stream(gic_context *ctx, unsigned long seed, double *host_out, size_t n);

/* The first n Gaussians of gic_gaussian_stream, written to the caller's
 * device array device_out (n doubles) without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workspace is a fixed 13 MiB device array per GPU, held until the
 * process ends, so n is at most 1048576 (larger n: error, nothing drawn);
 * calls that overlap on one GPU share it and race.  Synchronous on return. */
int gic_gaussian_stream_noalloc(gic_context *ctx, unsigned long seed, double *device_out, size_t n);

/* Stream control and per-kernel event timing (hipEvents on the context's
 * stream around each launch). */
int gic_synchronize(gic_context *ctx);
int gic_set_event_timing(gic_context *ctx, int enable);
int gic_get_event_timing(gic_context *ctx, const char *kernel, double *total_ms, long *count);
int gic_reset_event_timing(gic_context *ctx);
/* All timed kernels as text lines "name<TAB>total_ms<TAB>count<TAB>bytes",
 * bytes being the launches' summed algorithmic HBM traffic (each input
 * element read once, each output element written once; 0 for regions that
 * group several kernels); writes at most buflen bytes (NUL-terminated) and
 * the full length to *needed. */
int gic_list_event_timing(gic_context *ctx, char *buf, size_t buflen, size_t *needed);
/* Device bytes currently held by the context. */
int gic_device_bytes(gic_context *ctx, size_t *out);

#ifdef __cplusplus
}
#endif
#endif /* GIC_H */
