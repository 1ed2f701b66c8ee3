/*
 * libsfsim_atmo: MI355X (gfx950) atmospheric lookup-table precompute for sfsim.
 *
 * C ABI only: plain structs of doubles, int shapes, float buffers and sizes.  Every entry point
 * returns 0 on success or a negative SFSIM_ATMO_ERR_* code; the message of the last failure on a
 * context is returned by sfsim_atmo_last_error.  No pointer passed in is kept after a call returns.
 * Calls on one context are serialised on that context's HIP stream and are synchronous for the
 * caller; different contexts may be used from different threads.
 *
 * The reference builds these tables in Clojure with closures over sfsim.atmosphere functions
 * (src/clj/sfsim/atmosphere_lut.clj:43-105); closures cannot cross a C ABI, so the boundary is
 * the table level: each `interpolate-function` call of atmosphere_lut.clj:74-101 is one entry
 * point here, and sfsim_atmo_generate_luts replaces generate-atmosphere-luts as a whole.
 *
 * Table layout of every float buffer: grid order of the reference's nested vectors (row-major
 * over the space's shape, dimension 0 outermost), 3 floats (RGB) per entry, i.e. exactly
 * (pack-matrices table) of matrix.clj:130-134.  4D tables handed to generate_luts' outputs are
 * additionally tiled 4D->2D as image.clj:299-312 does for the .scatter files.
 */
#ifndef SFSIM_ATMO_H
#define SFSIM_ATMO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFSIM_ATMO_OK 0
#define SFSIM_ATMO_ERR_ARGUMENT (-1)  /* invalid size, step count, null pointer, unknown space */
#define SFSIM_ATMO_ERR_DEVICE (-2)    /* no usable gfx950 device / HIP runtime error */
#define SFSIM_ATMO_ERR_ALLOC (-3)     /* device or host allocation failed */
#define SFSIM_ATMO_ERR_NONFINITE (-4) /* a parameter or input table holds NaN/Inf */
#define SFSIM_ATMO_ERR_IO (-5)        /* file could not be written/read completely */

/* Planet map of atmosphere.clj:64-67 + sphere.clj:25 (:sfsim.sphere/centre, /radius,
 * :sfsim.atmosphere/height, /brightness).  The centre must be the origin: the reference's
 * is-above-horizon? (atmosphere.clj:95-102) and every table space assume it. */
typedef struct sfsim_atmo_planet {
  double centre[3];
  double radius;
  double height;
  double brightness[3];
} sfsim_atmo_planet;

/* Scattering component of atmosphere.clj:35-39.  quotient: 1.0 when the reference omits it;
 * g: 0.0 when omitted. */
typedef struct sfsim_atmo_scatter {
  double scatter_base[3];
  double scatter_scale;
  double scatter_quotient;
  double scatter_g;
} sfsim_atmo_scatter;

/* Everything generate-atmosphere-luts hard-codes (atmosphere_lut.clj:20-62).  Every field sits at its
 * natural alignment and the size (232 bytes) is a multiple of 8, so a packed description of the fields in
 * this order (coffi's [::mem/struct ...], ctypes) has exactly this layout, in arrays too. */
typedef struct sfsim_atmo_params {
  sfsim_atmo_planet planet;
  sfsim_atmo_scatter mie;      /* scatter = [mie rayleigh] (atmosphere_lut.clj:63) */
  sfsim_atmo_scatter rayleigh;
  double intensity[3];
  int transmittance_shape[2];    /* [height elevation] */
  int surface_radiance_shape[2]; /* [height sun-elevation] */
  int scatter_shape[4];          /* [height elevation sun-elevation heading] */
  int ray_steps;                 /* integral-ray steps; also the half-sphere steps of surface-radiance */
  int sphere_steps;              /* integral-sphere steps of point-scatter */
  int iterations;                /* scattering-order iterations */
  int reserved;                  /* set to 0; makes the 4-byte tail explicit: 232 bytes, no padding anywhere */
} sfsim_atmo_params;

typedef struct sfsim_atmo_ctx sfsim_atmo_ctx;

/* Lifecycle (the reference's native library keeps global state behind jolt_init/jolt_destroy,
 * src/c/sfsim/jolt.cc:137-182; here the state lives in a context). */
int sfsim_atmo_create(int device, sfsim_atmo_ctx** ctx);
int sfsim_atmo_destroy(sfsim_atmo_ctx* ctx);
const char* sfsim_atmo_last_error(const sfsim_atmo_ctx* ctx);

/* How the single-scatter stage gets the optical depth of each sample's sun segment (the transmittance
 * toward the sun of atmosphere.clj:154-160).  rows x cols in [16, 8192]: from a table of the segments'
 * midpoint sums over (horizon distance, exit distance), built at the start of each build and read with
 * bilinear interpolation in log2; with 640 x 768 (the default, 3.75 MiB) the final tables of the 5-order
 * workload stay within 5.8e-5 of the FP64 oracle (768 x 768: 5.4e-5, slightly slower; 1024 x 1024:
 * 4.6e-5, 6% slower).  0 x 0: every segment
 * summed node by node, as the reference does (about 7x the single-scatter time, optical depths exact to
 * FP32 rounding).  No counterpart in the
 * reference, which always sums per segment.  Takes effect at the next build. */
int sfsim_atmo_set_sun_table(sfsim_atmo_ctx* ctx, int rows, int cols);
int sfsim_atmo_get_sun_table(const sfsim_atmo_ctx* ctx, int* rows, int* cols);

/* T = (interpolate-function transmittance transmittance-space) table (atmosphere_lut.clj:66-67,74;
 * atmosphere.clj:114-128, 291-319).  out: transmittance_shape entries x 3 floats. */
int sfsim_atmo_transmittance_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* out);

/* dE0 = surface-radiance-base table (atmosphere_lut.clj:68-69,75; atmosphere.clj:131-137, 335-365). */
int sfsim_atmo_surface_radiance_base_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* out);

/* First-order tables of atmosphere_lut.clj:70-73,77-78: ray-scatter of point-scatter-component for
 * Rayleigh (with phase) and of strength-component for Mie (no phase) (atmosphere.clj:162-200),
 * computed in one pass.  Either output may be NULL. */
int sfsim_atmo_single_scatter_tables(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* rayleigh_out,
                                     float* mie_strength_out);

/* dJ = point-scatter table (atmosphere_lut.clj:88,91; atmosphere.clj:203-222).
 * dS is the previous order's ray-scatter table.  When mie_strength is non-NULL the first-iteration
 * function of atmosphere_lut.clj:79-84 is used instead: dS(q) + mie_strength(q) * phase(mie, w.l).
 * dE is the previous order's surface-radiance table. */
int sfsim_atmo_point_scatter_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dS,
                                   const float* mie_strength, const float* dE, float* out);

/* dS' = ray-scatter table of an interpolated dJ table (atmosphere_lut.clj:92,94; atmosphere.clj:192-200). */
int sfsim_atmo_ray_scatter_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dJ, float* out);

/* dE' = surface-radiance table of dS (atmosphere_lut.clj:89,93; atmosphere.clj:225-230), half-sphere
 * steps = ray_steps.  mie_strength as for point_scatter. */
int sfsim_atmo_surface_radiance_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dS,
                                      const float* mie_strength, float* out);

/* Re-sampling f(backward(i)) of interpolated tables on one space, the accumulations of
 * atmosphere_lut.clj:95-97 (b != NULL: a + b) and the final make-lookup-table calls of 98-101
 * (b == NULL).  space: 2 = transmittance, 3 = surface radiance, 4 = ray/point scatter. */
#define SFSIM_ATMO_SPACE_TRANSMITTANCE 2
#define SFSIM_ATMO_SPACE_SURFACE_RADIANCE 3
#define SFSIM_ATMO_SPACE_SCATTER 4
int sfsim_atmo_resample_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int space, const float* a,
                              const float* b, float* out);

/* generate-atmosphere-luts (atmosphere_lut.clj:43-105) end to end.  Outputs in the exact byte
 * order of the four .scatter files: transmittance [h][e] RGB, surface-radiance [h][s] RGB,
 * ray-scatter and mie-strength 4D->2D tiled (rows h*S+s, cols e*A+a) RGB.  Any output may be NULL. */
int sfsim_atmo_generate_luts(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* transmittance,
                             float* surface_radiance, float* ray_scatter, float* mie_strength);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): sfsim_atmo_generate_luts with the four
 * final tables written to the caller's DEVICE buffers (layouts and sizes as generate_luts, none NULL) and
 * every intermediate table in one static device array of SFSIM_ATMO_NOALLOC_SCRATCH_BYTES, which each GPU
 * the library runs on holds until the process ends.  The call allocates and frees no device memory; it
 * copies the quadrature directions and the ray order from the host and returns after the stream has drained.  A
 * workload whose intermediate tables do not fit the array is rejected with SFSIM_ATMO_ERR_ARGUMENT (the
 * reference's production workload takes 277 MiB of it, the 32x64x32x8 5-order workload 142 MiB).  All
 * calls on one GPU share its array, which keeps the last call's contents: calls that overlap on one GPU
 * (several contexts or threads) race and must be serialised by the caller. */
#define SFSIM_ATMO_NOALLOC_SCRATCH_BYTES ((size_t)288 << 20)
int sfsim_atmo_generate_luts_noalloc(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* transmittance,
                                     float* surface_radiance, float* ray_scatter, float* mie_strength);

/* Device-resident build for callers that keep the tables on the GPU (and for timing): runs the
 * whole generate-atmosphere-luts pipeline into context-owned device buffers without host copies;
 * sfsim_atmo_fetch then copies one final table out (which: 0 = transmittance, 1 = surface
 * radiance, 2 = ray scatter, 3 = mie strength; layouts as generate_luts).  The launches are captured
 * into a HIP graph on the first build with given parameters and replayed by later builds with the
 * same parameters (same results: the kernels are deterministic). */
int sfsim_atmo_build(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p);
/* The same build launched kernel by kernel with HIP events around each launch (fills the timings
 * that sfsim_atmo_stage_timings reports). */
int sfsim_atmo_build_timed(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p);
int sfsim_atmo_fetch(sfsim_atmo_ctx* ctx, int which, float* out, size_t count);

/* generate-atmosphere-luts for `count` atmospheres at once: the reference builds one atmosphere at a time,
 * parallel only over the first table dimension (pmap, interpolate.clj:72); here every launch of the
 * pipeline covers the entries of all the atmospheres (grid = entries x atmospheres).  The atmospheres must
 * share the planet's radius and height, the table shapes, the step counts and the iterations; their
 * scattering components, brightness and intensity may differ.  The geometry that depends only on the planet
 * and the grid is computed once; every other table exists once per atmosphere in context-owned device
 * memory.  Needs the sun optical-depth table (sfsim_atmo_set_sun_table with rows > 0, the default).
 * sfsim_atmo_fetch_batch copies one final table of one atmosphere out (which and layouts as
 * sfsim_atmo_fetch).  Results equal those of sfsim_atmo_generate_luts for each atmosphere alone. */
int sfsim_atmo_build_batch(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* atmospheres, int count);
int sfsim_atmo_fetch_batch(sfsim_atmo_ctx* ctx, int atmosphere, int which, float* out, size_t count);

/* Kernel durations of the last sfsim_atmo_build_timed (or seeded build), measured with HIP events
 * on the context stream and summed per kernel kind: ms[k] milliseconds and launches[k] launches of kind k, for
 * k < min(n, number of kinds); returns the number of kinds.  Kind names: sfsim_atmo_stage_name(k). */
int sfsim_atmo_stage_timings(const sfsim_atmo_ctx* ctx, double* ms, int* launches, int n);
const char* sfsim_atmo_stage_name(int kind);

/* Progress of a build (util.clj:441-449 progress-wrap and the "Iteration k/N" line of
 * atmosphere_lut.clj:80, which the reference prints while it evaluates the tables): fn(user, stage,
 * iteration, iterations) is called once the device has finished each stage of a build, in order, on a
 * thread of the HIP runtime (it must not call into the library or HIP): SFSIM_ATMO_PROGRESS_FIRST_ORDER
 * after the transmittance, surface-radiance-base and single-scattering tables (iteration 0), then
 * SFSIM_ATMO_PROGRESS_ITERATION after each scattering order (iteration 1..iterations), then
 * SFSIM_ATMO_PROGRESS_FINAL after the final tables (iteration 0).  Applies to generate_luts, build,
 * build_timed and build_batch of this context; while a callback is set, build launches the stages
 * directly instead of replaying its graph (same results).  fn NULL removes the callback. */
#define SFSIM_ATMO_PROGRESS_FIRST_ORDER 0
#define SFSIM_ATMO_PROGRESS_ITERATION 1
#define SFSIM_ATMO_PROGRESS_FINAL 2
typedef void (*sfsim_atmo_progress_fn)(void* user, int stage, int iteration, int iterations);
int sfsim_atmo_set_progress(sfsim_atmo_ctx* ctx, sfsim_atmo_progress_fn fn, void* user);

/* Work the kernels perform for one build of p (extinction samples, 4D and 2D lookups), counted on
 * the device by an instrumented build (not the timed one).  counts: [0] extinction samples of the
 * single-scatter kernel, [1] extinction samples of all kernels, [2] 4D lookups, [3] 2D lookups. */
int sfsim_atmo_count_work(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, unsigned long long* counts);

/* ---- Learned stand-in for scattering orders >= 2 (no counterpart in the reference) -------------
 * A small MLP (4 LUT coordinates + 6 atmosphere parameters, 8-band Fourier features, 3 x 128 SiLU,
 * bf16 matrix-core products with FP32 accumulation, Adam in FP32) trained on tables that the plain
 * GPU iteration produces.  It predicts the order sums P_k = dS_1 + ... + dS_k (k = 1..5) of the
 * scattering tables; the seeded build applies the iteration (atmosphere_lut.clj:86-97) m times to the
 * predicted partial sum of 5 - m orders and falls back to the plain iteration when the network's own
 * P_5 differs from the result by more than the tolerance. */
size_t sfsim_atmo_surrogate_size(void); /* number of FP32 parameters */
/* (Re)initialise the network deterministically from seed; batch: training / inference rows per pass. */
int sfsim_atmo_surrogate_reset(sfsim_atmo_ctx* ctx, unsigned long long seed, int batch);
int sfsim_atmo_surrogate_weights(sfsim_atmo_ctx* ctx, float* out, size_t count);
/* trained != 0 marks the weights as usable by sfsim_atmo_build_seeded. */
int sfsim_atmo_surrogate_set_weights(sfsim_atmo_ctx* ctx, const float* in, size_t count, int trained);
/* Forward pass over rows of 10 inputs in [0, 1] (outputs: rows x 16); with targets (rows x 16, first
 * 15 used) also the loss (mean over rows of half the squared error) and its gradient (size() floats). */
int sfsim_atmo_surrogate_eval(sfsim_atmo_ctx* ctx, const float* inputs, const float* targets, int rows,
                              float* outputs, float* grad, float* loss);
/* Trains on the first five orders of the plain builds of `count` atmospheres (iterations >= 5) sharing one
 * scatter shape: `steps` Adam
 * steps of one batch each, learning rate lr with a linear warm-up over the first 5% of the steps and
 * cosine decay after it (a fresh network's output biases start at the targets' means); losses
 * (nullable) receives one value per step. */
int sfsim_atmo_surrogate_train(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* atmospheres, int count, int steps,
                               double lr, unsigned long long seed, float* losses);
/* Device-resident build of p->iterations (5..32) scattering orders seeded by the network, which predicts
 * the order sums P_1..P_5: the partial sum of iterations - applications orders (1..4) comes from the network
 * and `applications` exact applications of the iteration (atmosphere_lut.clj:86-97) follow, so applications
 * is in iterations-4 .. iterations-1 (outputs as sfsim_atmo_build, read with sfsim_atmo_fetch).  residual
 * receives the network's own P_5 against the computed one, relative to the ray-scatter table; fallback is 1
 * when it exceeded the tolerance and the plain iteration ran instead. */
int sfsim_atmo_build_seeded(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int applications, double tolerance,
                            double* residual, int* fallback);
/* The same with the order sums P_1..P_5 supplied by the caller (5 RGB scatter tables one after the
 * other, e.g. those of a neighbouring atmosphere in a parameter sweep) instead of the network's. */
int sfsim_atmo_build_from_sums(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int applications, const float* sums,
                               double tolerance, double* residual, int* fallback);
/* The order sums P_1..P_5 of the plain iteration (iterations >= 5; only the first five orders run), and
 * the network's prediction of them: sums holds 5 scatter tables (RGB) one after the other. */
int sfsim_atmo_order_sums(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* sums);
int sfsim_atmo_surrogate_predict(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* sums);

/* util.clj:227-240 spit-floats / util.clj:188-203 slurp-floats: raw little-endian float32, no header. */
int sfsim_atmo_write_scatter(const char* path, const float* data, size_t count);
int sfsim_atmo_read_scatter(const char* path, float* data, size_t count);

/* Number of directions of integral-sphere (half = 0) / integral-half-sphere (half = 1) with the
 * given steps (sphere.clj:70-105). */
int sfsim_atmo_sphere_directions(int half, int steps);

/* The standard atmosphere that sfsim's aerodynamics reads (atmosphere.clj:425-506, after Hull,
 * "Fundamentals of Airplane Flight Mechanics"): temperature (K), pressure (Pa) and density (kg/m^3) at a
 * height above sea level (m), from the troposphere formulas up to 36089 ft, the lower-stratosphere ones up to
 * 65617 ft and the upper-stratosphere ones above (temperature-at-height, pressure-at-height,
 * density-at-height).  Any output may be NULL; a non-finite height is SFSIM_ATMO_ERR_NONFINITE.  Host code:
 * the reference's callers evaluate it one flight state at a time (aerodynamics.clj:424). */
int sfsim_atmo_standard_atmosphere(double height, double* temperature, double* pressure, double* density);
/* speed-of-sound (atmosphere.clj:508-511): 331.3 sqrt(T / 273.15) m/s for a temperature T in K. */
double sfsim_atmo_speed_of_sound(double temperature);

#ifdef __cplusplus
}
#endif

#endif /* SFSIM_ATMO_H */
