// C ABI of libsfsim_atmo: contexts, parameter validation, quadrature direction sets, the
// generate-atmosphere-luts driver (atmosphere_lut.clj:43-105) and the per-table entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sfsim_atmo.h"
#include "device_params.hpp"
#include "geometry.hpp"
#include "kernels.hpp"
#include "surrogate.hpp"

using namespace sfsim;

namespace {

enum Kind {
  kTransmittance = 0,
  kSurfaceRadianceBase,
  kSingleScatter,
  kDirectionGeometry,
  kPointScatter,
  kSurfaceRadiance,
  kRayScatter,
  kAccumulate,
  kFinal,
  kSurrogate,
  kRayGeometry,
  kNumKinds
};

const char* const kKindNames[kNumKinds] = {"transmittance",      "surface_radiance_base", "single_scatter",
                                           "direction_geometry", "point_scatter",         "surface_radiance",
                                           "ray_scatter",        "accumulate",            "final_resample",
                                           "surrogate",          "ray_geometry"};

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

// Default size of the sun optical-depth table: 640 horizon-distance rows x 768 exit-distance columns
// (3.75 MiB, within one XCD's 4 MiB L2): the final tables stay within 5.8e-5 of the FP64 oracle at the
// full workload (768 x 768: 5.4e-5 and ~5 us slower per build, its 4.5 MiB refetched beyond the L2;
// 1024 x 1024: 4.6e-5, 6% slower; fewer columns cost more accuracy than fewer rows: 768 x 640 9.8e-5).
constexpr int kSunRows = 640, kSunCols = 768;

// Device memory of sfsim_atmo_generate_luts_noalloc: every intermediate table of a build lives here, so
// that entry point allocates nothing.  Each GPU the library runs on holds its own copy of the array, at
// its own address, from the module's load until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: kernels get the address that hipGetSymbolAddress returned for
// the current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[SFSIM_ATMO_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current: looked up once per device.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace

// One progress report queued on the stream (sfsim_atmo_set_progress); the runtime calls the host function
// with its address once the stream reaches it.
struct ProgressNote {
  sfsim_atmo_progress_fn fn;
  void* user;
  int stage, iteration, iterations;
};

struct sfsim_atmo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  // progress callback and the reports of the build in flight (a deque keeps their addresses stable)
  sfsim_atmo_progress_fn progress_fn = nullptr;
  void* progress_user = nullptr;
  std::deque<ProgressNote> notes;
  // device address of the scratch array on this context's device (sfsim_atmo_generate_luts_noalloc)
  void* scratch = nullptr;
  // set on the stack contexts of batched builds and of sfsim_atmo_generate_luts_noalloc: every buffer is
  // pre-assigned (in the batch's memory or the scratch array), and a request that does not fit is an error
  // instead of a hipMalloc
  bool no_alloc = false;
  // tables (float4 per entry)
  DeviceBuffer T, dE[2], Eacc[2], R1, M1, dJ, dS[2], Sacc[2];
  // packed outputs (RGB floats): 0 transmittance, 1 surface radiance, 2 ray scatter (tiled), 3 mie (tiled)
  DeviceBuffer out[4];
  size_t out_count[4] = {0, 0, 0, 0};
  bool built = false;
  // stage-entry staging
  DeviceBuffer stage_in[3], stage_out4, stage_rgb;
  // quadrature directions
  DeviceBuffer sphere_dirs, half_dirs;
  DeviceBuffer geom, slab, slab_mie;  // point-scatter direction geometry and folded dS slabs
  int sphere_key = -1, half_key = -1, sphere_n = 0, half_n = 0;
  // single scatter's sun optical-depth table (rows x cols of log2 node sums; 0 x 0 sums every segment)
  int sun_rows = kSunRows, sun_cols = kSunCols;
  DeviceBuffer sun_depth;
  DeviceBuffer ray_geo;  // per-(view ray, sample) geometry of single and ray scatter
  DeviceBuffer rmap;     // per-entry 4D re-sampling lookups of the pipelines
  DeviceBuffer sunmap;   // per-entry sun terms of the point- and ray-scatter kernels (with rmap)
  DeviceBuffer sgeo;     // half-sphere direction geometry per surface-radiance height
  DeviceBuffer ray_order;  // ray-scatter visiting order (ascending impact parameter)
  DeviceBuffer Ssum[2];    // S + dS at the grid points, written by ray scatter (alternate iterations)
  double ray_order_key[4] = {-1.0, -1.0, -1.0, -1.0};  // (R, H, NH, NE) it was computed for
  DeviceBuffer counters;
  // surrogate: network state, training set (codes, targets), order sums, predictions, residual scratch
  Surrogate sg;
  DeviceBuffer codes, targets, sumS, cap[kMlpOrders], pred[kMaxSeededOrders], resid;
  double last_residual = 0.0;
  // graph of the last build() (replayed while parameters and buffers stay the same)
  hipGraphExec_t graph_exec = nullptr;
  sfsim_atmo_params graph_params;
  DevParams graph_dev;
  unsigned long long alloc_gen = 0, graph_gen = 0;
  // graph of the last network-seeded build
  hipGraphExec_t seeded_exec = nullptr;
  sfsim_atmo_params seeded_params;
  DevParams seeded_dev;
  unsigned long long seeded_gen = 0;
  int seeded_m = 0;
  const void* seeded_bufs[3] = {nullptr, nullptr, nullptr};
  // batched builds: one block of buffers per atmosphere (batch_stride bytes apart) after the shared
  // geometry, the per-atmosphere parameters on the device, and the final tables of atmosphere 0
  DeviceBuffer batch_mem, batch_params;
  int batch_count = 0;
  size_t batch_stride = 0;
  void* batch_out[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t batch_out_count[4] = {0, 0, 0, 0};
  bool batch_built = false;
  // timing
  std::vector<hipEvent_t> events;
  std::vector<int> event_kind;
  double kind_ms[kNumKinds] = {};
  int kind_launches[kNumKinds] = {};
};

namespace {

int fail(sfsim_atmo_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

int hip_fail(sfsim_atmo_ctx* ctx, const char* what, hipError_t e) {
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SFSIM_ATMO_ERR_ALLOC : SFSIM_ATMO_ERR_DEVICE;
  return fail(ctx, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                 \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, #expr, e_); \
  } while (0)

#define TRY(expr)             \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != 0) return rc_; \
  } while (0)

int ensure(sfsim_atmo_ctx* ctx, DeviceBuffer& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return 0;
  if (ctx->no_alloc)
    return fail(ctx, SFSIM_ATMO_ERR_ALLOC, "a build requested a buffer outside its preassigned layout");
  ctx->alloc_gen++;  // captured graphs hold buffer addresses
  if (b.ptr) {
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.ptr, bytes < 256 ? 256 : bytes);
  if (e != hipSuccess) return hip_fail(ctx, "hipMalloc", e);
  b.bytes = bytes < 256 ? 256 : bytes;
  return 0;
}

void release(DeviceBuffer& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

bool finite3(const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

long long count2(const int* s) { return (long long)s[0] * s[1]; }
long long count4(const int* s) { return (long long)s[0] * s[1] * s[2] * s[3]; }

int validate(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  if (!p) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "params is NULL");
  const sfsim_atmo_planet& pl = p->planet;
  if (!finite3(pl.centre) || !std::isfinite(pl.radius) || !std::isfinite(pl.height) || !finite3(pl.brightness) ||
      !finite3(p->intensity))
    return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "planet or intensity holds a non-finite value");
  const sfsim_atmo_scatter* comps[2] = {&p->mie, &p->rayleigh};
  for (const sfsim_atmo_scatter* c : comps)
    if (!finite3(c->scatter_base) || !std::isfinite(c->scatter_scale) || !std::isfinite(c->scatter_quotient) ||
        !std::isfinite(c->scatter_g))
      return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "scattering component holds a non-finite value");
  if (pl.centre[0] != 0.0 || pl.centre[1] != 0.0 || pl.centre[2] != 0.0)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "planet centre must be the origin (atmosphere.clj:95-102 assumes it)");
  if (!(pl.radius > 0.0) || !(pl.height > 0.0)) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "radius and height must be > 0");
  // the FP64 altitude series needs (Rt^2 - R^2) / R^2 <= 0.1
  if (pl.height * (2.0 * pl.radius + pl.height) > 0.1 * pl.radius * pl.radius)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "atmosphere height must be below 4.9% of the radius");
  for (const sfsim_atmo_scatter* c : comps) {
    if (!(c->scatter_scale > 0.0)) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "scatter-scale must be > 0");
    if (!(c->scatter_quotient > 0.0)) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "scatter-quotient must be > 0");
    if (!(std::fabs(c->scatter_g) < 1.0)) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "|scatter-g| must be < 1");
  }
  const int* shapes[3] = {p->transmittance_shape, p->surface_radiance_shape, p->scatter_shape};
  const int dims[3] = {2, 2, 4};
  for (int s = 0; s < 3; ++s)
    for (int d = 0; d < dims[s]; ++d)
      if (shapes[s][d] < 2 || shapes[s][d] > 65536)
        return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "every table dimension must be in [2, 65536]");
  if (count4(p->scatter_shape) > (1LL << 31) / 4 || count2(p->transmittance_shape) > (1LL << 30) ||
      count2(p->surface_radiance_shape) > (1LL << 30))
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "table too large");
  if ((long long)p->scatter_shape[2] * p->scatter_shape[3] > 1024)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sun-elevation x heading size must be <= 1024");
  if (p->ray_steps < 1 || p->ray_steps > kMaxSteps)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "ray_steps must be in [1, 256]");
  if (p->iterations < 0 || p->iterations > 1000) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "iterations must be in [0, 1000]");
  return 0;
}

// Direction sets need rings: integral-sphere has steps >> 1, integral-half-sphere steps >> 2
// (sphere.clj:96-105); with none the reference's reduction over no samples fails.
int validate_sphere(sfsim_atmo_ctx* ctx, int steps) {
  if ((steps >> 1) < 1 || sfsim_atmo_sphere_directions(0, steps) > kMaxSphereDirs)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sphere_steps must give 1..1000 directions (steps >= 2)");
  return 0;
}
int validate_half_sphere(sfsim_atmo_ctx* ctx, int steps) {
  if ((steps >> 2) < 1) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "ray_steps must be >= 4 for the surface-radiance half sphere");
  return 0;
}

bool finite_floats(const float* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// Polynomial of degree `deg` (2 or 3) through the deg + 1 Chebyshev nodes of [0, xh],
// xh = (Rt^2 - R^2) / R^2, of f(x) = k (R / 2) p(x), p(x) = 2 (sqrt(1 + x) - 1) / x = 2 / (1 + sqrt(1 + x)),
// so that k h = x f(x) for the altitude h = sqrt(R^2 + x R^2) - R.  Returns the largest relative error of
// the fit over 2001 points of [0, xh].  Cubic: 4e-10 for a 100 km atmosphere on Earth, 4e-8 at the
// largest xh = 0.1 the parameters may have; quadratic: 7e-8 for 100 km.
double altitude_fit(double R, double Rt, double k, int deg, double out[4]) {
  const double xh = (Rt * Rt - R * R) / (R * R);
  const int m = deg + 1;
  auto f = [&](double x) { return k * 0.5 * R * 2.0 / (1.0 + std::sqrt(1.0 + x)); };
  double A[4][5] = {};
  for (int i = 0; i < m; ++i) {
    const double x = 0.5 * xh * (1.0 + std::cos((2 * i + 1) * M_PI / (2.0 * m)));
    double v = 1.0;
    for (int j = 0; j < m; ++j, v *= x) A[i][j] = v;
    A[i][4] = f(x);
  }
  for (int c = 0; c < m; ++c) {  // Gauss-Jordan with partial pivoting
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    for (int j = 0; j < 5; ++j) std::swap(A[c][j], A[piv][j]);
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double q = A[r][c] / A[c][c];
      for (int j = c; j < 5; ++j) A[r][j] -= q * A[c][j];
    }
  }
  for (int j = 0; j < 4; ++j) out[j] = j < m ? A[j][4] / A[j][j] : 0.0;
  double worst = 0.0;
  for (int i = 0; i <= 2000; ++i) {
    const double x = xh * i / 2000.0;
    const double y = out[0] + x * (out[1] + x * (out[2] + x * out[3]));
    worst = std::fmax(worst, std::fabs(y / f(x) - 1.0));
  }
  return worst;
}

DevParams make_dev_params(const sfsim_atmo_params* p) {
  DevParams P;
  std::memset(&P, 0, sizeof(P));
  P.g.R = p->planet.radius;
  P.g.H = p->planet.height;
  P.g.Rt = p->planet.radius + p->planet.height;
  P.g.rho_max = std::sqrt(sqr(P.g.Rt) - sqr(P.g.R));
  const double log2e = 1.4426950408889634;
  for (int c = 0; c < 3; ++c) {
    P.ext_mie[c] = (float)(p->mie.scatter_base[c] / p->mie.scatter_quotient);
    P.ext_ray[c] = (float)(p->rayleigh.scatter_base[c] / p->rayleigh.scatter_quotient);
    P.ext_mie_l2[c] = (float)(log2e * p->mie.scatter_base[c] / p->mie.scatter_quotient);
    P.ext_ray_l2[c] = (float)(log2e * p->rayleigh.scatter_base[c] / p->rayleigh.scatter_quotient);
    P.sca_mie[c] = (float)p->mie.scatter_base[c];
    P.sca_ray[c] = (float)p->rayleigh.scatter_base[c];
    P.brightness_over_pi[c] = (float)(p->planet.brightness[c] / M_PI);
    P.intensity[c] = (float)p->intensity[c];
  }
  P.k_mie = (float)(-log2e / p->mie.scatter_scale);
  P.k_ray = (float)(-log2e / p->rayleigh.scatter_scale);
  P.inv_R = 1.0 / P.g.R;
  P.ray_over_mie = (float)(p->mie.scatter_scale / p->rayleigh.scatter_scale);
  altitude_fit(P.g.R, P.g.Rt, -log2e / p->mie.scatter_scale, 3, P.alt_poly_d);
  double quad[4];
  P.alt_deg = altitude_fit(P.g.R, P.g.Rt, -log2e / p->mie.scatter_scale, 2, quad) <= 1e-7 ? 2 : 3;
  for (int j = 0; j < 4; ++j) P.alt_poly[j] = (float)(P.alt_deg == 2 ? quad[j] : P.alt_poly_d[j]);
  P.inv_scale_mie = 1.0 / p->mie.scatter_scale;
  P.inv_scale_ray = 1.0 / p->rayleigh.scatter_scale;
  P.g_mie = (float)p->mie.scatter_g;
  P.g_ray = (float)p->rayleigh.scatter_g;
  P.sun_inv_den = (float)(1.0 / (1.0 - std::exp(-3.6)));
  for (int d = 0; d < 2; ++d) {
    P.nT[d] = p->transmittance_shape[d];
    P.nE[d] = p->surface_radiance_shape[d];
  }
  for (int d = 0; d < 4; ++d) P.nS[d] = p->scatter_shape[d];
  P.ray_steps = p->ray_steps;
  P.inv_steps = 1.0 / p->ray_steps;
  // SFSIM_ATMO_GRAZE_HEIGHT (metres, >= 0) overrides the altitude below which a segment's lowest
  // point selects the FP64 optical-depth path; 0 keeps every segment in FP32.
  P.graze_height = 2000.0;
  if (const char* env = std::getenv("SFSIM_ATMO_GRAZE_HEIGHT")) {
    char* end = nullptr;
    double v = std::strtod(env, &end);
    if (end != env && std::isfinite(v) && v >= 0.0) P.graze_height = v;
  }
  P.graze_u = P.graze_height * (2.0 * P.g.R + P.graze_height);
  return P;
}

// make_dev_params plus the context's sun optical-depth table (allocated by ensure_sun_depth).
DevParams dev_params(const sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  DevParams P = make_dev_params(p);
  if (ctx->sun_rows > 0) {
    P.sun_rows = ctx->sun_rows;
    P.sun_cols = ctx->sun_cols;
    P.sun_depth = (float2*)ctx->sun_depth.ptr;
  }
  P.ray_geo = (char*)ctx->ray_geo.ptr;
  P.rmap = (char*)ctx->rmap.ptr;  // filled by launch_resample_map at the start of a pipeline
  if (ctx->sunmap.ptr) {          // filled with rmap
    P.sun_ps = (float4*)ctx->sunmap.ptr;
    P.sun_rs = P.sun_ps + count4(p->scatter_shape);
    P.sun_ss = (double2*)(P.sun_rs + count4(p->scatter_shape));
  }
  P.sgeo = (float4*)ctx->sgeo.ptr;  // filled by launch_surface_geometry before the surface-radiance stages
  P.ray_order = (const int*)ctx->ray_order.ptr;  // set by ensure_ray_order
  return P;
}

// The view rays of the scatter grid in ascending impact parameter r0 sqrt(1 - mu0^2) (ties in grid
// order), uploaded when the planet or the grid changes.
int ensure_ray_order(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  const double key[4] = {p->planet.radius, p->planet.height, (double)p->scatter_shape[0],
                         (double)p->scatter_shape[1]};
  if (ctx->ray_order.ptr && std::memcmp(key, ctx->ray_order_key, sizeof(key)) == 0) return 0;
  const DevParams P = make_dev_params(p);
  const int NH = p->scatter_shape[0], NE = p->scatter_shape[1];
  std::vector<double> b((size_t)NH * NE);
  for (int ih = 0; ih < NH; ++ih) {
    const double r0 = index_to_radius(P.g, NH, ih);
    for (int ie = 0; ie < NE; ++ie) {
      double se, ce;
      bool above;
      index_to_elevation(P.g, NE, r0, ie, se, ce, above);
      b[(size_t)ih * NE + ie] = r0 * std::fabs(ce);
    }
  }
  std::vector<int> order(b.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return b[u] < b[v]; });
  TRY(ensure(ctx, ctx->ray_order, order.size() * sizeof(int)));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->ray_order.ptr, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(ctx->ray_order_key, key, sizeof(key));
  return 0;
}

// The surface-radiance direction geometry buffer (after ensure_dirs for the half sphere).
int ensure_surface_geometry(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  return ensure(ctx, ctx->sgeo, (size_t)p->surface_radiance_shape[0] * ctx->half_n * sizeof(float4));
}

int ensure_sun_depth(sfsim_atmo_ctx* ctx) {
  if (ctx->sun_rows <= 0) return 0;
  return ensure(ctx, ctx->sun_depth, (size_t)ctx->sun_rows * ctx->sun_cols * sizeof(float2));
}

// The sun table and the per-(view ray, sample) geometry buffer of the single/ray-scatter kernels.
int ensure_scatter_buffers(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  TRY(ensure_sun_depth(ctx));
  TRY(ensure_ray_order(ctx, p));
  TRY(ensure(ctx, ctx->rmap, resample_map_bytes(make_dev_params(p))));
  TRY(ensure(ctx, ctx->sunmap, (2 * sizeof(float4) + sizeof(double2)) * (size_t)count4(p->scatter_shape)));
  return ensure(ctx, ctx->ray_geo, ray_geometry_bytes(make_dev_params(p)));
}

// sphere.clj:62-105 for the normal (1 0 0) of every grid point: rings theta_i = range (i + 1/2) / N,
// weight (cos(theta - d) - cos(theta + d)) 2 pi / K_i, K_i = ceil(sin(theta_i) steps), phi_j = 2 pi (j + 1/2) / K_i,
// direction x n + y o1 + z o2 with the frame of oriented-matrix (matrix.clj:207-213, quaternion.clj:166-171).
std::vector<Dir> sphere_directions(int half, int steps) {
  std::vector<Dir> dirs;
  int theta_steps = half ? (steps >> 2) : (steps >> 1);
  double theta_range = half ? M_PI / 2 : M_PI;
  double n[3] = {1, 0, 0};
  // orthogonal: basis vector with the smallest |n . e| (first on ties), o1 = normalize(n x e), o2 = n x o1
  int best = 0;
  for (int i = 1; i < 3; ++i)
    if (std::fabs(n[i]) < std::fabs(n[best])) best = i;
  double e[3] = {0, 0, 0};
  e[best] = 1;
  double o1[3] = {n[1] * e[2] - n[2] * e[1], n[2] * e[0] - n[0] * e[2], n[0] * e[1] - n[1] * e[0]};
  double len = std::sqrt(o1[0] * o1[0] + o1[1] * o1[1] + o1[2] * o1[2]);
  for (double& v : o1) v /= len;
  double o2[3] = {n[1] * o1[2] - n[2] * o1[1], n[2] * o1[0] - n[0] * o1[2], n[0] * o1[1] - n[1] * o1[0]};
  double delta2 = theta_range / theta_steps / 2;
  for (int i = 0; i < theta_steps; ++i) {
    double theta = theta_range * ((0.5 + i) / theta_steps);
    double factor = std::cos(theta - delta2) - std::cos(theta + delta2);
    int ringsteps = (int)std::ceil(std::sin(theta) * steps);
    double weight = (2 * M_PI) / ringsteps;
    double ct = std::cos(theta), st = std::sin(theta);
    for (int j = 0; j < ringsteps; ++j) {
      double phi = 2 * M_PI * ((0.5 + j) / ringsteps);
      double x = ct, y = st * std::cos(phi), z = st * std::sin(phi);
      Dir d;
      d.x = n[0] * x + o1[0] * y + o2[0] * z;
      d.y = n[1] * x + o1[1] * y + o2[1] * z;
      d.z = n[2] * x + o1[2] * y + o2[2] * z;
      d.w = weight * factor;
      dirs.push_back(d);
    }
  }
  return dirs;
}

int ensure_dirs(sfsim_atmo_ctx* ctx, int half, int steps) {
  int& key = half ? ctx->half_key : ctx->sphere_key;
  int& count = half ? ctx->half_n : ctx->sphere_n;
  DeviceBuffer& buf = half ? ctx->half_dirs : ctx->sphere_dirs;
  if (key == steps) return 0;
  std::vector<Dir> d = sphere_directions(half, steps);
  TRY(ensure(ctx, buf, d.size() * sizeof(Dir)));
  HIP_TRY(ctx, hipMemcpyAsync(buf.ptr, d.data(), d.size() * sizeof(Dir), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  key = steps;
  count = (int)d.size();
  return 0;
}

// Direction geometry of the sphere quadrature and the slab scratch of the point-scatter kernels.
int prepare_point_scatter(sfsim_atmo_ctx* ctx, const DevParams& P, bool with_mie, WorkCounters* wc) {
  size_t slab = (size_t)P.nS[0] * ctx->sphere_n * P.nS[2] * P.nS[3] * sizeof(float4);
  TRY(ensure(ctx, ctx->geom, direction_geometry_bytes(P, ctx->sphere_n)));
  TRY(ensure(ctx, ctx->slab, slab));
  if (with_mie) TRY(ensure(ctx, ctx->slab_mie, slab));
  HIP_TRY(ctx, launch_direction_geometry(P, (const Dir*)ctx->sphere_dirs.ptr, ctx->sphere_n, ctx->geom.ptr, wc,
                                         ctx->stream));
  return 0;
}

int activate(sfsim_atmo_ctx* ctx) {
  if (!ctx) return SFSIM_ATMO_ERR_ARGUMENT;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return 0;
}

// ------------------------------------------------------------------ timed launches
struct Timer {
  sfsim_atmo_ctx* ctx;
  bool enabled;
  int next = 0;
  int begin(int kind) {
    if (!enabled) return 0;
    size_t need = (size_t)next + 2;
    while (ctx->events.size() < need) {
      hipEvent_t e;
      HIP_TRY(ctx, hipEventCreate(&e));
      ctx->events.push_back(e);
      ctx->event_kind.push_back(0);
    }
    ctx->event_kind[next] = kind;
    HIP_TRY(ctx, hipEventRecord(ctx->events[next], ctx->stream));
    return 0;
  }
  int end() {
    if (!enabled) return 0;
    HIP_TRY(ctx, hipEventRecord(ctx->events[next + 1], ctx->stream));
    next += 2;
    return 0;
  }
  int collect() {
    if (!enabled) return 0;
    for (int k = 0; k < kNumKinds; ++k) {
      ctx->kind_ms[k] = 0;
      ctx->kind_launches[k] = 0;
    }
    for (int i = 0; i < next; i += 2) {
      float ms = 0;
      HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->events[i], ctx->events[i + 1]));
      ctx->kind_ms[ctx->event_kind[i]] += ms;
      ctx->kind_launches[ctx->event_kind[i]] += 1;
    }
    return 0;
  }
};

#define LAUNCH(timer, kind, call)                      \
  do {                                                 \
    TRY((timer).begin(kind));                          \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(ctx, #call, e_); \
    TRY((timer).end());                                \
  } while (0)

// Validation, quadrature directions and buffers of a build: everything that allocates or copies
// from the host, so that the launches that follow can be captured into a graph.
int prepare_pipeline(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  TRY(validate(ctx, p));
  TRY(ensure_scatter_buffers(ctx, p));
  if (p->iterations > 0) {
    TRY(validate_sphere(ctx, p->sphere_steps));
    TRY(validate_half_sphere(ctx, p->ray_steps));
    TRY(ensure_dirs(ctx, 0, p->sphere_steps));
    TRY(ensure_dirs(ctx, 1, p->ray_steps));
    TRY(ensure_surface_geometry(ctx, p));
  }
  const size_t nT = (size_t)count2(p->transmittance_shape), nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  TRY(ensure(ctx, ctx->T, nT * sizeof(float4)));
  for (int i = 0; i < 2; ++i) {
    TRY(ensure(ctx, ctx->dE[i], nE * sizeof(float4)));
    TRY(ensure(ctx, ctx->Eacc[i], nE * sizeof(float4)));
    TRY(ensure(ctx, ctx->dS[i], nS * sizeof(float4)));
    TRY(ensure(ctx, ctx->Sacc[i], nS * sizeof(float4)));
  }
  TRY(ensure(ctx, ctx->R1, nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->M1, nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->dJ, nS * sizeof(float4)));
  if (p->iterations > 0)
    for (int i = 0; i < 2; ++i) TRY(ensure(ctx, ctx->Ssum[i], nS * sizeof(float4)));
  const size_t outc[4] = {nT * 3, nE * 3, nS * 3, nS * 3};
  for (int i = 0; i < 4; ++i) {
    TRY(ensure(ctx, ctx->out[i], outc[i] * sizeof(float)));
    ctx->out_count[i] = outc[i];
  }
  if (p->iterations > 0) {
    const size_t slab = nS / ((size_t)p->scatter_shape[1]) * ctx->sphere_n * sizeof(float4);  // NH x dirs x NL x NA
    TRY(ensure(ctx, ctx->geom, direction_geometry_bytes(make_dev_params(p), ctx->sphere_n)));
    TRY(ensure(ctx, ctx->slab, slab));
    TRY(ensure(ctx, ctx->slab_mie, slab));
  }
  ctx->built = false;
  return 0;
}

// The launches of the whole driver of atmosphere_lut.clj:43-105 on device buffers of the context
// (prepared by prepare_pipeline): no allocation, host copy or memset, so the sequence can be captured.
// capture (nullable): capture[k - 1] receives the sum dS_1 + ... + dS_k of the scattering tables
// (float4 tables, k = 1..min(iterations, 5)), the training targets of the surrogate.
// The per-build launches before the iterations: T, E0 and (with iterations) the quadrature directions'
// geometry, the ray geometry, the re-sampling map, the sun table, then R1/M1.  merged: the independent
// ones in one launch (graph capture; the timed and counted builds launch them one by one).
int enqueue_prologue(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const DevParams& P, WorkCounters* wc,
                     Timer& timer, bool merged) {
  auto f4 = [](DeviceBuffer& b) { return (float4*)b.ptr; };
  const bool dirs = p->iterations > 0;
  if (merged) {
    PrologueJobs J;
    J.T = f4(ctx->T);
    J.E0 = f4(ctx->dE[0]);
    J.sdirs = dirs ? (const Dir*)ctx->sphere_dirs.ptr : nullptr;
    J.sphere_n = ctx->sphere_n;
    J.geom = ctx->geom.ptr;
    J.hdirs = dirs ? (const Dir*)ctx->half_dirs.ptr : nullptr;
    J.half_n = ctx->half_n;
    J.wc = wc;
    LAUNCH(timer, kRayGeometry, launch_prologue(P, J, ctx->stream));
    LAUNCH(timer, kSingleScatter, launch_single_scatter(P, f4(ctx->R1), f4(ctx->M1), wc, ctx->stream, true));
    return 0;
  }
  LAUNCH(timer, kTransmittance, launch_transmittance(P, f4(ctx->T), wc, ctx->stream));
  LAUNCH(timer, kSurfaceRadianceBase, launch_surface_radiance_base(P, f4(ctx->dE[0]), wc, ctx->stream));
  if (dirs) {
    LAUNCH(timer, kDirectionGeometry, launch_direction_geometry(P, (const Dir*)ctx->sphere_dirs.ptr, ctx->sphere_n,
                                                                ctx->geom.ptr, wc, ctx->stream));
    LAUNCH(timer, kDirectionGeometry,
           launch_surface_geometry(P, (const Dir*)ctx->half_dirs.ptr, ctx->half_n, ctx->stream));
  }
  LAUNCH(timer, kRayGeometry, launch_ray_geometry(P, wc, ctx->stream));
  LAUNCH(timer, kAccumulate, launch_resample_map(P, ctx->stream));
  LAUNCH(timer, kSingleScatter, launch_single_scatter(P, f4(ctx->R1), f4(ctx->M1), wc, ctx->stream));
  return 0;
}

void progress_host(void* arg) {
  const ProgressNote* n = (const ProgressNote*)arg;
  n->fn(n->user, n->stage, n->iteration, n->iterations);
}

// Queues a progress report behind the launches enqueued so far (nothing without a callback).
int note_progress(sfsim_atmo_ctx* ctx, int stage, int iteration, int iterations) {
  if (!ctx->progress_fn) return 0;
  ctx->notes.push_back(ProgressNote{ctx->progress_fn, ctx->progress_user, stage, iteration, iterations});
  HIP_TRY(ctx, hipLaunchHostFunc(ctx->stream, progress_host, &ctx->notes.back()));
  return 0;
}

int enqueue_pipeline(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const DevParams& P, WorkCounters* wc,
                     Timer& timer, float4* const* capture, bool merged = false) {
  const size_t nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  (void)nE;
  auto f4 = [](DeviceBuffer& b) { return (float4*)b.ptr; };
  const Dir* sdirs = (const Dir*)ctx->sphere_dirs.ptr;

  TRY(enqueue_prologue(ctx, p, P, wc, timer, merged));
  TRY(note_progress(ctx, SFSIM_ATMO_PROGRESS_FIRST_ORDER, 0, p->iterations));
  // dS <- R1 + M1 * phase (a function); S <- R1; E <- 0; dE <- E0
  // The accumulations E <- table(E + dE), S <- table(S + dS) (atmosphere_lut.clj:95-97) ride in the
  // ray-scatter launches: iteration k writes Ssum[k % 2] = S + dS_k, with S = R1 for the first and
  // S = table(Ssum of the iteration before) after that, and E_k = table(dE_k + E_{k-1}) in an extra
  // workgroup; the last S re-sampling is a launch of its own.
  const float4* S_in = f4(ctx->R1);
  int S_resample = 0;
  const float4* E_cur = nullptr;
  int dE_cur = 0, dS_cur = -1, e_next = 0;
  for (int it = 0; it < p->iterations; ++it) {
    const bool first = it == 0;
    const float4* dS_tab = first ? f4(ctx->R1) : f4(ctx->dS[dS_cur]);
    const float4* mie = first ? f4(ctx->M1) : nullptr;
    // the partials of atmosphere_lut.clj:88-89 bind the previous dS and dE; the surface radiance of
    // dS rides in the point-scatter launch when its LDS slab fits
    const int dE_new = 1 - dE_cur;
    const bool sr_along = sr_fits_point_scatter(P, mie != nullptr);
    const SrJob sr{0, ctx->half_n, dS_tab, mie, f4(ctx->dE[dE_new])};
    LAUNCH(timer, kPointScatter,
           launch_point_scatter(P, sdirs, ctx->sphere_n, ctx->geom.ptr, dS_tab, mie, f4(ctx->dE[dE_cur]), f4(ctx->slab),
                                f4(ctx->slab_mie), f4(ctx->dJ), wc, ctx->stream, sr_along ? &sr : nullptr));
    if (!sr_along)
      LAUNCH(timer, kSurfaceRadiance,
             launch_surface_radiance(P, ctx->half_n, dS_tab, mie, f4(ctx->dE[dE_new]), wc, ctx->stream));
    const int dS_new = dS_cur < 0 ? 0 : 1 - dS_cur;
    RsAccum acc;
    acc.S_in = S_in;
    acc.S_out = f4(ctx->Ssum[it & 1]);
    acc.S_resample = S_resample;
    acc.dE = f4(ctx->dE[dE_new]);
    acc.E_in = E_cur;
    acc.E_out = f4(ctx->Eacc[e_next]);
    LAUNCH(timer, kRayScatter, launch_ray_scatter(P, f4(ctx->dJ), f4(ctx->dS[dS_new]), wc, ctx->stream, acc));
    dE_cur = dE_new;
    dS_cur = dS_new;
    E_cur = f4(ctx->Eacc[e_next]);
    e_next = 1 - e_next;
    S_in = f4(ctx->Ssum[it & 1]);
    S_resample = 1;
    if (capture) {
      const float4* d = f4(ctx->dS[dS_cur]);
      float4* sum = f4(ctx->sumS);
      if (it == 0)
        HIP_TRY(ctx, hipMemcpyAsync(sum, d, nS * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
      else
        HIP_TRY(ctx, launch_add4(sum, d, sum, (long long)nS, ctx->stream));
      const int order = it + 1;
      if (order <= kMlpOrders)
        HIP_TRY(ctx, hipMemcpyAsync(capture[order - 1], sum, nS * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
    }
    TRY(note_progress(ctx, SFSIM_ATMO_PROGRESS_ITERATION, it + 1, p->iterations));
  }
  // the last S re-sampling and the final make-lookup-table calls and file layouts
  // (atmosphere_lut.clj:97-105), the independent ones in one launch: T, E and M1 beside S <- table(S + dS)
  ResampleJob jobs[kMaxResampleJobs];
  int nj = 0;
  jobs[nj++] = ResampleJob{2, f4(ctx->T), nullptr, nullptr, (float*)ctx->out[0].ptr, 0, nullptr};
  jobs[nj++] = ResampleJob{4, f4(ctx->M1), nullptr, nullptr, (float*)ctx->out[3].ptr, 1, nullptr};
  if (E_cur) jobs[nj++] = ResampleJob{3, E_cur, nullptr, nullptr, (float*)ctx->out[1].ptr, 0, nullptr};
  if (S_resample) {  // S_in: the last iteration's S + dS
    jobs[nj++] = ResampleJob{4, S_in, nullptr, f4(ctx->Sacc[0]), nullptr, 0, nullptr};
    LAUNCH(timer, kFinal, launch_resamples(P, jobs, nj, ctx->stream));
    LAUNCH(timer, kFinal, launch_resample(P, 4, f4(ctx->Sacc[0]), nullptr, nullptr, (float*)ctx->out[2].ptr, 1,
                                          ctx->stream));
  } else {  // S = R1
    jobs[nj++] = ResampleJob{4, S_in, nullptr, nullptr, (float*)ctx->out[2].ptr, 1, nullptr};
    LAUNCH(timer, kFinal, launch_resamples(P, jobs, nj, ctx->stream));
  }
  if (!E_cur)  // every atmosphere's surface-radiance table (blocks atm_stride apart in a batch)
    for (int a = 0; a < std::max(P.batch, 1); ++a)
      LAUNCH(timer, kFinal, launch_fill_zero((float*)((char*)ctx->out[1].ptr + (size_t)a * P.atm_stride),
                                             (long long)ctx->out_count[1], ctx->stream));
  TRY(note_progress(ctx, SFSIM_ATMO_PROGRESS_FINAL, 0, p->iterations));
  return 0;
}

// The whole driver, launched eagerly (timed: per-kernel events; wc: work counters).
int run_pipeline(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, WorkCounters* wc, bool timed,
                 float4* const* capture = nullptr) {
  TRY(prepare_pipeline(ctx, p));
  const DevParams P = dev_params(ctx, p);
  Timer timer{ctx, timed};
  if (!ctx->notes.empty()) {  // reports of an earlier build, run by the time its stream is idle
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->notes.clear();
  }
  TRY(enqueue_pipeline(ctx, p, P, wc, timer, capture));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  TRY(timer.collect());
  ctx->built = true;
  return 0;
}

void drop_graph(sfsim_atmo_ctx* ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  ctx->graph_exec = nullptr;
}

// The whole driver as one graph launch: captured on the first build with these parameters and
// buffers, replayed afterwards (the kernels are deterministic, so replays equal eager builds).
int run_pipeline_graph(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  TRY(prepare_pipeline(ctx, p));
  const DevParams P = dev_params(ctx, p);
  const bool hit = ctx->graph_exec && ctx->graph_gen == ctx->alloc_gen &&
                   std::memcmp(&ctx->graph_params, p, sizeof(*p)) == 0 &&
                   std::memcmp(&ctx->graph_dev, &P, sizeof(P)) == 0;
  if (!hit) {
    drop_graph(ctx);
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    Timer timer{ctx, false};
    const int rc = enqueue_pipeline(ctx, p, P, nullptr, timer, nullptr, true);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != 0) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(ctx, e);
    const hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      ctx->graph_exec = nullptr;
      return hip_fail(ctx, "hipGraphInstantiate", ei);
    }
    ctx->graph_params = *p;
    ctx->graph_dev = P;
    ctx->graph_gen = ctx->alloc_gen;
  }
  HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->built = true;
  return 0;
}

// ------------------------------------------------------------------ surrogate-seeded build
AtmosphereCode code_of(const sfsim_atmo_params* p) {
  return atmosphere_code(p->rayleigh.scatter_base, p->rayleigh.scatter_scale, p->mie.scatter_base[0],
                         p->mie.scatter_scale, p->mie.scatter_g, p->planet.brightness[0]);
}

// The 5-order set with the network standing in for part of the iteration.  atmosphere_lut.clj:86-97 is
// the linear recurrence x_k = L x_{k-1} on x_k = (dS_k, dE_k): dS_k = S[J[dS_{k-1}, dE_{k-1}]],
// dE_k = E[dS_{k-1}], x_0 = (dS_0, E0).  Its partial sums X_j = x_0 + ... + x_j obey
// X_{j+1} = x_0 + L X_j, and with P_j = dS_1 + ... + dS_j the sum X_j is
// (dS_0 + P_j, E0 + E[dS_0 + P_{j-1}]).  For N iterations, seeding X_{N-m} with the network's P_{N-m},
// P_{N-m-1} and applying the recurrence m times gives X_N with the network's error multiplied by L^m (a
// contraction); the network predicts P_1..P_5, so m = `applications` is in N-4..N-1 (the seeded order is
// at most 4 and P_5 is computed).  The network's own P_5 against the computed sum is the residual; above
// `tol` the plain iteration runs instead.  The final ray-scatter table then follows the plain
// iteration's accumulation, re-sampled order by order, on P_1..P_j of the seed and the computed
// P_{j+1}..P_N.

// Validation and buffers of a seeded build (everything that allocates or copies from the host but the
// caller's order sums).
int prepare_seeded(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int m, const float* host_sums) {
  TRY(validate(ctx, p));
  const int N = p->iterations;
  if (N < kMlpOrders || N > kMaxSeededOrders)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "a seeded build has 5..32 scattering orders");
  if (m < N - (kMlpOrders - 1) || m > N - 1)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "applications must be in iterations-4 .. iterations-1");
  if (!host_sums && !ctx->sg.trained) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "surrogate is not trained");
  if (host_sums && !finite_floats(host_sums, (size_t)kMlpOrders * count4(p->scatter_shape) * 3))
    return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "order sums hold NaN/Inf");
  TRY(validate_sphere(ctx, p->sphere_steps));
  TRY(validate_half_sphere(ctx, p->ray_steps));
  TRY(ensure_dirs(ctx, 0, p->sphere_steps));
  TRY(ensure_dirs(ctx, 1, p->ray_steps));
  TRY(ensure_surface_geometry(ctx, p));
  TRY(ensure_scatter_buffers(ctx, p));
  const DevParams P = dev_params(ctx, p);
  const size_t nT = (size_t)count2(p->transmittance_shape), nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  TRY(ensure(ctx, ctx->T, nT * sizeof(float4)));
  for (int i = 0; i < 2; ++i) {
    TRY(ensure(ctx, ctx->dE[i], nE * sizeof(float4)));
    TRY(ensure(ctx, ctx->Eacc[i], nE * sizeof(float4)));
    TRY(ensure(ctx, ctx->dS[i], nS * sizeof(float4)));
    TRY(ensure(ctx, ctx->Sacc[i], nS * sizeof(float4)));
  }
  TRY(ensure(ctx, ctx->R1, nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->M1, nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->dJ, nS * sizeof(float4)));
  for (int i = 0; i < N; ++i) TRY(ensure(ctx, ctx->pred[i], nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->resid, residual_scratch_floats((long long)nS) * sizeof(float)));
  if (host_sums) TRY(ensure(ctx, ctx->stage_rgb, nS * 3 * sizeof(float)));
  const size_t outc[4] = {nT * 3, nE * 3, nS * 3, nS * 3};
  for (int i = 0; i < 4; ++i) {
    TRY(ensure(ctx, ctx->out[i], outc[i] * sizeof(float)));
    ctx->out_count[i] = outc[i];
  }
  TRY(ensure(ctx, ctx->geom, direction_geometry_bytes(P, ctx->sphere_n)));
  const size_t slab = nS / (size_t)p->scatter_shape[1] * ctx->sphere_n * sizeof(float4);  // NH x dirs x NL x NA
  TRY(ensure(ctx, ctx->slab, slab));
  TRY(ensure(ctx, ctx->slab_mie, slab));
  if (!host_sums) HIP_TRY(ctx, surrogate_reserve_tables(ctx->sg, P));
  ctx->built = false;
  return 0;
}

// The launches of a seeded build: T, E0, R1/M1, direction geometry, the seed (network prediction, or the
// caller's sums uploaded), m applications, the residual and the final tables.  With host_sums == nullptr
// there is no host copy, allocation or memset, so the sequence can be captured.
int enqueue_seeded(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const DevParams& P, int m, const float* host_sums,
                   Timer& timer, bool merged = false) {
  const size_t nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  auto f4 = [](DeviceBuffer& b) { return (float4*)b.ptr; };
  const Dir* sdirs = (const Dir*)ctx->sphere_dirs.ptr;
  const long long n4 = (long long)nS, e4 = (long long)nE;
  const int N = p->iterations;
  const int j = N - m;  // seeded partial sum X_j, 1 <= j <= 4
  TRY(enqueue_prologue(ctx, p, P, nullptr, timer, merged));
  // the network's order sums: P_1..P_4 into pred[0..3] (the ones below the seed feed the final
  // table's re-sampling chain), its P_5 into dS[1] (residual)
  float4* want[kMlpOrders];
  for (int k = 0; k + 1 < kMlpOrders; ++k) want[k] = f4(ctx->pred[k]);
  want[kMlpOrders - 1] = f4(ctx->dS[1]);
  if (host_sums) {
    for (int k = 0; k < kMlpOrders; ++k) {
      HIP_TRY(ctx, hipMemcpyAsync(ctx->stage_rgb.ptr, host_sums + (size_t)k * nS * 3, nS * 3 * sizeof(float),
                                  hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(ctx, launch_unpack_rgb((const float*)ctx->stage_rgb.ptr, want[k], n4, ctx->stream));
    }
  } else {
    TRY(timer.begin(kSurrogate));
    HIP_TRY(ctx, surrogate_predict_tables(ctx->sg, P, code_of(p), want, ctx->stream));
    TRY(timer.end());
  }
  // X_j: S part dS_0 + P_j carried as A = R1 + P_j (the Mie phase term enters through M1), E part
  // E0 + E[dS_0 + P_{j-1}]
  float4* A = f4(ctx->Sacc[0]);
  float4* B = f4(ctx->Sacc[1]);
  float4* Esum = f4(ctx->Eacc[0]);
  LAUNCH(timer, kAccumulate, launch_add4(f4(ctx->R1), f4(ctx->pred[j - 1]), A, n4, ctx->stream));
  if (j >= 2) {
    LAUNCH(timer, kAccumulate, launch_add4(f4(ctx->R1), f4(ctx->pred[j - 2]), B, n4, ctx->stream));
    LAUNCH(timer, kSurfaceRadiance,
           launch_surface_radiance(P, ctx->half_n, B, f4(ctx->M1), f4(ctx->dE[1]), nullptr, ctx->stream));
  } else {
    LAUNCH(timer, kSurfaceRadiance, launch_surface_radiance(P, ctx->half_n, f4(ctx->R1), f4(ctx->M1), f4(ctx->dE[1]),
                                                            nullptr, ctx->stream));
  }
  LAUNCH(timer, kAccumulate, launch_add4(f4(ctx->dE[0]), f4(ctx->dE[1]), Esum, e4, ctx->stream));
  // m applications of X <- x_0 + L X: application a leaves the computed P_{j+1+a} in pred[j + a]; after
  // the last, dE[1] = dE_1 + ... + dE_5
  for (int a = 0; a < m; ++a) {
    float4* Pk = f4(ctx->pred[j + a]);
    const bool sr_along = sr_fits_point_scatter(P, true);
    const SrJob sr{0, ctx->half_n, A, f4(ctx->M1), f4(ctx->dE[1])};
    LAUNCH(timer, kPointScatter,
           launch_point_scatter(P, sdirs, ctx->sphere_n, ctx->geom.ptr, A, f4(ctx->M1), Esum, f4(ctx->slab),
                                f4(ctx->slab_mie), f4(ctx->dJ), nullptr, ctx->stream, sr_along ? &sr : nullptr));
    LAUNCH(timer, kRayScatter, launch_ray_scatter(P, f4(ctx->dJ), Pk, nullptr, ctx->stream));
    if (!sr_along)
      LAUNCH(timer, kSurfaceRadiance,
             launch_surface_radiance(P, ctx->half_n, A, f4(ctx->M1), f4(ctx->dE[1]), nullptr, ctx->stream));
    if (a + 1 < m) {
      LAUNCH(timer, kAccumulate, launch_add4(f4(ctx->R1), Pk, A, n4, ctx->stream));
      LAUNCH(timer, kAccumulate, launch_add4(f4(ctx->dE[0]), f4(ctx->dE[1]), Esum, e4, ctx->stream));
    }
  }
  // residual of the network's own P_5 against the computed one, relative to R1 + P_5
  TRY(timer.begin(kSurrogate));
  HIP_TRY(ctx, launch_residual(f4(ctx->R1), f4(ctx->pred[kMlpOrders - 1]), f4(ctx->dS[1]), n4, 1e-6f,
                               (float*)ctx->resid.ptr, ctx->stream));
  TRY(timer.end());
  // the plain iteration's accumulation S_k = tab(S_{k-1} + dS_k), S_0 = R1 (atmosphere_lut.clj:95-97), on
  // the orders dS_k = P_k - P_{k-1}: re-sampling at forward(backward(i)) is not the identity where the
  // sun direction is clamped, so the order-by-order chain is what the plain tables hold there.  Each
  // step re-samples S + (P_k - P_{k-1}) in one launch; the last one shares its launch with the final
  // tables of T, E and M1 (atmosphere_lut.clj:98-105)
  const float4* S_cur = f4(ctx->R1);
  int s_next = 0;
  for (int k = 1; k <= N; ++k) {
    ResampleJob jobs[kMaxResampleJobs];
    int nj = 0;
    jobs[nj++] = ResampleJob{4, S_cur, f4(ctx->pred[k - 1]), f4(ctx->Sacc[s_next]), nullptr, 0,
                             k > 1 ? f4(ctx->pred[k - 2]) : nullptr};
    if (k == N) {
      jobs[nj++] = ResampleJob{2, f4(ctx->T), nullptr, nullptr, (float*)ctx->out[0].ptr, 0, nullptr};
      jobs[nj++] = ResampleJob{3, f4(ctx->dE[1]), nullptr, nullptr, (float*)ctx->out[1].ptr, 0, nullptr};
      jobs[nj++] = ResampleJob{4, f4(ctx->M1), nullptr, nullptr, (float*)ctx->out[3].ptr, 1, nullptr};
    }
    LAUNCH(timer, k == N ? kFinal : kAccumulate, launch_resamples(P, jobs, nj, ctx->stream));
    S_cur = f4(ctx->Sacc[s_next]);
    s_next = 1 - s_next;
  }
  LAUNCH(timer, kFinal, launch_resample(P, 4, S_cur, nullptr, nullptr, (float*)ctx->out[2].ptr, 1, ctx->stream));
  return 0;
}

void drop_seeded_graph(sfsim_atmo_ctx* ctx) {
  if (ctx->seeded_exec) (void)hipGraphExecDestroy(ctx->seeded_exec);
  ctx->seeded_exec = nullptr;
}

// host_sums (nullable): the order sums P_1..P_5 given by the caller (5 RGB scatter tables) instead of
// the network's; that build runs kernel by kernel with event timings.  The network-seeded build is
// captured into a graph on its first call with given parameters and m and replayed afterwards (the
// weights are read at replay, so training in between needs no re-capture).
int run_seeded(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int m, double tol, double* residual, int* fallback,
               const float* host_sums = nullptr) {
  TRY(prepare_seeded(ctx, p, m, host_sums));
  const DevParams P = dev_params(ctx, p);
  if (host_sums) {
    Timer timer{ctx, true};
    TRY(enqueue_seeded(ctx, p, P, m, host_sums, timer));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    TRY(timer.collect());
  } else {
    const void* bufs[3] = {ctx->sg.theta, ctx->sg.theta_bf, ctx->sg.ztab};
    const bool hit = ctx->seeded_exec && ctx->seeded_gen == ctx->alloc_gen && ctx->seeded_m == m &&
                     std::memcmp(ctx->seeded_bufs, bufs, sizeof(bufs)) == 0 &&
                     std::memcmp(&ctx->seeded_params, p, sizeof(*p)) == 0 &&
                     std::memcmp(&ctx->seeded_dev, &P, sizeof(P)) == 0;
    if (!hit) {
      drop_seeded_graph(ctx);
      HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      Timer timer{ctx, false};
      const int rc = enqueue_seeded(ctx, p, P, m, nullptr, timer, true);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != 0) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      HIP_TRY(ctx, e);
      const hipError_t ei = hipGraphInstantiate(&ctx->seeded_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        ctx->seeded_exec = nullptr;
        return hip_fail(ctx, "hipGraphInstantiate", ei);
      }
      ctx->seeded_gen = ctx->alloc_gen;
      ctx->seeded_m = m;
      std::memcpy(ctx->seeded_bufs, bufs, sizeof(bufs));
      ctx->seeded_params = *p;
      ctx->seeded_dev = P;
    }
    HIP_TRY(ctx, hipGraphLaunch(ctx->seeded_exec, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  float rr[8];
  HIP_TRY(ctx, hipMemcpy(rr, ctx->resid.ptr, sizeof(rr), hipMemcpyDeviceToHost));
  const float r = rr[4];
  *residual = r;
  ctx->last_residual = r;
  if (!(r <= tol)) {
    *fallback = 1;
    return run_pipeline_graph(ctx, p);
  }
  *fallback = 0;
  ctx->built = true;
  return 0;
}

// ------------------------------------------------------------------ stage entry helpers

// Upload a host RGB table into float4 staging buffer `slot`.
int upload_table(sfsim_atmo_ctx* ctx, int slot, const float* host, size_t entries, const float4** dev) {
  if (!host) {
    *dev = nullptr;
    return 0;
  }
  if (!finite_floats(host, entries * 3)) return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "input table holds NaN/Inf");
  TRY(ensure(ctx, ctx->stage_rgb, entries * 3 * sizeof(float)));
  TRY(ensure(ctx, ctx->stage_in[slot], entries * sizeof(float4)));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->stage_rgb.ptr, host, entries * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, launch_unpack_rgb((const float*)ctx->stage_rgb.ptr, (float4*)ctx->stage_in[slot].ptr, (long long)entries,
                                 ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *dev = (const float4*)ctx->stage_in[slot].ptr;
  return 0;
}

int download_table(sfsim_atmo_ctx* ctx, const float4* dev, float* host, size_t entries) {
  if (!host) return 0;
  TRY(ensure(ctx, ctx->stage_rgb, entries * 3 * sizeof(float)));
  HIP_TRY(ctx, launch_pack_rgb(dev, (float*)ctx->stage_rgb.ptr, (long long)entries, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(host, ctx->stage_rgb.ptr, entries * 3 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

float4* stage_out(sfsim_atmo_ctx* ctx, size_t entries, int* rc) {
  *rc = ensure(ctx, ctx->stage_out4, entries * sizeof(float4));
  return (float4*)ctx->stage_out4.ptr;
}

// Carves buffers out of device memory at `base` (256-byte aligned offsets); with base == nullptr only the
// size is counted.
struct Carver {
  char* base;
  size_t off = 0;
  void take(DeviceBuffer& b, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (base) {
      b.ptr = base + off;
      b.bytes = bytes;
    }
    off += (bytes + 255) & ~(size_t)255;
  }
};

// The buffers of a build that depend only on the planet and the grid (one set for all the atmospheres
// of a batch): ray order, re-sampling and sun maps, quadrature directions, half-sphere geometry.
size_t layout_shared(sfsim_atmo_ctx* c, const sfsim_atmo_params* p, char* base) {
  Carver k{base};
  const DevParams P = make_dev_params(p);
  const size_t nS = (size_t)count4(p->scatter_shape);
  k.take(c->ray_order, (size_t)p->scatter_shape[0] * p->scatter_shape[1] * sizeof(int));
  k.take(c->rmap, resample_map_bytes(P));
  k.take(c->sunmap, (2 * sizeof(float4) + sizeof(double2)) * nS);
  if (p->iterations > 0) {
    const int sn = sfsim_atmo_sphere_directions(0, p->sphere_steps), hn = sfsim_atmo_sphere_directions(1, p->ray_steps);
    k.take(c->sphere_dirs, (size_t)sn * sizeof(Dir));
    k.take(c->half_dirs, (size_t)hn * sizeof(Dir));
    k.take(c->sgeo, (size_t)p->surface_radiance_shape[0] * hn * sizeof(float4));
  }
  return k.off;
}

// The buffers of one atmosphere's build that prepare_pipeline asks for (with_outputs: also the four packed
// final tables); the same layout for every atmosphere of a batch, one block each.
size_t layout_atmosphere(sfsim_atmo_ctx* c, const sfsim_atmo_params* p, char* base, bool with_outputs) {
  Carver k{base};
  const DevParams P = make_dev_params(p);
  const size_t nT = (size_t)count2(p->transmittance_shape), nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  if (c->sun_rows > 0) k.take(c->sun_depth, (size_t)c->sun_rows * c->sun_cols * sizeof(float2));
  k.take(c->ray_geo, ray_geometry_bytes(P));
  if (p->iterations > 0) {
    const int sn = sfsim_atmo_sphere_directions(0, p->sphere_steps);
    k.take(c->geom, direction_geometry_bytes(P, sn));
    const size_t slab = nS / (size_t)p->scatter_shape[1] * sn * sizeof(float4);
    k.take(c->slab, slab);
    k.take(c->slab_mie, slab);
    for (int i = 0; i < 2; ++i) k.take(c->Ssum[i], nS * sizeof(float4));
  }
  k.take(c->T, nT * sizeof(float4));
  for (int i = 0; i < 2; ++i) {
    k.take(c->dE[i], nE * sizeof(float4));
    k.take(c->Eacc[i], nE * sizeof(float4));
    k.take(c->dS[i], nS * sizeof(float4));
    k.take(c->Sacc[i], nS * sizeof(float4));
  }
  k.take(c->R1, nS * sizeof(float4));
  k.take(c->M1, nS * sizeof(float4));
  k.take(c->dJ, nS * sizeof(float4));
  if (with_outputs) {
    const size_t outc[4] = {nT * 3, nE * 3, nS * 3, nS * 3};
    for (int i = 0; i < 4; ++i) k.take(c->out[i], outc[i] * sizeof(float));
  }
  return k.off;
}

// Every buffer that prepare_pipeline asks for (the outputs aside), in one piece of device memory at
// `base`; returns the bytes they take (base == nullptr: the size only).
size_t noalloc_layout(sfsim_atmo_ctx* c, const sfsim_atmo_params* p, char* base) {
  const size_t shared = layout_shared(c, p, base);
  return shared + layout_atmosphere(c, p, base ? base + shared : nullptr, false);
}

}  // namespace

extern "C" {

int sfsim_atmo_create(int device, sfsim_atmo_ctx** out) {
  if (!out) return SFSIM_ATMO_ERR_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SFSIM_ATMO_ERR_DEVICE;
  if (device < 0 || device >= n) return SFSIM_ATMO_ERR_ARGUMENT;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SFSIM_ATMO_ERR_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SFSIM_ATMO_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return SFSIM_ATMO_ERR_DEVICE;
  sfsim_atmo_ctx* ctx = new (std::nothrow) sfsim_atmo_ctx();
  if (!ctx) return SFSIM_ATMO_ERR_ALLOC;
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return SFSIM_ATMO_ERR_DEVICE;
  }
  // the scratch array's address on this device (the device is current); without it only
  // sfsim_atmo_generate_luts_noalloc fails, with the reason
  const hipError_t se = scratch_address(device, &ctx->scratch);
  if (se != hipSuccess) {
    ctx->scratch = nullptr;
    ctx->error = std::string("scratch array unavailable: ") + hipGetErrorString(se);
    (void)hipGetLastError();
  }
  *out = ctx;
  return 0;
}

int sfsim_atmo_destroy(sfsim_atmo_ctx* ctx) {
  if (!ctx) return SFSIM_ATMO_ERR_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DeviceBuffer* bufs[] = {&ctx->T,      &ctx->dE[0],   &ctx->dE[1],       &ctx->Eacc[0],     &ctx->Eacc[1],
                          &ctx->R1,     &ctx->M1,      &ctx->dJ,          &ctx->dS[0],       &ctx->dS[1],
                          &ctx->Sacc[0], &ctx->Sacc[1], &ctx->out[0],     &ctx->out[1],      &ctx->out[2],
                          &ctx->out[3], &ctx->stage_in[0], &ctx->stage_in[1], &ctx->stage_in[2], &ctx->stage_out4,
                          &ctx->stage_rgb, &ctx->sphere_dirs, &ctx->half_dirs, &ctx->counters,
                          &ctx->geom,   &ctx->slab,    &ctx->slab_mie};
  for (DeviceBuffer* b : bufs) release(*b);
  DeviceBuffer* more[] = {&ctx->codes,   &ctx->targets, &ctx->sumS,    &ctx->cap[0],  &ctx->cap[1],
                          &ctx->cap[2],  &ctx->cap[3],  &ctx->cap[4],  &ctx->resid,  &ctx->sun_depth, &ctx->ray_geo, &ctx->rmap, &ctx->sunmap, &ctx->sgeo, &ctx->ray_order, &ctx->Ssum[0], &ctx->Ssum[1],
                          &ctx->batch_mem, &ctx->batch_params};
  for (DeviceBuffer* b : more) release(*b);
  for (DeviceBuffer& b : ctx->pred) release(b);
  surrogate_free(ctx->sg);
  drop_graph(ctx);
  drop_seeded_graph(ctx);
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char* sfsim_atmo_last_error(const sfsim_atmo_ctx* ctx) {
  if (!ctx) return "context is NULL";
  return ctx->error.c_str();
}

int sfsim_atmo_set_sun_table(sfsim_atmo_ctx* ctx, int rows, int cols) {
  if (!ctx) return SFSIM_ATMO_ERR_ARGUMENT;
  const bool per_segment = rows == 0 && cols == 0;
  if (!per_segment && (rows < 16 || cols < 16 || rows > 8192 || cols > 8192))
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sun table rows and cols must both be 0 or both in [16, 8192]");
  ctx->sun_rows = rows;
  ctx->sun_cols = cols;
  return 0;
}

int sfsim_atmo_get_sun_table(const sfsim_atmo_ctx* ctx, int* rows, int* cols) {
  if (!ctx || !rows || !cols) return SFSIM_ATMO_ERR_ARGUMENT;
  *rows = ctx->sun_rows;
  *cols = ctx->sun_cols;
  return 0;
}

int sfsim_atmo_transmittance_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (!out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "out is NULL");
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count2(p->transmittance_shape);
  int rc;
  float4* o = stage_out(ctx, n, &rc);
  TRY(rc);
  HIP_TRY(ctx, launch_transmittance(P, o, nullptr, ctx->stream));
  return download_table(ctx, o, out, n);
}

int sfsim_atmo_surface_radiance_base_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (!out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "out is NULL");
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count2(p->surface_radiance_shape);
  int rc;
  float4* o = stage_out(ctx, n, &rc);
  TRY(rc);
  HIP_TRY(ctx, launch_surface_radiance_base(P, o, nullptr, ctx->stream));
  return download_table(ctx, o, out, n);
}

int sfsim_atmo_single_scatter_tables(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* rayleigh_out,
                                     float* mie_out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  TRY(ensure_scatter_buffers(ctx, p));
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count4(p->scatter_shape);
  TRY(ensure(ctx, ctx->R1, n * sizeof(float4)));
  TRY(ensure(ctx, ctx->M1, n * sizeof(float4)));
  ctx->built = false;
  HIP_TRY(ctx, launch_ray_geometry(P, nullptr, ctx->stream));
  HIP_TRY(ctx, launch_resample_map(P, ctx->stream));  // with it the per-entry sun terms
  HIP_TRY(ctx, launch_single_scatter(P, (float4*)ctx->R1.ptr, (float4*)ctx->M1.ptr, nullptr, ctx->stream));
  TRY(download_table(ctx, (const float4*)ctx->R1.ptr, rayleigh_out, n));
  return download_table(ctx, (const float4*)ctx->M1.ptr, mie_out, n);
}

int sfsim_atmo_point_scatter_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dS,
                                   const float* mie_strength, const float* dE, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  TRY(validate_sphere(ctx, p->sphere_steps));
  if (!dS || !dE || !out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "dS, dE and out must not be NULL");
  TRY(ensure_dirs(ctx, 0, p->sphere_steps));
  TRY(ensure_scatter_buffers(ctx, p));
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count4(p->scatter_shape), ne = (size_t)count2(p->surface_radiance_shape);
  const float4 *ddS, *dM, *ddE;
  TRY(upload_table(ctx, 0, dS, n, &ddS));
  TRY(upload_table(ctx, 1, mie_strength, n, &dM));
  TRY(upload_table(ctx, 2, dE, ne, &ddE));
  int rc;
  float4* o = stage_out(ctx, n, &rc);
  TRY(rc);
  TRY(prepare_point_scatter(ctx, P, dM != nullptr, nullptr));
  HIP_TRY(ctx, launch_resample_map(P, ctx->stream));  // with it the per-entry sun terms
  HIP_TRY(ctx, launch_point_scatter(P, (const Dir*)ctx->sphere_dirs.ptr, ctx->sphere_n, ctx->geom.ptr, ddS, dM, ddE,
                                    (float4*)ctx->slab.ptr, (float4*)ctx->slab_mie.ptr, o, nullptr, ctx->stream));
  return download_table(ctx, o, out, n);
}

int sfsim_atmo_ray_scatter_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dJ, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (!dJ || !out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "dJ and out must not be NULL");
  TRY(ensure_scatter_buffers(ctx, p));
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count4(p->scatter_shape);
  const float4* ddJ;
  TRY(upload_table(ctx, 0, dJ, n, &ddJ));
  int rc;
  float4* o = stage_out(ctx, n, &rc);
  TRY(rc);
  HIP_TRY(ctx, launch_ray_geometry(P, nullptr, ctx->stream));
  HIP_TRY(ctx, launch_resample_map(P, ctx->stream));  // with it the per-entry sun terms
  HIP_TRY(ctx, launch_ray_scatter(P, ddJ, o, nullptr, ctx->stream));
  return download_table(ctx, o, out, n);
}

int sfsim_atmo_surface_radiance_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, const float* dS,
                                      const float* mie_strength, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  TRY(validate_half_sphere(ctx, p->ray_steps));
  if (!dS || !out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "dS and out must not be NULL");
  TRY(ensure_dirs(ctx, 1, p->ray_steps));
  TRY(ensure_surface_geometry(ctx, p));
  DevParams P = dev_params(ctx, p);
  size_t n = (size_t)count4(p->scatter_shape), ne = (size_t)count2(p->surface_radiance_shape);
  const float4 *ddS, *dM;
  TRY(upload_table(ctx, 0, dS, n, &ddS));
  TRY(upload_table(ctx, 1, mie_strength, n, &dM));
  int rc;
  float4* o = stage_out(ctx, ne, &rc);
  TRY(rc);
  HIP_TRY(ctx, launch_surface_geometry(P, (const Dir*)ctx->half_dirs.ptr, ctx->half_n, ctx->stream));
  HIP_TRY(ctx, launch_surface_radiance(P, ctx->half_n, ddS, dM, o, nullptr, ctx->stream));
  return download_table(ctx, o, out, ne);
}

int sfsim_atmo_resample_table(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int space, const float* a,
                              const float* b, float* out) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (space != 2 && space != 3 && space != 4) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "space must be 2, 3 or 4");
  if (!a || !out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "a and out must not be NULL");
  DevParams P = dev_params(ctx, p);
  P.rmap = nullptr;  // a single re-sampling computes its lookups itself
  size_t n = space == 4   ? (size_t)count4(p->scatter_shape)
             : space == 2 ? (size_t)count2(p->transmittance_shape)
                          : (size_t)count2(p->surface_radiance_shape);
  const float4 *da, *db;
  TRY(upload_table(ctx, 0, a, n, &da));
  TRY(upload_table(ctx, 1, b, n, &db));
  int rc;
  float4* o = stage_out(ctx, n, &rc);
  TRY(rc);
  HIP_TRY(ctx, launch_resample(P, space, da, db, o, nullptr, 0, ctx->stream));
  return download_table(ctx, o, out, n);
}

int sfsim_atmo_build(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  TRY(activate(ctx));
  if (ctx->progress_fn) return run_pipeline(ctx, p, nullptr, false);  // host reports between the stages
  return run_pipeline_graph(ctx, p);
}

int sfsim_atmo_set_progress(sfsim_atmo_ctx* ctx, sfsim_atmo_progress_fn fn, void* user) {
  if (!ctx) return SFSIM_ATMO_ERR_ARGUMENT;
  ctx->progress_fn = fn;
  ctx->progress_user = fn ? user : nullptr;
  return 0;
}

int sfsim_atmo_build_timed(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p) {
  TRY(activate(ctx));
  return run_pipeline(ctx, p, nullptr, true);
}

int sfsim_atmo_fetch(sfsim_atmo_ctx* ctx, int which, float* out, size_t count) {
  TRY(activate(ctx));
  if (which < 0 || which > 3 || !out) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "which must be 0..3 and out non-NULL");
  if (!ctx->built) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "no completed build on this context");
  if (count != ctx->out_count[which]) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "count does not match the table size");
  HIP_TRY(ctx, hipMemcpyAsync(out, ctx->out[which].ptr, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int sfsim_atmo_generate_luts(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* transmittance,
                             float* surface_radiance, float* ray_scatter, float* mie_strength) {
  TRY(activate(ctx));
  TRY(run_pipeline(ctx, p, nullptr, false));
  float* outs[4] = {transmittance, surface_radiance, ray_scatter, mie_strength};
  for (int i = 0; i < 4; ++i)
    if (outs[i]) TRY(sfsim_atmo_fetch(ctx, i, outs[i], ctx->out_count[i]));
  return 0;
}

int sfsim_atmo_generate_luts_noalloc(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* transmittance,
                                     float* surface_radiance, float* ray_scatter, float* mie_strength) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (!transmittance || !surface_radiance || !ray_scatter || !mie_strength)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "the four output device buffers must not be NULL");
  if (p->iterations > 0) {
    TRY(validate_sphere(ctx, p->sphere_steps));
    TRY(validate_half_sphere(ctx, p->ray_steps));
  }
  if (!ctx->scratch)
    return fail(ctx, SFSIM_ATMO_ERR_DEVICE, ctx->error.empty() ? "scratch array unavailable" : ctx->error);
  // a context on the stack whose buffers all lie in the scratch array: the driver's ensure() calls find
  // them in place (and fail instead of allocating if one were missing)
  sfsim_atmo_ctx sh;
  sh.device = ctx->device;
  sh.stream = ctx->stream;
  sh.no_alloc = true;
  sh.sun_rows = ctx->sun_rows;
  sh.sun_cols = ctx->sun_cols;
  const size_t need = noalloc_layout(&sh, p, nullptr);
  if (need > SFSIM_ATMO_NOALLOC_SCRATCH_BYTES) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "the workload needs %zu bytes of scratch, more than the %zu of the no-alloc entry point",
                  need, (size_t)SFSIM_ATMO_NOALLOC_SCRATCH_BYTES);
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, msg);
  }
  noalloc_layout(&sh, p, (char*)ctx->scratch);
  const size_t nT = (size_t)count2(p->transmittance_shape), nE = (size_t)count2(p->surface_radiance_shape);
  const size_t nS = (size_t)count4(p->scatter_shape);
  const size_t outc[4] = {nT * 3, nE * 3, nS * 3, nS * 3};
  float* outs[4] = {transmittance, surface_radiance, ray_scatter, mie_strength};
  for (int i = 0; i < 4; ++i) {
    sh.out[i].ptr = outs[i];
    sh.out[i].bytes = outc[i] * sizeof(float);
  }
  const int rc = run_pipeline(&sh, p, nullptr, false);
  if (rc) ctx->error = sh.error;
  return rc;
}

int sfsim_atmo_build_batch(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* ps, int count) {
  TRY(activate(ctx));
  if (!ps || count < 1 || count > 65535) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "count must be in [1, 65535]");
  const sfsim_atmo_params* p0 = &ps[0];
  for (int a = 0; a < count; ++a) {
    const sfsim_atmo_params* p = &ps[a];
    TRY(validate(ctx, p));
    const bool same =
        p->planet.radius == p0->planet.radius && p->planet.height == p0->planet.height &&
        std::memcmp(p->transmittance_shape, p0->transmittance_shape, sizeof(p->transmittance_shape)) == 0 &&
        std::memcmp(p->surface_radiance_shape, p0->surface_radiance_shape, sizeof(p->surface_radiance_shape)) == 0 &&
        std::memcmp(p->scatter_shape, p0->scatter_shape, sizeof(p->scatter_shape)) == 0 &&
        p->ray_steps == p0->ray_steps && p->sphere_steps == p0->sphere_steps && p->iterations == p0->iterations;
    if (!same)
      return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT,
                  "the atmospheres of a batch must share the planet's radius and height, the table shapes, the step "
                  "counts and the iterations");
  }
  if (ctx->sun_rows <= 0)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "batched builds read the sun optical depth from its table (rows > 0)");
  if (p0->iterations > 0) {
    TRY(validate_sphere(ctx, p0->sphere_steps));
    TRY(validate_half_sphere(ctx, p0->ray_steps));
  }
  ctx->batch_built = false;
  // the shared geometry, then one block of buffers per atmosphere
  sfsim_atmo_ctx sh;
  sh.device = ctx->device;
  sh.stream = ctx->stream;
  sh.no_alloc = true;
  sh.sun_rows = ctx->sun_rows;
  sh.sun_cols = ctx->sun_cols;
  sh.progress_fn = ctx->progress_fn;
  sh.progress_user = ctx->progress_user;
  const size_t shared = layout_shared(&sh, p0, nullptr), block = layout_atmosphere(&sh, p0, nullptr, true);
  TRY(ensure(ctx, ctx->batch_mem, shared + (size_t)count * block));
  TRY(ensure(ctx, ctx->batch_params, (size_t)count * sizeof(DevParams)));
  char* base = (char*)ctx->batch_mem.ptr;
  layout_shared(&sh, p0, base);
  layout_atmosphere(&sh, p0, base + shared, true);
  auto fail_sh = [&](int rc) {
    ctx->error = sh.error;
    return rc;
  };
  int rc = prepare_pipeline(&sh, p0);
  if (rc) return fail_sh(rc);
  // per-atmosphere parameters: each atmosphere's own ray geometry and sun table, the shared maps
  const DevParams* dparams = (const DevParams*)ctx->batch_params.ptr;
  std::vector<DevParams> host(count);
  for (int a = 0; a < count; ++a) {
    DevParams Pa = dev_params(&sh, &ps[a]);
    Pa.ray_geo = (char*)sh.ray_geo.ptr + (size_t)a * block;
    Pa.sun_depth = (float2*)((char*)sh.sun_depth.ptr + (size_t)a * block);
    Pa.batch = count;
    Pa.atm_stride = block;
    Pa.batch_params = dparams;
    host[a] = Pa;
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->batch_params.ptr, host.data(), (size_t)count * sizeof(DevParams),
                              hipMemcpyHostToDevice, ctx->stream));
  Timer timer{&sh, false};
  rc = enqueue_pipeline(&sh, p0, host[0], nullptr, timer, nullptr, true);
  // the progress reports queued so far point into sh: wait for them on the failure path as well
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (rc) return fail_sh(rc);
  HIP_TRY(ctx, es);
  ctx->batch_count = count;
  ctx->batch_stride = block;
  for (int i = 0; i < 4; ++i) {
    ctx->batch_out[i] = sh.out[i].ptr;
    ctx->batch_out_count[i] = sh.out_count[i];
  }
  ctx->batch_built = true;
  return 0;
}

int sfsim_atmo_fetch_batch(sfsim_atmo_ctx* ctx, int atmosphere, int which, float* out, size_t count) {
  TRY(activate(ctx));
  if (!ctx->batch_built) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "no completed batched build on this context");
  if (atmosphere < 0 || atmosphere >= ctx->batch_count || which < 0 || which > 3 || !out)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "atmosphere must be in the batch, which 0..3 and out non-NULL");
  if (count != ctx->batch_out_count[which])
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "count does not match the table size");
  const char* src = (const char*)ctx->batch_out[which] + (size_t)atmosphere * ctx->batch_stride;
  HIP_TRY(ctx, hipMemcpyAsync(out, src, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int sfsim_atmo_stage_timings(const sfsim_atmo_ctx* ctx, double* ms, int* launches, int n) {
  if (!ctx) return SFSIM_ATMO_ERR_ARGUMENT;
  for (int k = 0; k < n && k < kNumKinds; ++k) {
    if (ms) ms[k] = ctx->kind_ms[k];
    if (launches) launches[k] = ctx->kind_launches[k];
  }
  return kNumKinds;
}

const char* sfsim_atmo_stage_name(int kind) {
  if (kind < 0 || kind >= kNumKinds) return "";
  return kKindNames[kind];
}

int sfsim_atmo_count_work(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, unsigned long long* counts) {
  TRY(activate(ctx));
  if (!counts) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "counts is NULL");
  TRY(ensure(ctx, ctx->counters, sizeof(WorkCounters)));
  HIP_TRY(ctx, hipMemsetAsync(ctx->counters.ptr, 0, sizeof(WorkCounters), ctx->stream));
  TRY(run_pipeline(ctx, p, (WorkCounters*)ctx->counters.ptr, false));
  WorkCounters wc;
  HIP_TRY(ctx, hipMemcpy(&wc, ctx->counters.ptr, sizeof(wc), hipMemcpyDeviceToHost));
  counts[0] = wc.single_samples;
  counts[1] = wc.samples;
  counts[2] = wc.lookups4;
  counts[3] = wc.lookups2;
  return 0;
}

// ------------------------------------------------------------------ surrogate entry points
size_t sfsim_atmo_surrogate_size(void) { return MlpLayout::total; }

int sfsim_atmo_surrogate_reset(sfsim_atmo_ctx* ctx, unsigned long long seed, int batch) {
  TRY(activate(ctx));
  if (batch < 64 || batch > (1 << 20) || batch % 64 != 0)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "batch must be a multiple of 64 in [64, 1048576]");
  HIP_TRY(ctx, surrogate_alloc(ctx->sg, batch));
  HIP_TRY(ctx, surrogate_init(ctx->sg, seed, ctx->stream));
  return 0;
}

int sfsim_atmo_surrogate_weights(sfsim_atmo_ctx* ctx, float* out, size_t count) {
  TRY(activate(ctx));
  if (!ctx->sg.theta) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "surrogate not initialised");
  if (!out || count != MlpLayout::total) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "count must be the parameter count");
  HIP_TRY(ctx, hipMemcpyAsync(out, ctx->sg.theta, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int sfsim_atmo_surrogate_set_weights(sfsim_atmo_ctx* ctx, const float* in, size_t count, int trained) {
  TRY(activate(ctx));
  if (!ctx->sg.theta) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "surrogate not initialised");
  if (!in || count != MlpLayout::total) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "count must be the parameter count");
  if (!finite_floats(in, count)) return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "weights hold NaN/Inf");
  HIP_TRY(ctx, hipMemcpyAsync(ctx->sg.theta, in, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, surrogate_sync_bf16(ctx->sg, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->sg.trained = trained != 0;
  return 0;
}

int sfsim_atmo_surrogate_eval(sfsim_atmo_ctx* ctx, const float* inputs, const float* targets, int rows,
                              float* outputs, float* grad, float* loss) {
  TRY(activate(ctx));
  Surrogate& sg = ctx->sg;
  if (!sg.theta) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "surrogate not initialised");
  if (!inputs || rows < 1 || rows > sg.batch_capacity)
    return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "rows must be in [1, batch capacity] and inputs non-NULL");
  if (!finite_floats(inputs, (size_t)rows * kMlpIn)) return fail(ctx, SFSIM_ATMO_ERR_NONFINITE, "inputs hold NaN/Inf");
  TRY(ensure(ctx, ctx->stage_rgb, (size_t)rows * kMlpIn * sizeof(float)));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->stage_rgb.ptr, inputs, (size_t)rows * kMlpIn * sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, surrogate_encode_rows(sg, (const float*)ctx->stage_rgb.ptr, rows, ctx->stream));
  HIP_TRY(ctx, surrogate_forward(sg, rows, ctx->stream));
  if (outputs)
    HIP_TRY(ctx, hipMemcpyAsync(outputs, sg.Y, (size_t)rows * kMlpOut * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
  if (targets) {
    HIP_TRY(ctx, hipMemcpyAsync(sg.T, targets, (size_t)rows * kMlpOut * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, surrogate_backward(sg, rows, ctx->stream));
    if (grad)
      HIP_TRY(ctx, hipMemcpyAsync(grad, sg.grad, MlpLayout::total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (loss) {
      int parts = (rows * kMlpOut + 255) / 256;
      std::vector<float> h(parts);
      HIP_TRY(ctx, hipMemcpyAsync(h.data(), sg.loss, parts * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      double t = 0;
      for (float v : h) t += v;
      *loss = (float)t;
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int sfsim_atmo_surrogate_train(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* atmospheres, int count, int steps,
                               double lr, unsigned long long seed, float* losses) {
  TRY(activate(ctx));
  if (!atmospheres || count < 1 || steps < 0) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "need >= 1 atmosphere");
  if (!ctx->sg.theta) TRY(sfsim_atmo_surrogate_reset(ctx, seed, 8192));
  const sfsim_atmo_params& p0 = atmospheres[0];
  for (int a = 0; a < count; ++a) {
    TRY(validate(ctx, &atmospheres[a]));
    if (atmospheres[a].iterations < kMlpOrders)
      return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "training needs builds of at least 5 orders");
    if (std::memcmp(atmospheres[a].scatter_shape, p0.scatter_shape, sizeof(p0.scatter_shape)) != 0)
      return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "all training atmospheres need the same scatter shape");
  }
  const size_t nS = (size_t)count4(p0.scatter_shape);
  TRY(ensure(ctx, ctx->sumS, nS * sizeof(float4)));
  for (int i = 0; i < kMlpOrders; ++i) TRY(ensure(ctx, ctx->cap[i], nS * sizeof(float4)));
  TRY(ensure(ctx, ctx->targets, (size_t)count * nS * kMlpUsed * sizeof(float)));
  TRY(ensure(ctx, ctx->codes, (size_t)count * sizeof(AtmosphereCode)));
  std::vector<AtmosphereCode> codes(count);
  // training data: the plain GPU iteration of every atmosphere (order sums P_1..P_5)
  float4* caps[kMlpOrders];
  for (int k = 0; k < kMlpOrders; ++k) caps[k] = (float4*)ctx->cap[k].ptr;
  for (int a = 0; a < count; ++a) {
    sfsim_atmo_params five = atmospheres[a];  // P_1..P_5 are the first five orders of any longer build
    five.iterations = kMlpOrders;
    TRY(run_pipeline(ctx, &five, nullptr, false, caps));
    HIP_TRY(ctx, surrogate_make_targets(caps, (long long)nS, (float*)ctx->targets.ptr + (size_t)a * nS * kMlpUsed,
                                        ctx->stream));
    codes[a] = code_of(&atmospheres[a]);
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->codes.ptr, codes.data(), count * sizeof(AtmosphereCode), hipMemcpyHostToDevice,
                              ctx->stream));
  const DevParams P = dev_params(ctx, &p0);
  Surrogate& sg = ctx->sg;
  // a fresh network (no Adam step taken yet) starts from the targets' means
  if (sg.adam_step == 0)
    HIP_TRY(ctx, surrogate_init_output_bias(sg, (const float*)ctx->targets.ptr, (long long)count * (long long)nS,
                                            ctx->stream));
  const int rows = sg.batch_capacity;
  const int parts = (rows * kMlpOut + 255) / 256;
  std::vector<float> part(parts);
  const int warmup = steps / 20;  // linear warm-up over 5% of the steps, then cosine decay
  for (int step = 0; step < steps; ++step) {
    float lr_t = step < warmup ? (float)(lr * (step + 1) / warmup)
                               : (float)(lr * 0.5 * (1.0 + std::cos(M_PI * (step - warmup) / (steps - warmup))));
    HIP_TRY(ctx, surrogate_gather_batch(sg, P, (const AtmosphereCode*)ctx->codes.ptr, count,
                                        (const float*)ctx->targets.ptr, seed, step, rows, ctx->stream));
    HIP_TRY(ctx, surrogate_forward(sg, rows, ctx->stream));
    HIP_TRY(ctx, surrogate_backward(sg, rows, ctx->stream));
    HIP_TRY(ctx, surrogate_adam(sg, lr_t, ctx->stream));
    if (losses) {
      HIP_TRY(ctx, hipMemcpyAsync(part.data(), sg.loss, parts * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      double t = 0;
      for (float v : part) t += v;
      losses[step] = (float)t;
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  sg.trained = true;
  ctx->built = false;
  return 0;
}

int sfsim_atmo_build_seeded(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int applications, double tolerance,
                            double* residual, int* fallback) {
  TRY(activate(ctx));
  double r = 0;
  int f = 0;
  TRY(run_seeded(ctx, p, applications, tolerance, &r, &f));
  if (residual) *residual = r;
  if (fallback) *fallback = f;
  return 0;
}

int sfsim_atmo_build_from_sums(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, int applications, const float* sums,
                               double tolerance, double* residual, int* fallback) {
  TRY(activate(ctx));
  if (!sums) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sums is NULL");
  double r = 0;
  int f = 0;
  TRY(run_seeded(ctx, p, applications, tolerance, &r, &f, sums));
  if (residual) *residual = r;
  if (fallback) *fallback = f;
  return 0;
}

int sfsim_atmo_order_sums(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* sums) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (p->iterations < kMlpOrders) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "order sums need a build of at least 5 orders");
  if (!sums) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sums is NULL");
  const size_t nS = (size_t)count4(p->scatter_shape);
  TRY(ensure(ctx, ctx->sumS, nS * sizeof(float4)));
  float4* caps[kMlpOrders];
  for (int k = 0; k < kMlpOrders; ++k) {
    TRY(ensure(ctx, ctx->cap[k], nS * sizeof(float4)));
    caps[k] = (float4*)ctx->cap[k].ptr;
  }
  sfsim_atmo_params five = *p;
  five.iterations = kMlpOrders;
  TRY(run_pipeline(ctx, &five, nullptr, false, caps));
  for (int k = 0; k < kMlpOrders; ++k) TRY(download_table(ctx, caps[k], sums + (size_t)k * nS * 3, nS));
  return 0;
}

int sfsim_atmo_surrogate_predict(sfsim_atmo_ctx* ctx, const sfsim_atmo_params* p, float* sums) {
  TRY(activate(ctx));
  TRY(validate(ctx, p));
  if (!ctx->sg.theta) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "surrogate not initialised");
  if (!sums) return fail(ctx, SFSIM_ATMO_ERR_ARGUMENT, "sums is NULL");
  const size_t nS = (size_t)count4(p->scatter_shape);
  float4* out[kMlpOrders];
  for (int k = 0; k < kMlpOrders; ++k) {
    TRY(ensure(ctx, ctx->pred[k], nS * sizeof(float4)));
    out[k] = (float4*)ctx->pred[k].ptr;
  }
  const DevParams P = dev_params(ctx, p);
  HIP_TRY(ctx, surrogate_predict_tables(ctx->sg, P, code_of(p), out, ctx->stream));
  for (int k = 0; k < kMlpOrders; ++k) TRY(download_table(ctx, out[k], sums + (size_t)k * nS * 3, nS));
  return 0;
}

int sfsim_atmo_write_scatter(const char* path, const float* data, size_t count) {
  if (!path || (!data && count)) return SFSIM_ATMO_ERR_ARGUMENT;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SFSIM_ATMO_ERR_IO;
  // little-endian float32 (util.clj:227-240); gfx950 hosts are little-endian x86-64
  size_t w = count ? std::fwrite(data, sizeof(float), count, f) : 0;
  int rc = std::fclose(f);
  return (w == count && rc == 0) ? 0 : SFSIM_ATMO_ERR_IO;
}

int sfsim_atmo_read_scatter(const char* path, float* data, size_t count) {
  if (!path || (!data && count)) return SFSIM_ATMO_ERR_ARGUMENT;
  FILE* f = std::fopen(path, "rb");
  if (!f) return SFSIM_ATMO_ERR_IO;
  size_t r = count ? std::fread(data, sizeof(float), count, f) : 0;
  int extra = std::fgetc(f);
  std::fclose(f);
  return (r == count && extra == EOF) ? 0 : SFSIM_ATMO_ERR_IO;
}

// Units of the reference (units.clj:14-19): degrees Rankine, feet, pounds-force, slugs.
namespace {
constexpr double kRankine = 5.0 / 9.0, kFoot = 0.3048, kPoundForce = 0.45359237 * 9.8066500286389,
                 kSlug = 14.5939029372;
double isa_temperature(double feet) {  // atmosphere.clj:425-450
  if (feet <= 36089.0) return (518.69 - 3.5662e-3 * feet) * kRankine;
  if (feet <= 65617.0) return 389.99 * kRankine;
  return (389.99 + 5.4864e-4 * (feet - 65617.0)) * kRankine;
}
}  // namespace

int sfsim_atmo_standard_atmosphere(double height, double* temperature, double* pressure, double* density) {
  if (!std::isfinite(height)) return SFSIM_ATMO_ERR_NONFINITE;
  const double ft = height / kFoot, psf = kPoundForce / (kFoot * kFoot), spcf = kSlug / (kFoot * kFoot * kFoot);
  const double t = isa_temperature(ft), tr = t / kRankine;
  double pr, rho;
  if (ft <= 36089.0) {  // atmosphere.clj:453-456, 480-483
    pr = 1.1376e-11 * std::pow(tr, 5.2560) * psf;
    rho = 6.6277e-15 * std::pow(tr, 4.2560) * spcf;
  } else if (ft <= 65617.0) {  // atmosphere.clj:459-462, 486-489: density from the pressure in lbf / ft^2
    const double p_psf = 2678.4 * std::exp(-4.8063e-5 * ft);
    pr = p_psf * psf;
    rho = 1.4939e-6 * p_psf * spcf;
  } else {  // atmosphere.clj:465-468, 492-494
    pr = 3.7930e+90 * std::pow(tr, -34.164) * psf;
    rho = 2.2099e+87 * std::pow(tr, -35.164) * spcf;
  }
  if (temperature) *temperature = t;
  if (pressure) *pressure = pr;
  if (density) *density = rho;
  return 0;
}

double sfsim_atmo_speed_of_sound(double temperature) { return 331.3 * std::sqrt(temperature / 273.15); }

int sfsim_atmo_sphere_directions(int half, int steps) {
  int theta_steps = half ? (steps >> 2) : (steps >> 1);
  if (theta_steps < 1) return 0;
  double theta_range = half ? M_PI / 2 : M_PI;
  int total = 0;
  for (int i = 0; i < theta_steps; ++i)
    total += (int)std::ceil(std::sin(theta_range * ((0.5 + i) / theta_steps)) * steps);
  return total;
}

}  // extern "C"
