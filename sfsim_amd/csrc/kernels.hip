// HIP kernels (gfx950) of the atmosphere lookup-table precompute.
//
// Mapping: one table entry per lane.  The 4D kernels run one workgroup per view ray
// (height index, elevation index) of the scatter grid, in XCD-aware order: the ray geometry, the
// per-sample altitudes and the view-ray transmittance T(x, p_i) depend only on the ray, so they are
// computed once per build (ray_geometry_kernel) and every lane (sun-elevation, heading entry) of that
// ray reuses them.  Quadratures follow the reference (midpoint nodes of ray.clj:19-30, latitude rings
// of sphere.clj:62-105, O(n^2) transmittance of atmosphere.clj:192-200); only loop-invariant work is
// shared, and single scattering reads each sample's sun optical depth from a table of the same node
// sums (sun_depth_kernel).  Geometry (intersections, the sun direction at each sample) is FP64; the
// extinction sums along rays are FP32 with compensated accumulation, switching to FP64 for segments
// whose lowest point is interior and below DevParams::graze_height.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "kernels.hpp"

namespace sfsim {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxRayChunk = 8;  // samples per slab chunk of the ray-scatter kernel

__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
// min(floor(x), hi) for a coordinate already clamped to x >= 0, as one med3 (fminf would quiet a
// signalling NaN in hi on every call)
__device__ __forceinline__ float lower_cell(float x, float hi) { return __builtin_amdgcn_fmed3f(floorf(x), hi, 0.0f); }

// top_exit_distance (geometry.hpp) for a point inside the atmosphere (h <= H), branch-free, with the
// FP64 square root and reciprocal from the FP32 hardware estimates refined by one Newton step each
// (relative error ~1e-14, against ~1e-16 for the correctly rounded library sequences that cost several
// times as many instructions).  The discriminant pd^2 + (Rt^2 - r^2) lies in [0, 1e14]; it is clamped
// below at 1e-30 (an exit distance under 1e-14 m in place of the zero it stands for) so that the FP32
// estimates stay in their normal range.
__device__ __forceinline__ double sun_exit_distance(const Geo& g, double h, double pd) {
  const double c = (g.H - h) * (2.0 * g.R + g.H + h);
  const double disc = fmax(fma(pd, pd, c), 1e-30);
  double y = (double)__builtin_amdgcn_rsqf((float)disc);
  y = y * fma(-0.5 * disc * y, y, 1.5);
  const double l2 = disc * y;
  double z = (double)__builtin_amdgcn_rcpf((float)(pd + l2));
  z = z * fma(-(pd + l2), z, 2.0);
  return pd > 0.0 ? c * z : l2 - pd;
}

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
  float y = x - comp;
  float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// atmosphere.clj:56-61 in FP32
__device__ __forceinline__ float phase_f(float g, float mu) {
  float g2 = g * g;
  float x = (1.0f + g2) - 2.0f * g * mu;
  return (3.0f * (1.0f - g2) * (1.0f + mu * mu)) / (8.0f * 3.14159265358979f * (2.0f + g2) * (x * sqrtf(x)));
}

// Altitude along a ray, centred on the segment's lowest point.  With u(t) = r(t)^2 - R^2 =
// h0 (2R + h0) + t (2 pd0 + t) and any shift t = s + ts:  u = u(ts) + s (b' + s), b' = 2 pd0 + 2 ts.
// Taking ts = clamp(-pd0, 0, D) (the point of the segment closest to the planet centre) leaves no
// cancellation: near ts, where the density is largest, u ~ u(ts) is exact, and elsewhere s and
// b' + s have the same sign.  h = u / (R + sqrt(R^2 + u)).
struct Segment {
  float c0, b, s0, dt;  // FP32 node constants in units of R: x_j = c0 + s_j (b + s_j), s_j = j dt + s0
  double c0d, bd, s0d, dtd;  // the same in metres (u_j = c0d + s_j (bd + s_j)) for the FP64 sums
  bool graze;           // lowest point interior and below the FP64 altitude threshold
};

// The midpoint nodes of a segment of length D in P.ray_steps steps (every segment of the path has that many).
__device__ __forceinline__ Segment make_segment(const DevParams& P, double h0, double pd0, double D) {
  Segment g;
  double ts = fmin(fmax(-pd0, 0.0), D);
  g.c0d = h0 * (2.0 * P.g.R + h0) + ts * (2.0 * pd0 + ts);
  g.bd = 2.0 * pd0 + 2.0 * ts;
  g.dtd = D * P.inv_steps;
  g.s0d = 0.5 * g.dtd - ts;
  g.c0 = (float)(g.c0d * P.inv_R * P.inv_R);
  g.b = (float)(g.bd * P.inv_R);
  g.s0 = (float)(g.s0d * P.inv_R);
  g.dt = (float)(g.dtd * P.inv_R);
  g.graze = (-pd0 > 0.0) && (-pd0 < D) && (g.c0d < P.graze_u);
  return g;
}

// Sums over the n midpoint nodes of exp(-h / H_mie) and exp(-h / H_rayleigh) with FP32 altitudes;
// 8-node partial sums, Kahan-compensated across them.  Node j sits at s_j = s0 + j dt (units of R,
// measured from the segment's lowest point), x_j = c0 + s_j (b + s_j) = (r_j^2 - R^2) / R^2, and the
// Mie exponent log2(exp(-h_j / H_mie)) is x_j times a polynomial of degree DEG in x_j
// (DevParams::alt_poly, alt_deg), so a node costs DEG + 3 FMAs, two multiplies, an add and the two
// exponentials.
template <int DEG>
__device__ float2 density_sums_f32_deg(const DevParams& P, const Segment& g, int n) {
  const float a0 = P.alt_poly[0], a1 = P.alt_poly[1], a2 = P.alt_poly[2], a3 = P.alt_poly[3];
  const float rho = P.ray_over_mie;
  auto exponent = [&](float x) {
    const float p = DEG == 3 ? fmaf(fmaf(fmaf(a3, x, a2), x, a1), x, a0) : fmaf(fmaf(a2, x, a1), x, a0);
    return x * p;
  };
  float sm = 0.f, sr = 0.f, cm = 0.f, cr = 0.f;
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    const float base = fmaf((float)j, g.dt, g.s0);
    float pm = 0.f, pr = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float s = fmaf((float)q, g.dt, base);
      const float em = exponent(fmaf(s, s + g.b, g.c0));
      const float m = fexp2(em), r = fexp2(em * rho);
      if (q == 0) {
        pm = m;
        pr = r;
      } else {
        pm += m;
        pr += r;
      }
    }
    kahan_add(sm, cm, pm);
    kahan_add(sr, cr, pr);
  }
  if (j < n) {
    float pm = 0.f, pr = 0.f;
    for (; j < n; ++j) {
      const float s = fmaf((float)j, g.dt, g.s0);
      const float em = exponent(fmaf(s, s + g.b, g.c0));
      pm += fexp2(em);
      pr += fexp2(em * rho);
    }
    kahan_add(sm, cm, pm);
    kahan_add(sr, cr, pr);
  }
  return make_float2(sm, sr);
}

__device__ __forceinline__ float2 density_sums_f32(const DevParams& P, const Segment& g, int n) {
  return P.alt_deg == 2 ? density_sums_f32_deg<2>(P, g, n) : density_sums_f32_deg<3>(P, g, n);
}

// Same sums with FP64 node positions, altitude exponents and accumulation (near-horizon and grazing
// segments): x_j = u_j / R^2 from the metre-valued constants and the FP64 cubic of DevParams; the
// exponentials are FP32 (their argument is rounded to FP32, a relative 6e-8).
__device__ double2 density_sums_f64(const DevParams& P, const Segment& g, int n) {
  const double inv_R2 = P.inv_R * P.inv_R;
  const double a0 = P.alt_poly_d[0], a1 = P.alt_poly_d[1], a2 = P.alt_poly_d[2], a3 = P.alt_poly_d[3];
  const float rho = P.ray_over_mie;
  double sm = 0.0, sr = 0.0;
  for (int j = 0; j < n; ++j) {
    const double s = fma((double)j, g.dtd, g.s0d);
    const double x = fma(s, s + g.bd, g.c0d) * inv_R2;
    const float em = (float)(x * fma(fma(fma(a3, x, a2), x, a1), x, a0));
    sm += (double)fexp2(em);
    sr += (double)fexp2(em * rho);
  }
  return make_double2(sm, sr);
}

// exp(-(beta_mie * integral of the Mie density + beta_rayleigh * ...)) per channel from the two sums,
// as exp2 of the log2(e)-scaled exponent.
__device__ __forceinline__ float3 transmittance_of(const DevParams& P, float sm, float sr, double D) {
  if (!(D > 0.0)) return make_float3(1.f, 1.f, 1.f);
  return make_float3(fexp2(-fmaf(P.ext_mie_l2[0], sm, P.ext_ray_l2[0] * sr)),
                     fexp2(-fmaf(P.ext_mie_l2[1], sm, P.ext_ray_l2[1] * sr)),
                     fexp2(-fmaf(P.ext_mie_l2[2], sm, P.ext_ray_l2[2] * sr)));
}

__device__ __forceinline__ float3 segment_transmittance_f32(const DevParams& P, const Segment& g, double D, int n) {
  float2 s = density_sums_f32(P, g, n);
  const float dt = (float)g.dtd;
  return transmittance_of(P, s.x * dt, s.y * dt, D);
}

__device__ __forceinline__ float3 segment_transmittance_f64(const DevParams& P, const Segment& g, double D, int n) {
  double2 s = density_sums_f64(P, g, n);
  return transmittance_of(P, (float)(s.x * g.dtd), (float)(s.y * g.dtd), D);
}

// Transmittance of the segment of length D from altitude h0 along a unit direction with
// dot(point, direction) = pd0, n midpoint samples (atmosphere.clj:114-125).  The FP64 variant is
// chosen per wavefront (when any active lane holds a grazing segment) so that the two loops never
// diverge within a wave.
__device__ float3 segment_transmittance(const DevParams& P, double h0, double pd0, double D, int n) {
  Segment g = make_segment(P, h0, pd0, D);
  if (__any(g.graze)) return segment_transmittance_f64(P, g, D, n);
  return segment_transmittance_f32(P, g, D, n);
}

// ------------------------------------------------------------------ multilinear lookups
// interpolate.clj:75-98: per dimension i = clip(c), u = floor(i), v = clip(u + 1), s = i - u.
struct Lerp {
  int i0, i1;
  float s;
};

__device__ __forceinline__ Lerp make_lerp(float c, int N) {
  float i = fminf(fmaxf(c, 0.0f), (float)(N - 1));
  float u = floorf(i);
  Lerp L;
  L.i0 = (int)u;
  L.i1 = min(L.i0 + 1, N - 1);
  L.s = i - u;
  return L;
}

// The two corners of a lerp by index parity: idx[p], w[p] for the corner whose index has parity p
// (i0 with weight 1 - s, i1 with weight s); at the upper edge (i1 == i0, s == 0) the one node carries
// weight 1 and the other parity is empty (idx -1, weight 0).
__device__ __forceinline__ void corner_by_parity(const Lerp& L, int idx[2], float w[2]) {
  const bool odd = (L.i0 & 1) != 0;
  const int other = L.i1 != L.i0 ? L.i1 : -1;
  const float w_other = L.i1 != L.i0 ? L.s : 0.0f, w_lower = 1.0f - L.s;
  idx[0] = odd ? other : L.i0;
  w[0] = odd ? w_other : w_lower;
  idx[1] = odd ? L.i0 : other;
  w[1] = odd ? w_lower : w_other;
}

__device__ __forceinline__ float3 mix3(float3 a, float3 b, float s) {
  float r = 1.0f - s;
  return make_float3(a.x * r + b.x * s, a.y * r + b.y * s, a.z * r + b.z * s);
}

__device__ __forceinline__ float3 ld3(const float4* __restrict__ T, long long i) {
  float4 v = T[i];
  return make_float3(v.x, v.y, v.z);
}

// 2D table [N0][N1]
__device__ __forceinline__ float3 lookup2(const float4* __restrict__ T, int N1, Lerp a, Lerp b) {
  float3 u = mix3(ld3(T, (long long)a.i0 * N1 + b.i0), ld3(T, (long long)a.i0 * N1 + b.i1), b.s);
  float3 v = mix3(ld3(T, (long long)a.i1 * N1 + b.i0), ld3(T, (long long)a.i1 * N1 + b.i1), b.s);
  return mix3(u, v, a.s);
}

// lookup2 when one of the two weights may be exactly zero (a grid coordinate that is an integer):
// mix(u, v, 0) = u exactly, so the row or column that it would mix in is not read.  Same result as
// lookup2, bit for bit.
__device__ __forceinline__ float3 lookup2_skip(const float4* __restrict__ T, int N1, Lerp a, Lerp b) {
  if (a.s == 0.0f) return mix3(ld3(T, (long long)a.i0 * N1 + b.i0), ld3(T, (long long)a.i0 * N1 + b.i1), b.s);
  if (b.s == 0.0f) return mix3(ld3(T, (long long)a.i0 * N1 + b.i0), ld3(T, (long long)a.i1 * N1 + b.i0), a.s);
  return lookup2(T, N1, a, b);
}

// 4D table [N0][N1][N2][N3], nested like interpolate-value (outer dimension mixed last)
__device__ __forceinline__ float3 lookup4(const float4* __restrict__ T, int N1, int N2, int N3, Lerp a, Lerp b, Lerp c,
                                          Lerp d) {
  float3 r[2];
#pragma unroll
  for (int ia = 0; ia < 2; ++ia) {
    long long ba = (long long)(ia ? a.i1 : a.i0) * N1;
    float3 rb[2];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      long long bb = (ba + (ib ? b.i1 : b.i0)) * N2;
      long long c0 = (bb + c.i0) * N3, c1 = (bb + c.i1) * N3;
      float3 u = mix3(ld3(T, c0 + d.i0), ld3(T, c0 + d.i1), d.s);
      float3 v = mix3(ld3(T, c1 + d.i0), ld3(T, c1 + d.i1), d.s);
      rb[ib] = mix3(u, v, c.s);
    }
    r[ia] = mix3(rb[0], rb[1], b.s);
  }
  return mix3(r[0], r[1], a.s);
}

// lookup4 of the texels ld(i) without the reads that a zero weight would mix in: mix3(u, v, 0) is u
// exactly, so each dimension whose weight is zero reads one side only.  Same operations in the same order
// as lookup4 where they are not skipped, so the same result bit for bit.
template <class Ld>
__device__ __forceinline__ float3 lookup4_skip(Ld ld, int N1, int N2, int N3, Lerp a, Lerp b, Lerp c, Lerp d) {
  auto row = [&](int i, int j, int k) {
    const long long base = (((long long)i * N1 + j) * N2 + k) * N3;
    const float3 u = ld(base + d.i0);
    return d.s == 0.0f ? u : mix3(u, ld(base + d.i1), d.s);
  };
  auto cell = [&](int i, int j) {
    const float3 u = row(i, j, c.i0);
    return c.s == 0.0f ? u : mix3(u, row(i, j, c.i1), c.s);
  };
  auto side = [&](int i) {
    const float3 u = cell(i, b.i0);
    return b.s == 0.0f ? u : mix3(u, cell(i, b.i1), b.s);
  };
  const float3 r0 = side(a.i0);
  return a.s == 0.0f ? r0 : mix3(r0, side(a.i1), a.s);
}

__device__ __forceinline__ void count(WorkCounters* wc, unsigned long long single, unsigned long long samples,
                                      unsigned long long l4, unsigned long long l2) {
  if (wc) {
    if (single) atomicAdd(&wc->single_samples, single);
    if (samples) atomicAdd(&wc->samples, samples);
    if (l4) atomicAdd(&wc->lookups4, l4);
    if (l2) atomicAdd(&wc->lookups2, l2);
  }
}

// XCD-aware order of the view rays: workgroups are dispatched to the 8 XCDs round robin (workgroup b
// to XCD b mod 8), each with its own L2.  Giving XCD x the contiguous range of rays
// [x rays / 8, (x + 1) rays / 8) (4 of 32 heights at the full workload) keeps the tables that a height
// range reads (dJ cells, point-scatter slabs, sun-table rows) in fewer L2s.  A bijection when rays is a
// multiple of 8; identity otherwise.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_ray(int b, int rays) {
  if (rays % kXcds != 0) return b;
  return (b % kXcds) * (rays / kXcds) + b / kXcds;
}

// ------------------------------------------------------------------------------ kernels
// T table: atmosphere_lut.clj:74 over transmittance-space (atmosphere.clj:291-319)
__device__ __forceinline__ void transmittance_body(const DevParams& P, float4* __restrict__ out,
                                                               WorkCounters* wc, unsigned blk) {
  const int NH = P.nT[0], NE = P.nT[1];
  int idx = blk * blockDim.x + threadIdx.x;
  if (idx >= NH * NE) return;
  int ih = idx / NE, ie = idx - ih * NE;
  double r0 = index_to_radius(P.g, NH, ih), h0 = index_to_altitude(P.g, NH, ih);
  double se, ce;
  bool above;
  index_to_elevation(P.g, NE, r0, ie, se, ce, above);
  double pd0 = r0 * se;
  double D = above ? top_exit_distance(P.g, h0, pd0) : ground_distance(P.g, h0, pd0);
  float3 T = segment_transmittance(P, h0, pd0, D, P.ray_steps);
  out[idx] = make_float4(T.x, T.y, T.z, 0.f);
  count(wc, 0, P.ray_steps, 0, 0);
}

__global__ void __launch_bounds__(kBlock) transmittance_kernel(const DevParams P, float4* __restrict__ out,
                                                               WorkCounters* wc) {
  transmittance_body(P, out, wc, blockIdx.x);
}

// dE0 table: surface-radiance-base (atmosphere.clj:131-137) over surface-radiance-space
__device__ __forceinline__ void surface_radiance_base_body(const DevParams& P, float4* __restrict__ out,
                                                                       WorkCounters* wc, unsigned blk) {
  const int NH = P.nE[0], NL = P.nE[1];
  int idx = blk * blockDim.x + threadIdx.x;
  if (idx >= NH * NL) return;
  int ih = idx / NL, il = idx - ih * NL;
  double r0 = index_to_radius(P.g, NH, ih), h0 = index_to_altitude(P.g, NH, ih);
  double s = index_to_sin_sun_elevation(NL, il);
  double pd0 = r0 * s;  // dot(x, l) with l = (s, c, 0)
  double D = top_exit_distance(P.g, h0, pd0);
  float3 T = segment_transmittance(P, h0, pd0, D, P.ray_steps);
  float cosf_ = (float)fmax(0.0, s);  // max(0, normal . l), normal = (1 0 0)
  out[idx] = make_float4(T.x * P.intensity[0] * cosf_, T.y * P.intensity[1] * cosf_, T.z * P.intensity[2] * cosf_, 0.f);
  count(wc, 0, P.ray_steps, 0, 0);
}

__global__ void __launch_bounds__(kBlock) surface_radiance_base_kernel(const DevParams P, float4* __restrict__ out,
                                                                       WorkCounters* wc) {
  surface_radiance_base_body(P, out, wc, blockIdx.x);
}

// Per-ray setup shared by the 4D kernels: grid point x = (r0 0 0), view v = (se ce 0), ray length D.
struct Ray {
  double r0, h0, se, ce, pd0, D;
  bool above;
};

__device__ __forceinline__ Ray make_ray(const DevParams& P, int ih, int ie) {
  Ray r;
  r.r0 = index_to_radius(P.g, P.nS[0], ih);
  r.h0 = index_to_altitude(P.g, P.nS[0], ih);
  index_to_elevation(P.g, P.nS[1], r.r0, ie, r.se, r.ce, r.above);
  r.pd0 = r.r0 * r.se;
  r.D = r.above ? top_exit_distance(P.g, r.h0, r.pd0) : ground_distance(P.g, r.h0, r.pd0);
  return r;
}

// R1 (Rayleigh with phase) and M1 (Mie strength) in one pass: atmosphere_lut.clj:70-78,
// atmosphere.clj:154-200.  One wavefront per (view ray, group of 64 (sun elevation, heading)
// entries); groups of high sun elevation (most samples lit) are launched first so that the light
// groups fill the end of the grid.  The ray's samples are processed in chunks of kSSChunk: each lane
// marks the samples where its sun is above the horizon, the wavefront compacts the lit
// (entry, sample) pairs into an LDS list and all 64 lanes take items from it, so lanes never idle on
// unlit samples; each entry then adds its results in sample order (deterministic sums).
constexpr int kSSWave = 64;
constexpr int kSSChunk = 8;

__device__ __forceinline__ int wave_exclusive_scan(int v, int lane) {
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x - v;
}

__global__ void __launch_bounds__(kSSWave) single_scatter_kernel(const DevParams P, float4* __restrict__ R1,
                                                                 float4* __restrict__ M1, WorkCounters* wc) {
  extern __shared__ unsigned char s_raw[];
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int NLA = NL * NA;
  const int n = P.ray_steps;
  double* s_t = (double*)s_raw;                  // [n] sample distance along the ray
  double* s_h = s_t + n;                         // [n] sample altitude
  float4* s_w = (float4*)(s_h + n);              // [n] T(x, p_i) rgb, w: exp(-h_i / H_mie)
  float* s_bray = (float*)(s_w + n);             // [n] exp(-h_i / H_rayleigh)
  float* s_res = s_bray + ((n + 3) & ~3);  // [3][64 * kSSChunk] sun transmittance per item (r, g, b planes)
  double* s_xl = (double*)(s_res + 3 * kSSWave * kSSChunk);  // [64] dot(x, l) of each lane's entry
  double* s_vl = s_xl + kSSWave;                          // [64] dot(v, l)
  unsigned short* s_item = (unsigned short*)(s_vl + kSSWave);  // lane << 8 | chunk sample

  const int rays = NH * NE;
  const int groups = (NLA + kSSWave - 1) / kSSWave;
  const int group = groups - 1 - blockIdx.x / rays;
  const int rayid = blockIdx.x - (groups - 1 - group) * rays;
  const int ih = rayid / NE, ie = rayid - ih * NE;
  const int lane = threadIdx.x;
  const Ray ray = make_ray(P, ih, ie);
  const double dt = ray.D / n;
  for (int i = lane; i < n; i += kSSWave) {
    double t = (i + 0.5) * dt;
    double h = altitude_along(P.g, ray.h0, ray.pd0, t);
    s_t[i] = t;
    s_h[i] = h;
    float3 T = segment_transmittance(P, ray.h0, ray.pd0, t, n);
    s_w[i] = make_float4(T.x, T.y, T.z, (float)exp(-h * P.inv_scale_mie));
    s_bray[i] = (float)exp(-h * P.inv_scale_ray);
  }
  const int k = group * kSSWave + lane;
  const bool valid = k < NLA;
  double vl = 0.0, xl = 0.0;
  if (valid) {
    int is = k / NA, ia = k - is * NA;
    double sin_s = index_to_sin_sun_elevation(NL, is);
    double l0, l1, l2;
    index_to_sun_direction(NA, ray.se, ray.ce, sin_s, ia, l0, l1, l2);
    vl = ray.se * l0 + ray.ce * l1;
    xl = ray.r0 * l0;
  }
  s_xl[lane] = xl;
  s_vl[lane] = vl;
  __syncthreads();
  float3 accR = make_float3(0.f, 0.f, 0.f), accM = make_float3(0.f, 0.f, 0.f);
  unsigned long long samples = 0;
  for (int i0 = 0; i0 < n; i0 += kSSChunk) {
    const int cn = min(kSSChunk, n - i0);
    // lit samples (atmosphere.clj:154-160), split by the optical-depth path their sun segment takes:
    // FP64 when its lowest point (at distance -dot(p_i, l) when that is positive) lies below
    // graze_height, FP32 otherwise
    unsigned m32 = 0, m64 = 0;
    if (valid)
      for (int c = 0; c < cn; ++c) {
        const int i = i0 + c;
        const double h = s_h[i], pl = xl + s_t[i] * vl;
        if (is_above_horizon(P.g, h, pl)) {
          if (pl < 0.0 && h * (2.0 * P.g.R + h) - pl * pl < P.graze_u)
            m64 |= 1u << c;
          else
            m32 |= 1u << c;
        }
      }
    const int c32 = __popc(m32), c64 = __popc(m64);
    const int o32 = wave_exclusive_scan(c32, lane), o64 = wave_exclusive_scan(c64, lane);
    const int t32 = __shfl(o32 + c32, kSSWave - 1, 64), t64 = __shfl(o64 + c64, kSSWave - 1, 64);
    {
      unsigned m = m32;
      for (int q = 0; q < c32; ++q) {
        int c = __ffs(m) - 1;
        m &= m - 1;
        s_item[o32 + q] = (unsigned short)((lane << 8) | c);
      }
      m = m64;
      for (int q = 0; q < c64; ++q) {
        int c = __ffs(m) - 1;
        m &= m - 1;
        s_item[t32 + o64 + q] = (unsigned short)((lane << 8) | c);
      }
    }
    __syncthreads();
    // sun transmittance T(p_i, l) of every listed item: the FP32 list, then the FP64 list, so the two
    // loops never share a wavefront
    for (int it = lane; it < t32; it += kSSWave) {
      const unsigned item = s_item[it];
      const int owner = item >> 8, i = i0 + (int)(item & 255);
      const double h = s_h[i];
      const double pl = s_xl[owner] + s_t[i] * s_vl[owner];  // dot(p_i, l)
      const double Ds = sun_exit_distance(P.g, h, pl);
      const float3 Ts = segment_transmittance_f32(P, make_segment(P, h, pl, Ds), Ds, n);
      s_res[it] = Ts.x;
      s_res[it + kSSWave * kSSChunk] = Ts.y;
      s_res[it + 2 * kSSWave * kSSChunk] = Ts.z;
    }
    for (int it = t32 + lane; it < t32 + t64; it += kSSWave) {
      const unsigned item = s_item[it];
      const int owner = item >> 8, i = i0 + (int)(item & 255);
      const double h = s_h[i];
      const double pl = s_xl[owner] + s_t[i] * s_vl[owner];
      const double Ds = sun_exit_distance(P.g, h, pl);
      const float3 Ts = segment_transmittance_f64(P, make_segment(P, h, pl, Ds), Ds, n);
      s_res[it] = Ts.x;
      s_res[it + kSSWave * kSSChunk] = Ts.y;
      s_res[it + 2 * kSSWave * kSSChunk] = Ts.z;
    }
    __syncthreads();
    {
      // each entry adds its lit samples in sample order (deterministic sums)
      unsigned m = m32 | m64;
      int r32 = o32, r64 = t32 + o64;
      while (m) {
        const int c = __ffs(m) - 1;
        m &= m - 1;
        const int idx = ((m32 >> c) & 1u) ? r32++ : r64++;
        const float3 Ts = make_float3(s_res[idx], s_res[idx + kSSWave * kSSChunk], s_res[idx + 2 * kSSWave * kSSChunk]);
        const float4 w = s_w[i0 + c];
        const float br = s_bray[i0 + c];
        const float3 t3 = make_float3(w.x * Ts.x, w.y * Ts.y, w.z * Ts.z);
        accR.x += t3.x * br; accR.y += t3.y * br; accR.z += t3.z * br;
        accM.x += t3.x * w.w; accM.y += t3.y * w.w; accM.z += t3.z * w.w;
      }
    }
    samples += (unsigned long long)(c32 + c64) * n;
    __syncthreads();
  }
  if (valid) {
    const long long o = (long long)rayid * NLA + k;
    float a = (float)dt;
    float pr = phase_f(P.g_ray, (float)vl) * a;
    R1[o] = make_float4(accR.x * P.sca_ray[0] * P.intensity[0] * pr, accR.y * P.sca_ray[1] * P.intensity[1] * pr,
                        accR.z * P.sca_ray[2] * P.intensity[2] * pr, 0.f);
    M1[o] = make_float4(accM.x * P.sca_mie[0] * P.intensity[0] * a, accM.y * P.sca_mie[1] * P.intensity[1] * a,
                        accM.z * P.sca_mie[2] * P.intensity[2] * a, 0.f);
  }
  unsigned long long view = (lane < n) ? (unsigned long long)n * ((n - lane + kSSWave - 1) / kSSWave) : 0;
  count(wc, samples, samples + view, 0, 0);
}

// ------------------------------------------------------------ tabulated sun optical depths
// The sun segment of a sample p (atmosphere.clj:154-160: transmittance from p toward l to the top of
// the atmosphere, ray_steps midpoint nodes) depends on p and l only through r = |p| and pl = dot(p, l),
// so its two node sums (Mie and Rayleigh density times the node spacing) are functions of two
// variables.  Rows: a = rho / rho_max, rho = sqrt(r^2 - R^2), uniform in the horizon distance (the
// altitude grid is finest at the ground); columns: b = (D - D_min) / (D_max - D_min) for the exit
// distance D, D_min = Rt - r (sun at the zenith) and D_max = rho + rho_max (sun on the horizon), which
// spans every direction in which the sun is above the horizon.  Entries hold log2 of the node sums
// (FP64 for grazing segments, compensated FP32 otherwise); interpolation is bilinear in log2
// (exponential in altitude between grid points).
__device__ __forceinline__ void sun_depth_body(const DevParams& P, unsigned blk) {
  const int rows = P.sun_rows, cols = P.sun_cols;
  const int idx = blk * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int ia = idx / cols, ib = idx - ia * cols;
  const double rho = P.g.rho_max * ia / (rows - 1);
  const double h = rho * rho / (P.g.R + sqrt(P.g.R * P.g.R + rho * rho));
  const double c = (P.g.rho_max - rho) * (P.g.rho_max + rho);  // Rt^2 - r^2
  const double dmin = P.g.H - h, dmax = rho + P.g.rho_max;
  const double D = dmin + (dmax - dmin) * ib / (cols - 1);
  float2 out = make_float2(-256.f, -256.f);  // empty segment (the top of the atmosphere, sun at the zenith)
  if (D > 0.0) {
    const double pl = (c - D * D) / (2.0 * D);  // from D = -pl + sqrt(pl^2 + c)
    const Segment g = make_segment(P, h, pl, D);
    double sm, sr;
    if (__any(g.graze)) {  // the same FP32/FP64 choice as segment_transmittance
      const double2 sum = density_sums_f64(P, g, P.ray_steps);
      sm = sum.x * g.dtd;
      sr = sum.y * g.dtd;
    } else {
      const float2 sum = density_sums_f32(P, g, P.ray_steps);
      sm = (double)sum.x * g.dtd;
      sr = (double)sum.y * g.dtd;
    }
    out = make_float2((float)fmax(log2(sm), -256.0), (float)fmax(log2(sr), -256.0));
  }
  P.sun_depth[idx] = out;
}

__global__ void __launch_bounds__(kBlock) sun_depth_kernel(const DevParams P) {
  sun_depth_body(P, blockIdx.x);
}

// Geometry of every (view ray, sample) pair, one thread each: the sample position, T(x, p_i) over
// ray_steps midpoint nodes, the height and elevation coordinates of the dJ lookup, the terms of the
// sun-elevation coordinate and, with a sun table, the sample's row and exit-distance terms in it.
// Shared by the single-scatter kernel and the ray-scatter launches of all iterations.
__device__ __forceinline__ void ray_geometry_body(const DevParams& P, WorkCounters* wc, unsigned blk) {
  const int n = P.ray_steps, NH = P.nS[0], NE = P.nS[1];
  const long long g = (long long)blk * blockDim.x + threadIdx.x;
  if (g >= (long long)NH * NE * n) return;
  const int rayid = (int)(g / n), i = (int)(g - (long long)rayid * n);
  const int ih = rayid / NE, ie = rayid - ih * NE;
  const Ray ray = make_ray(P, ih, ie);
  const double t = (i + 0.5) * (ray.D / n);
  const double h = altitude_along(P.g, ray.h0, ray.pd0, t), r = P.g.R + h;
  const float3 T = segment_transmittance(P, ray.h0, ray.pd0, t, n);
  const RayGeoView V = ray_geo_view(P);
  const double k3 = -3.0 * 1.4426950408889634;
  V.th[g] = make_double2(t, h);
  V.T[g] = make_float4(T.x, T.y, T.z, (float)exp(-h * P.inv_scale_mie));
  const float hc = (float)altitude_to_index(P.g, NH, h);
  const float ec = (float)elevation_to_index(P.g, NE, r, ray.pd0 + t, ray.above);
  V.rs[g] = make_float4(hc, ec, (float)(k3 / r), (float)(k3 * t / r));
  {  // the fold's dJ nodes by slot: the corner with height parity ph and elevation parity pe goes to slot
     // 2 ph + pe (at the upper edge of the grid, where i1 == i0 and the weight of i1 is zero, the
     // coinciding corners share one slot and the slots of the other parity stay unused)
    const int NLA = P.nS[2] * P.nS[3];
    const Lerp Lh = make_lerp(hc, NH), Le = make_lerp(ec, NE);
    int hs[2], es[2];
    float hw[2], ew[2];
    corner_by_parity(Lh, hs, hw);
    corner_by_parity(Le, es, ew);
    auto off = [&](int ph, int pe) { return hs[ph] < 0 || es[pe] < 0 ? -1 : (hs[ph] * NE + es[pe]) * NLA; };
    auto wt = [&](int ph, int pe) { return hs[ph] < 0 || es[pe] < 0 ? 0.0f : ew[pe] * hw[ph]; };
    V.fold[g] = FoldStep{make_int4(off(0, 0), off(0, 1), off(1, 0), off(1, 1)),
                         make_float4(wt(0, 0), wt(0, 1), wt(1, 0), wt(1, 1))};
  }
  if (P.sun_rows > 0) {
    const int cols = P.sun_cols;
    // clamped like horizon_distance: the sample of a ray that grazes the ground at its tangent point (the
    // middle sample of an odd number of them) has h = 0 up to FP64 rounding, of either sign
    const double rho = sqrt(fmax(0.0, h * (2.0 * P.g.R + h)));
    const double dmin = P.g.H - h, dmax = rho + P.g.rho_max;
    V.ss[g] = make_float4((float)((P.g.H - h) * (2.0 * P.g.R + P.g.H + h)), (float)dmin,
                          (float)((cols - 1) / (dmax - dmin)), (float)exp(-h * P.inv_scale_ray));
    const float x = (float)(rho / P.g.rho_max * (P.sun_rows - 1));
    const float row = fminf(floorf(x), (float)(P.sun_rows - 2));
    V.row[g] = make_float2(x - row, __int_as_float((int)row * cols));
  }
  count(wc, 0, n, 0, 0);
}

__global__ void __launch_bounds__(kBlock) ray_geometry_kernel(const DevParams P, WorkCounters* wc) {
  ray_geometry_body(P, wc, blockIdx.x);
}

// Single scatter with the sun segments' optical depths read from the sun_depth table: one block per
// view ray, one thread per (sun elevation, heading) entry, each adding its lit samples in sample order.
// The ray's samples (t, h, T(x, p_i), sun-table terms) come from the per-build geometry through LDS.
constexpr int kSSTBlock = 256;

template <bool kB>
__global__ void __launch_bounds__(kSSTBlock) single_scatter_table_kernel(
    const DevParams P0, const DevParams* __restrict__ PB, const double2* __restrict__ geo_th,
    const float4* __restrict__ geo_T, const float4* __restrict__ geo_ss, const float2* __restrict__ geo_row,
    float4* __restrict__ R1, float4* __restrict__ M1, WorkCounters* wc) {
  const DevParams& P = params_of<kB>(P0, PB);
  if constexpr (kB) {  // this atmosphere's tables and ray geometry (offsets of the restrict arguments, whose
                       // loads stay provably unclobbered and uniform: scalar loads)
    geo_th = atm_ptr<kB>(geo_th, P0);
    geo_T = atm_ptr<kB>(geo_T, P0);
    geo_ss = atm_ptr<kB>(geo_ss, P0);
    geo_row = atm_ptr<kB>(geo_row, P0);
    R1 = atm_ptr<kB>(R1, P);
    M1 = atm_ptr<kB>(M1, P);
  }
  const int NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int NLA = NL * NA;
  const int n = P.ray_steps, cols = P.sun_cols;
  const int rayid = xcd_ray(blockIdx.x, P.nS[0] * NE);
  const int ih = rayid / NE, ie = rayid - ih * NE;
  const Ray ray = make_ray(P, ih, ie);
  const double dt = ray.D / n;
  // this ray's samples from the per-build geometry, staged in LDS
  extern __shared__ unsigned char s_raw[];
  double2* s_th = (double2*)s_raw;      // [n] t_i, h_i
  float4* s_w = (float4*)(s_th + n);    // [n] T(x, p_i) rgb, w: exp(-h_i / H_mie)
  float4* s_g = s_w + n;                // [n] Rt^2 - r_i^2, D_min, (cols - 1) / (D_max - D_min), exp(-h_i / H_rayleigh)
  float2* s_r = (float2*)(s_g + n);     // [n] sun-table row fraction, start of the row (bits of an int)
  const long long g0 = (long long)rayid * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s_th[i] = geo_th[g0 + i];
    s_w[i] = geo_T[g0 + i];
    s_g[i] = geo_ss[g0 + i];
    s_r[i] = geo_row[g0 + i];
  }
  __syncthreads();
  const float2* __restrict__ tab = P.sun_depth;
  unsigned long long lit = 0;
  for (int k = threadIdx.x; k < NLA; k += blockDim.x) {
    // dot(v, l) and dot(x, l) in FP64 from the per-build sun map (the same FP64 evaluation)
    const double2 vx = P.sun_ss[(long long)rayid * NLA + k];
    const double vl = vx.x, xl = vx.y;
    float3 accR = make_float3(0.f, 0.f, 0.f), accM = make_float3(0.f, 0.f, 0.f);
    // branch-free: a sample with the sun below the horizon (atmosphere.clj:154-160) reads the table at
    // clamped, in-range coordinates and adds exactly zero, so consecutive samples' reads overlap
#pragma unroll 2
    for (int i = 0; i < n; ++i) {
      const double2 th = s_th[i];
      const double h = th.y, pl = xl + th.x * vl;  // dot(p_i, l)
      const bool sun = is_above_horizon(P.g, h, pl);
      lit += sun;
      const float on = sun ? 1.0f : 0.0f;
      const float4 gm = s_g[i];
      const float2 rr = s_r[i];
      const float plf = (float)pl;
      const float q = __builtin_amdgcn_sqrtf(fmaf(plf, plf, gm.x));
      const float D = plf > 0.f ? gm.x * __builtin_amdgcn_rcpf(plf + q) : q - plf;
      const float y = fminf(fmaxf((D - gm.y) * gm.z, 0.f), (float)(cols - 1));
      const float col = lower_cell(y, (float)(cols - 2));
      const float fy = y - col;
      const int base = __float_as_int(rr.y) + (int)col;
      const float2 v00 = tab[base], v01 = tab[base + 1], v10 = tab[base + cols], v11 = tab[base + cols + 1];
      const float lo_m = fmaf(fy, v01.x - v00.x, v00.x), hi_m = fmaf(fy, v11.x - v10.x, v10.x);
      const float lo_r = fmaf(fy, v01.y - v00.y, v00.y), hi_r = fmaf(fy, v11.y - v10.y, v10.y);
      const float sm = fexp2(fmaf(rr.x, hi_m - lo_m, lo_m)), sr = fexp2(fmaf(rr.x, hi_r - lo_r, lo_r));
      const float3 Ts = transmittance_of(P, sm, sr, 1.0);
      const float4 w = s_w[i];
      const float3 t3 = make_float3(w.x * Ts.x * on, w.y * Ts.y * on, w.z * Ts.z * on);
      accR.x += t3.x * gm.w; accR.y += t3.y * gm.w; accR.z += t3.z * gm.w;
      accM.x += t3.x * w.w; accM.y += t3.y * w.w; accM.z += t3.z * w.w;
    }
    const long long o = (long long)rayid * NLA + k;
    const float a = (float)dt;
    const float pr = phase_f(P.g_ray, (float)vl) * a;
    R1[o] = make_float4(accR.x * P.sca_ray[0] * P.intensity[0] * pr, accR.y * P.sca_ray[1] * P.intensity[1] * pr,
                        accR.z * P.sca_ray[2] * P.intensity[2] * pr, 0.f);
    M1[o] = make_float4(accM.x * P.sca_mie[0] * P.intensity[0] * a, accM.y * P.sca_mie[1] * P.intensity[1] * a,
                        accM.z * P.sca_mie[2] * P.intensity[2] * a, 0.f);
  }
  count(wc, lit, 0, 0, 0);
}

// dS' = ray-scatter of an interpolated dJ table (atmosphere_lut.clj:92,94; atmosphere.clj:192-200).
// One workgroup per view ray; lanes over (sun elevation, heading) entries, up to EPT per lane.
// At sample p_i the height and elevation coordinates of the dJ lookup are the same for every
// entry of the ray, so the workgroup first folds those two dimensions into a (sun elevation,
// heading) slab in LDS (coalesced reads of four dJ slabs), Q samples at a time; each lane then
// interpolates its own sun-elevation / heading coordinates in that slab.
// One entry's samples i0 .. i0 + qn - 1 of the folded slabs.  The sun-elevation coordinate of
// sun-elevation-to-index (atmosphere.clj:322-326) at p_i, (NL - 1) max(0, (1 - exp(-3 s - 0.6)) / (1 - exp(-3.6)))
// with s = dot(p_i, l) / |p_i| = (xl + t_i vl) / r_i, is A - A exp2(alpha_i xl + beta_i vl + c) with
// A = (NL - 1) / (1 - exp(-3.6)), alpha_i = -3 log2(e) / r_i, beta_i = alpha_i t_i, c = -0.6 log2(e), clamped
// to [0, NL - 1] (interpolate.clj:75-98 clips the coordinate); the lower row is min(floor, NL - 2), so
// the upper row is always the next one (weight 1 on it at the clamped top).  LDS byte offsets are
// formed in FP32 (exact below 2^24).
struct SunCoord {
  float A, c, top;  // A, c, NL - 1
};

__device__ __forceinline__ float sun_coord(const SunCoord& sc, float2 ab, float xl, float vl) {
  const float e = fexp2(fmaf(ab.y, vl, fmaf(ab.x, xl, sc.c)));
  return __builtin_amdgcn_fmed3f(fmaf(-sc.A, e, sc.A), 0.0f, sc.top);
}

// Sun-elevation grids with fewer rows than this take the exact form below.  Their rows lie far apart in
// sun elevation, so neighbouring rows can differ by orders of magnitude (a sun below the horizon against
// one above it), and the ~1e-7 index units of rounding in sun_coord's FP32 form weight such a row by as
// much: where the true coordinate is integral (a vertical ray keeps its sun elevation) or clamped (the
// exponent cancels to 0 at sin = -0.2), the FP64 oracle gives that row no weight at all.
constexpr int kExactSunRows = 16;

// sun_coord in FP64: z = -3 dot(p_i, l) / r_i - 0.6 from the sample's t_i, h_i and the entry's dot(x, l),
// dot(v, l) (1 / r_i: the FP32 estimate and one FP64 Newton step, relative error ~1e-14), the coordinate
// (NL - 1) (1 - e^z) / (1 - e^-3.6) with 1 - e^z = -expm1(z) (no cancellation near the clamp), clamped, and
// its lower row and weight, the weight formed in FP64 before rounding (a coordinate just above a row, say
// 1 + 5e-6, would lose 1% of its weight to the FP32 spacing at 1).  A64: (NL - 1) / (1 - e^-3.6).
__device__ __forceinline__ void sun_coord_exact(const SunCoord& sc, double A64, double R, double2 th, double xl,
                                                double vl, float& row, float& s) {
  const double r = R + th.y;
  double inv = (double)__builtin_amdgcn_rcpf((float)r);
  inv = fma(inv, fma(-r, inv, 1.0), inv);
  const double a = -3.0 * inv;
  const double z = fma(a * th.x, vl, fma(a, xl, -0.6));
  const double c = fmin(fmax(-expm1(z) * A64, 0.0), (double)sc.top);
  const double r0 = fmin(floor(c), (double)sc.top - 1.0);
  row = (float)r0;
  s = (float)(c - r0);
}

__device__ __forceinline__ float3 lerp3(float4 a, float4 b, float s) {
  return make_float3(fmaf(s, b.x - a.x, a.x), fmaf(s, b.y - a.y, a.y), fmaf(s, b.z - a.z, a.z));
}

__device__ __forceinline__ float4 lds4(const float4* base, unsigned byte_offset) {
  return *(const float4*)((const char*)base + byte_offset);
}

// Samples whose per-sample terms (the lookups' T and sun-coordinate terms, the fold's steps) are requested
// together, before the barrier that precedes their chunk: one wait per group instead of one per sample.
constexpr int kStepGroup = 4;

// Per-sample terms of kStepGroup consecutive samples for the lookups, in scalar registers: T(x, p_i) and
// (alpha_i, beta_i), requested together (before the barrier that precedes their chunk).
struct SampleGroup {
  float4 T[kStepGroup];
  float2 ab[kStepGroup];
  double2 th[kStepGroup];  // t_i, h_i (the exact sun coordinate only)
};

template <bool kExact>
__device__ __forceinline__ void load_sample_group(const float4* __restrict__ gT, const float4* __restrict__ gRS,
                                                  const double2* __restrict__ gTH, int i, int n, SampleGroup& g) {
#pragma unroll
  for (int q = 0; q < kStepGroup; ++q) {
    const int j = min(i + q, n - 1);  // in bounds
    g.T[q] = gT[j];
    const float4 r = gRS[j];
    g.ab[q] = make_float2(r.z, r.w);
    if (kExact) g.th[q] = gTH[j];
  }
}

// The entry's sun-elevation coordinate terms: FP32 (xl, vl) and, for the exact form, FP64 ones, R and A64.
struct SunTerms {
  float xl, vl;
  double xl64, vl64, R, A64;
};

// ray_scatter_samples over samples b .. min(b + kStepGroup, qn) - 1 of the chunk (b a multiple of
// kStepGroup), their terms in g: the same operations in the same order.
template <bool kColumn, bool kExact>
__device__ __forceinline__ float3 ray_scatter_sample_group(const SunCoord& sc, const float4* s_slab,
                                                           const SampleGroup& g, int b, int qn, float row_bytes,
                                                           float col_bytes, int NLA, const SunTerms& st, Lerp La,
                                                           float3 acc) {
  const unsigned rb = (unsigned)row_bytes;
  const float slab_bytes = (float)(NLA * (int)sizeof(float4));
  float q_bytes = fmaf((float)b, slab_bytes, col_bytes);  // byte offset of the entry's column in slab b (exact)
#pragma unroll
  for (int q = 0; q < kStepGroup; ++q, q_bytes += slab_bytes) {
    if (b + q < qn) {
      const float4 w = g.T[q];
      float row, s;
      if (kExact) {
        sun_coord_exact(sc, st.A64, st.R, g.th[q], st.xl64, st.vl64, row, s);
      } else {
        const float c = sun_coord(sc, g.ab[q], st.xl, st.vl);
        row = lower_cell(c, sc.top - 1.0f);
        s = c - row;
      }
      const unsigned off = (unsigned)fmaf(row, row_bytes, q_bytes);
      float3 v;
      if (kColumn) {
        v = lerp3(lds4(s_slab, off), lds4(s_slab, off + rb), s);
      } else {
        const unsigned d = (unsigned)(La.i1 - La.i0) * (unsigned)sizeof(float4);
        const float3 u = lerp3(lds4(s_slab, off), lds4(s_slab, off + d), La.s);
        const float3 t = lerp3(lds4(s_slab, off + rb), lds4(s_slab, off + rb + d), La.s);
        v = make_float3(fmaf(s, t.x - u.x, u.x), fmaf(s, t.y - u.y, u.y), fmaf(s, t.z - u.z, u.z));
      }
      acc.x = fmaf(w.x, v.x, acc.x);
      acc.y = fmaf(w.y, v.y, acc.y);
      acc.z = fmaf(w.z, v.z, acc.z);
    }
  }
  return acc;
}

template <bool kColumn, bool kExact>
__device__ __forceinline__ float3 ray_scatter_samples(const SunCoord& sc, const float4* s_slab,
                                                      const float4* __restrict__ gT, const float4* __restrict__ gRS,
                                                      const double2* __restrict__ gTH, int i0, int qn, float row_bytes,
                                                      float col_bytes, int NLA, const SunTerms& st, Lerp La,
                                                      float3 acc) {
  const unsigned rb = (unsigned)row_bytes;
  const float slab_bytes = (float)(NLA * (int)sizeof(float4));
  float q_bytes = col_bytes;  // byte offset of the entry's column in slab q
  for (int q = 0; q < qn; ++q, q_bytes += slab_bytes) {
    const int i = i0 + q;
    const float4 w = gT[i];  // workgroup-uniform: scalar loads
    const float4 g = gRS[i];
    float row, s;
    if (kExact) {
      sun_coord_exact(sc, st.A64, st.R, gTH[i], st.xl64, st.vl64, row, s);
    } else {
      const float c = sun_coord(sc, make_float2(g.z, g.w), st.xl, st.vl);
      row = lower_cell(c, sc.top - 1.0f);
      s = c - row;
    }
    const unsigned off = (unsigned)fmaf(row, row_bytes, q_bytes);
    float3 v;
    if (kColumn) {
      v = lerp3(lds4(s_slab, off), lds4(s_slab, off + rb), s);
    } else {
      const unsigned d = (unsigned)(La.i1 - La.i0) * (unsigned)sizeof(float4);
      const float3 u = lerp3(lds4(s_slab, off), lds4(s_slab, off + d), La.s);
      const float3 t = lerp3(lds4(s_slab, off + rb), lds4(s_slab, off + rb + d), La.s);
      v = make_float3(fmaf(s, t.x - u.x, u.x), fmaf(s, t.y - u.y, u.y), fmaf(s, t.z - u.z, u.z));
    }
    acc.x = fmaf(w.x, v.x, acc.x);
    acc.y = fmaf(w.y, v.y, acc.y);
    acc.z = fmaf(w.z, v.z, acc.z);
  }
  return acc;
}

// Register slots of the fold: slot ((h & 1) << 1) | (e & 1) holds dJ at node (h, e) for the folding
// thread's entry.  The four corners of a cell have distinct parities, so every sample's corners occupy
// the four slots, and a corner shared with the previous sample is already in its slot: a slot is
// re-read only when its node changes (about 60 node reads per ray of 64 samples instead of four per
// sample that changes cell).  The nodes and weights of every sample come from the per-build ray geometry.
struct FoldSlots {
  float4 v[4];
};

__device__ __forceinline__ void fold_slots_init(FoldSlots& f) {
  for (int s = 0; s < 4; ++s) f.v[s] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The per-build fold steps of one ray, read-only while a ray-scatter launch runs.  The single-atmosphere
// kernel reads them through its restrict argument; the batched one through the constant address space,
// where the compiler cannot otherwise prove them unclobbered and would use vector loads.
typedef __attribute__((address_space(4))) const int* ConstFoldSteps;

__device__ __forceinline__ FoldStep fold_step(const FoldStep* __restrict__ fs, int i) { return fs[i]; }

__device__ __forceinline__ FoldStep fold_step(ConstFoldSteps fs, int i) {
  const ConstFoldSteps p = fs + 8 * i;  // FoldStep: 8 dwords
  FoldStep st;
  st.off = make_int4(p[0], p[1], p[2], p[3]);
  st.wt = make_float4(__int_as_float(p[4]), __int_as_float(p[5]), __int_as_float(p[6]), __int_as_float(p[7]));
  return st;
}

// The folded value of one entry: the weighted sum of its four slots.
__device__ __forceinline__ float4 fold_value(const FoldSlots& f, const float4 w) {
  const float4 a = f.v[0], b = f.v[1], c = f.v[2], d = f.v[3];
  return make_float4(fmaf(w.w, d.x, fmaf(w.z, c.x, fmaf(w.y, b.x, w.x * a.x))),
                     fmaf(w.w, d.y, fmaf(w.z, c.y, fmaf(w.y, b.y, w.x * a.y))),
                     fmaf(w.w, d.z, fmaf(w.z, c.z, fmaf(w.y, b.z, w.x * a.z))), 0.f);
}

// Fold steps of kStepGroup consecutive samples, held in scalar registers: requested together, one wait
// for all of them instead of one per sample.
struct StepGroup {
  FoldStep s[kStepGroup];
};

template <class Steps>
__device__ __forceinline__ void load_step_group(const Steps fs, int i, int n, StepGroup& g) {
#pragma unroll
  for (int q = 0; q < kStepGroup; ++q) g.s[q] = fold_step(fs, min(i + q, n - 1));  // in bounds
}

// Fold of the chunk of samples i0 .. i0 + qn - 1 into `slab` by the nt folding threads (t = 0 .. nt - 1):
// slab[q][kk] = sum over slots of weight x dJ[node][kk].  fs: the ray's per-build fold steps (node offsets
// and weights by slot; workgroup-uniform, read with scalar loads, so are the branches).  KPT == 1 (the
// kernel with one entry per lookup lane, NLA <= nt = 256): the slots (f, f2) carry over from one of the
// thread's samples to the next and a slot is re-read only where its node changed; prev: the node offsets of
// the thread's previous sample; sg holds the steps of the chunk's first kStepGroup samples on entry and
// those of the next chunk's on return (n: samples of the ray).  KPT > 1: every sample reads its nodes.
template <int KPT, class Steps>
__device__ __forceinline__ void ray_scatter_fold(const float4* __restrict__ dJ, const Steps fs,
                                                 float4* slab, int i0, int qn, int NLA, int t, int nt, FoldSlots& f,
                                                 FoldSlots& f2, int4& prev, StepGroup& sg, int n) {
  if (KPT == 1) {
    // the four fold wavefronts split the ray's samples by parity, which halves the chain of dependent
    // node loads that each of them walks: wavefronts 0-1 fold the even samples, 2-3 the odd ones, each lane
    // two entries of its pair's 256 (f: the first, f2: the second), the slots carrying over from the
    // wavefront's previous sample, two samples back.  Lanes past the table read the last entry and store
    // the same bits as the last entry's lane to the same place (no branch around the store, which would
    // give the loop a path that skips the wait for the slot loads).
    const int w = t >> 6, lane = t & 63, par = w >> 1;
    const int kl0 = min((w & 1) * 128 + lane, NLA - 1), kl1 = min((w & 1) * 128 + 64 + lane, NLA - 1);
    // sg holds the steps of samples i0 .. i0 + 3 (requested before the barrier that precedes this chunk)
    for (int b = 0; b < qn; b += kStepGroup) {
      if (b > 0) load_step_group(fs, i0 + b, n, sg);
#pragma unroll
      for (int q = 0; q < kStepGroup; ++q) {
        if (b + q < qn && ((i0 + b + q) & 1) == par) {
          const FoldStep st = sg.s[q];
          const int4 o = st.off;
          if (o.x >= 0 && o.x != prev.x) {
            f.v[0] = dJ[o.x + kl0];
            f2.v[0] = dJ[o.x + kl1];
          }
          if (o.y >= 0 && o.y != prev.y) {
            f.v[1] = dJ[o.y + kl0];
            f2.v[1] = dJ[o.y + kl1];
          }
          if (o.z >= 0 && o.z != prev.z) {
            f.v[2] = dJ[o.z + kl0];
            f2.v[2] = dJ[o.z + kl1];
          }
          if (o.w >= 0 && o.w != prev.w) {
            f.v[3] = dJ[o.w + kl0];
            f2.v[3] = dJ[o.w + kl1];
          }
          prev = o;
          slab[(b + q) * NLA + kl0] = fold_value(f, st.wt);
          slab[(b + q) * NLA + kl1] = fold_value(f2, st.wt);
        }
      }
    }
    // the next chunk's first steps, in flight across the barrier
    if (i0 + qn < n) load_step_group(fs, i0 + qn, n, sg);
  } else {  // KPT (even) entries per thread, NLA <= KPT nt: every sample reads its four nodes
    for (int q = 0; q < qn; ++q) {
      const FoldStep st = fold_step(fs, i0 + q);
      const int4 o = st.off;
      const float4 w = st.wt;
      // an unused slot (offset -1, weight 0) reads node 0, which adds zero
      const float4* __restrict__ d0 = dJ + max(o.x, 0);
      const float4* __restrict__ d1 = dJ + max(o.y, 0);
      const float4* __restrict__ d2 = dJ + max(o.z, 0);
      const float4* __restrict__ d3 = dJ + max(o.w, 0);
      // two entries per round, eight loads in flight (entries past the table repeat the last one, which
      // stores the same bits to the same place)
#pragma unroll
      for (int e = 0; e < KPT; e += 2) {
        const int k0 = min(t + e * nt, NLA - 1), k1 = min(t + (e + 1) * nt, NLA - 1);
        const float4 a0 = d0[k0], b0 = d1[k0], c0 = d2[k0], e0 = d3[k0];
        const float4 a1 = d0[k1], b1 = d1[k1], c1 = d2[k1], e1 = d3[k1];
        FoldSlots f0, f1;
        f0.v[0] = a0;
        f0.v[1] = b0;
        f0.v[2] = c0;
        f0.v[3] = e0;
        f1.v[0] = a1;
        f1.v[1] = b1;
        f1.v[2] = c1;
        f1.v[3] = e1;
        slab[q * NLA + k0] = fold_value(f0, w);
        slab[q * NLA + k1] = fold_value(f1, w);
      }
    }
  }
}

__device__ float3 resample4_value(const DevParams& P, const float4* __restrict__ a, const float4* __restrict__ b,
                                 long long idx, const float4* __restrict__ c = nullptr);
__device__ float3 resample4_value_lean(const DevParams& P, const float4* __restrict__ a, long long idx);
__device__ float3 resample2_value(const DevParams& P, int space, const float4* __restrict__ a,
                                 const float4* __restrict__ b, long long idx);

// Warp-specialised: waves 0-3 fold, waves 4-7 look up.  The folded slabs are double-buffered: while the
// lookup waves read chunk c from one half, the fold waves fill the other half with chunk c + 1, and one
// barrier per chunk hands both halves over (chunk c + 1 overwrites the half that chunk c - 1's lookups
// read, finished before the previous barrier).
constexpr int kRSBlock = 512;
constexpr int kRSHalf = kRSBlock / 2;

// Eight waves per SIMD (four workgroups per CU, as the LDS allows): at most 64 VGPRs, the re-sampling
// of the accumulations spilling a few registers outside the loops.
template <int EPT, bool kB, bool kExact>
__global__ void __launch_bounds__(kRSBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) ray_scatter_kernel(
    const DevParams P0, const DevParams* __restrict__ PB, const float4* __restrict__ dJ,
    const double2* __restrict__ geo_th, const float4* __restrict__ geo_T, const float4* __restrict__ geo_rs,
    const FoldStep* __restrict__ geo_fold, float4* __restrict__ out, const RsAccum accum_arg, int Q,
    WorkCounters* wc) {
  const DevParams& P = params_of<kB>(P0, PB);
  if constexpr (kB) {  // this atmosphere's tables and ray geometry (the accumulation tables where they are used);
                       // offsets of the restrict arguments, whose loads stay provably unclobbered and uniform
    geo_th = atm_ptr<kB>(geo_th, P0);
    geo_T = atm_ptr<kB>(geo_T, P0);
    geo_rs = atm_ptr<kB>(geo_rs, P0);
    // the fold steps depend only on the planet and the grid, which the atmospheres of a batch share: all
    // read atmosphere 0's copy (one copy in the caches instead of one per atmosphere)
    dJ = atm_ptr<kB>(dJ, P);
    out = atm_ptr<kB>(out, P);
  }
  // dynamic LDS: 2 x Q x (NL x NA) slab entries (chunk c in half c % 2)
  extern __shared__ float4 s_slab[];
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int NLA = NL * NA;
  const int n = P.ray_steps;
  if ((int)blockIdx.x >= NH * NE) {  // the extra workgroup: E <- tab(dE + E) (atmosphere_lut.clj:95-96)
    const long long nE = (long long)P.nE[0] * P.nE[1];
    const float4* dE = atm_ptr<kB>(accum_arg.dE, P);
    const float4* E_in = atm_ptr<kB>(accum_arg.E_in, P);
    float4* E_out = atm_ptr<kB>(accum_arg.E_out, P);
    for (long long i = threadIdx.x; i < nE; i += blockDim.x) {
      const float3 v = resample2_value(P, 3, dE, E_in, i);
      E_out[i] = make_float4(v.x, v.y, v.z, 0.f);
    }
    return;
  }
  // XCD-aware, in ascending impact parameter: every sample of a straight ray has the same
  // |p| sqrt(1 - mu^2), so the (height, elevation) cells that a ray folds lie along one curve, and an XCD
  // given a contiguous range of impact parameters reads a band of the dJ table
  const int ord = xcd_ray(blockIdx.x, NH * NE);
  const int rayid = P.ray_order ? P.ray_order[ord] : ord;
  const bool folder = threadIdx.x < kRSHalf;  // whole wavefronts
  const int t = folder ? (int)threadIdx.x : (int)threadIdx.x - kRSHalf;
  // this ray's samples in the per-build geometry: T(x, p_i) rgb and (height coordinate, elevation
  // coordinate, alpha_i, beta_i)
  const float4* __restrict__ gT = geo_T + (long long)rayid * n;
  const float4* __restrict__ gRS = geo_rs + (long long)rayid * n;
  const int chunks = (n + Q - 1) / Q;
  // with one entry per lane the fold waves, idle during the last chunk, re-sample S for the epilogue
  // (0.5% per launch); with several entries per lane it is slower there and stays with the lookup waves
  constexpr bool kFoldResamples = EPT == 1;
  if (folder) {
    const FoldStep* __restrict__ gS = geo_fold + (long long)rayid * n;
    const auto gF = [&] {
      if constexpr (kB) return (ConstFoldSteps)gS;
      else return gS;
    }();
    FoldSlots fs, fs2;
    fold_slots_init(fs);
    fold_slots_init(fs2);
    int4 prev = make_int4(-1, -1, -1, -1);
    StepGroup sg;
    if constexpr (EPT == 1) load_step_group(gF, 0, n, sg);
    ray_scatter_fold<EPT>(dJ, gF, s_slab, 0, min(Q, n), NLA, t, kRSHalf, fs, fs2, prev, sg, n);
    __syncthreads();
    for (int c = 0; c + 1 < chunks; ++c) {
      ray_scatter_fold<EPT>(dJ, gF, s_slab + ((c + 1) & 1) * Q * NLA, (c + 1) * Q, min(Q, n - (c + 1) * Q), NLA,
                            t, kRSHalf, fs, fs2, prev, sg, n);
      __syncthreads();
    }
    if (kFoldResamples && accum_arg.S_out) {
      // while the last chunk's lookups run: the running sum S of atmosphere_lut.clj:95-97 at this ray's grid
      // points into the half that the lookups of the chunk before the last have finished with, for the
      // lookup waves to add their dS to after the barrier (the fold slots are dead by now)
      const float4* S_in = atm_ptr<kB>(accum_arg.S_in, P);
      float4* su = s_slab + (chunks & 1) * Q * NLA;
      const long long base = (long long)rayid * NLA;
      for (int kk = t; kk < NLA; kk += kRSHalf) {
        const float3 u = accum_arg.S_resample ? resample4_value_lean(P, S_in, base + kk) : ld3(S_in, base + kk);
        su[kk] = make_float4(u.x, u.y, u.z, 0.f);
      }
    }
    __syncthreads();
    return;
  }
  // the ray's step D / n: twice its first sample's t = 0.5 D / n (FP64, exact), from the per-build geometry
  const double dt = 2.0 * geo_th[(long long)rayid * n].x;
  const SunCoord sc = {(float)((NL - 1) * (double)P.sun_inv_den), (float)(-0.6 * 1.4426950408889634), (float)(NL - 1)};
  const float row_bytes = (float)(NA * (int)sizeof(float4));
  // per-entry constants: heading coordinate and the terms of dot(p_i, l)
  SunTerms st[EPT];
  Lerp La[EPT];
  float3 acc[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    int k = t + q * kRSHalf;
    acc[q] = make_float3(0.f, 0.f, 0.f);
    st[q].vl = 0.f;
    st[q].xl = 0.f;
    st[q].vl64 = 0.0;
    st[q].xl64 = 0.0;
    st[q].R = P.g.R;
    st[q].A64 = kExact ? (NL - 1) / (1.0 - exp(-3.6)) : 0.0;
    La[q] = make_lerp(0.f, NA);
    if (k < NLA) {
      // the per-build sun terms (v . l, x . l, the heading coordinate, rounded once from FP64)
      const float4 m = P.sun_rs[(long long)rayid * NLA + k];
      st[q].vl = m.x;
      st[q].xl = m.y;
      La[q] = make_lerp(m.z, NA);
      if (kExact) {  // the same dot products in FP64
        const double2 d = P.sun_ss[(long long)rayid * NLA + k];
        st[q].vl64 = d.x;
        st[q].xl64 = d.y;
      }
    }
  }
  SampleGroup sgp;
  if constexpr (EPT == 1) load_sample_group<kExact>(gT, gRS, geo_th + (long long)rayid * n, 0, n, sgp);
  __syncthreads();  // chunk 0 folded
  for (int c = 0; c < chunks; ++c) {
    const int i0 = c * Q, qn = min(Q, n - i0);
    const float4* slab = s_slab + (c & 1) * Q * NLA;
    if constexpr (EPT == 1) {
      for (int b = 0; b < qn; b += kStepGroup) {
        if (b > 0) load_sample_group<kExact>(gT, gRS, geo_th + (long long)rayid * n, i0 + b, n, sgp);
        if (t < NLA) {
          const float col_bytes = (float)(La[0].i0 * (int)sizeof(float4));
          // a heading coordinate with zero weight (every entry but the clamped ones) reads one column;
          // the choice is made per wavefront so that a wavefront never runs both loops
          if (__all(La[0].s == 0.0f))
            acc[0] = ray_scatter_sample_group<true, kExact>(sc, slab, sgp, b, qn, row_bytes, col_bytes, NLA, st[0],
                                                            La[0], acc[0]);
          else
            acc[0] = ray_scatter_sample_group<false, kExact>(sc, slab, sgp, b, qn, row_bytes, col_bytes, NLA, st[0],
                                                             La[0], acc[0]);
        }
      }
      if (c + 1 < chunks)  // in flight across the barrier
        load_sample_group<kExact>(gT, gRS, geo_th + (long long)rayid * n, i0 + qn, n, sgp);
    } else {  // several entries per lane: the terms per sample (a group would spill)
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        if (t + e * kRSHalf < NLA) {
          const float col_bytes = (float)(La[e].i0 * (int)sizeof(float4));
          if (__all(La[e].s == 0.0f))
            acc[e] = ray_scatter_samples<true, kExact>(sc, slab, gT, gRS, geo_th + (long long)rayid * n, i0, qn,
                                                       row_bytes, col_bytes, NLA, st[e], La[e], acc[e]);
          else
            acc[e] = ray_scatter_samples<false, kExact>(sc, slab, gT, gRS, geo_th + (long long)rayid * n, i0, qn,
                                                        row_bytes, col_bytes, NLA, st[e], La[e], acc[e]);
        }
      }
    }
    __syncthreads();
  }
  const long long base = (long long)rayid * NLA;
  const float a = (float)dt;
  unsigned long long lookups = 0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = t + e * kRSHalf;
    if (k < NLA) {
      const float4 d = make_float4(acc[e].x * a, acc[e].y * a, acc[e].z * a, 0.f);
      out[base + k] = d;
      if (accum_arg.S_out) {  // S + dS at this grid point
        float3 u;
        if (kFoldResamples) {
          const float4 v = s_slab[(chunks & 1) * Q * NLA + k];
          u = make_float3(v.x, v.y, v.z);
        } else {
          const float4* S_in = atm_ptr<kB>(accum_arg.S_in, P);
          u = accum_arg.S_resample ? resample4_value_lean(P, S_in, base + k) : ld3(S_in, base + k);
        }
        atm_ptr<kB>(accum_arg.S_out, P)[base + k] = make_float4(u.x + d.x, u.y + d.y, u.z + d.z, 0.f);
      }
      lookups += n;
    }
  }
  count(wc, 0, 0, lookups, 0);
}

// dJ = point-scatter over the sphere quadrature (atmosphere_lut.clj:88,91; atmosphere.clj:203-222).
// Every grid point of a height index shares its position and quadrature frame, so the per-direction
// geometry (ray extremity P, surface flag, T(x, P), elevation coordinate of omega at x) depends only
// on (height, direction) and is computed once per build (direction_geometry_kernel).  The dS lookup
// at (x, omega, l) has height and elevation coordinates that depend only on (height, direction) as
// well: point_scatter_combine_kernel folds them into a (sun elevation, heading) slab per
// (height, direction) once per iteration, shared by every view ray of that height.
struct DirGeom {
  float4 omega;   // direction xyz, w: elevation coordinate of omega at x (flag: not surface)
  float4 ground;  // T(x, P) * brightness / pi (0 for sky directions), w: 1 for surface directions
  float4 phat;    // P / |P| (w unused: P is on the ground, height row 0 of the surface-radiance table)
};

__device__ __forceinline__ void direction_geometry_body(const DevParams& P, const Dir* __restrict__ dirs,
                                                                    int ndirs, DirGeom* __restrict__ geom,
                                                                    WorkCounters* wc, unsigned blk) {
  const int NH = P.nS[0], NE = P.nS[1];
  const int n = P.ray_steps;
  int idx = blk * blockDim.x + threadIdx.x;
  bool active = idx < NH * ndirs;
  int ih = active ? idx / ndirs : 0, j = active ? idx - ih * ndirs : 0;
  const double r0 = index_to_radius(P.g, NH, ih), h0 = index_to_altitude(P.g, NH, ih);
  Dir d = dirs[j];
  double pd = r0 * d.x;  // dot(x, omega)
  bool above = is_above_horizon(P.g, h0, pd);
  double D = above ? top_exit_distance(P.g, h0, pd) : ground_distance(P.g, h0, pd);
  double hP = altitude_along(P.g, h0, pd, D);
  bool surface = 2.0 * hP < P.g.H;  // atmosphere.clj:88-92
  // every lane evaluates the transmittance (uniform control flow for the wavefront vote inside);
  // sky directions discard it
  float3 T = segment_transmittance(P, h0, pd, D, n);
  if (!active) return;
  DirGeom g;
  g.omega = make_float4((float)d.x, (float)d.y, (float)d.z, (float)elevation_to_index(P.g, NE, r0, pd, !surface));
  if (surface) {
    g.ground = make_float4(T.x * P.brightness_over_pi[0], T.y * P.brightness_over_pi[1], T.z * P.brightness_over_pi[2],
                           1.f);
    double px = r0 + d.x * D, py = d.y * D, pz = d.z * D;
    double pn = 1.0 / sqrt(px * px + py * py + pz * pz);
    g.phat = make_float4((float)(px * pn), (float)(py * pn), (float)(pz * pn), 0.f);
  } else {
    g.ground = make_float4(0.f, 0.f, 0.f, 0.f);
    g.phat = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  geom[idx] = g;
  count(wc, 0, surface ? n : 0, 0, 0);
}

__global__ void __launch_bounds__(kBlock) direction_geometry_kernel(const DevParams P, const Dir* __restrict__ dirs,
                                                                    int ndirs, DirGeom* __restrict__ geom,
                                                                    WorkCounters* wc) {
  direction_geometry_body(P, dirs, ndirs, geom, wc, blockIdx.x);
}

// slab[h][j][s][a] = dS at (height of x_h, elevation coordinate of omega_j) for every (s, a) grid
// point; with mie != nullptr the Mie strength slab is folded the same way into slab_mie.
template <bool kB>
__global__ void __launch_bounds__(kBlock) point_scatter_combine_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                                       const DirGeom* __restrict__ geom, int ndirs,
                                                                       const float4* __restrict__ dS,
                                                                       const float4* __restrict__ mie,
                                                                       float4* __restrict__ slab,
                                                                       float4* __restrict__ slab_mie) {
  const DevParams& P = params_of<kB>(P0, PB);
  geom = atm_ptr<kB>(geom, P);
  dS = atm_ptr<kB>(dS, P);
  mie = atm_ptr<kB>(mie, P);
  slab = atm_ptr<kB>(slab, P);
  slab_mie = atm_ptr<kB>(slab_mie, P);
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int NLA = NL * NA;
  const int b = xcd_ray(blockIdx.x, NH * ndirs);  // the XCD that reads this height's slabs (same ranges)
  const int ih = b / ndirs;
  const double r0 = index_to_radius(P.g, NH, ih);
  const Lerp Lh = make_lerp((float)height_to_index(P.g, NH, r0), NH);
  const Lerp Le = make_lerp(geom[b].omega.w, NE);
  const long long sh = (long long)NE * NLA;
  const long long o00 = Lh.i0 * sh + (long long)Le.i0 * NLA, o01 = Lh.i0 * sh + (long long)Le.i1 * NLA;
  const long long o10 = Lh.i1 * sh + (long long)Le.i0 * NLA, o11 = Lh.i1 * sh + (long long)Le.i1 * NLA;
  float4* dst = slab + (long long)b * NLA;
  float4* dstm = slab_mie ? slab_mie + (long long)b * NLA : nullptr;
  for (int k = threadIdx.x; k < NLA; k += blockDim.x) {
    float3 c = mix3(mix3(ld3(dS, o00 + k), ld3(dS, o01 + k), Le.s), mix3(ld3(dS, o10 + k), ld3(dS, o11 + k), Le.s), Lh.s);
    dst[k] = make_float4(c.x, c.y, c.z, 0.f);
    if (mie) {
      float3 m = mix3(mix3(ld3(mie, o00 + k), ld3(mie, o01 + k), Le.s), mix3(ld3(mie, o10 + k), ld3(mie, o11 + k), Le.s),
                      Lh.s);
      dstm[k] = make_float4(m.x, m.y, m.z, 0.f);
    }
  }
}

// One workgroup per view ray (x, v): quadrature weights sum_c beta_c(h(x)) phase_c(v . omega) w_j
// and the direction geometry in LDS, lanes over (sun elevation, heading).
struct DirLds {
  float4 omega;   // direction xyz, w: 1 for surface directions
  float4 weight;  // rgb quadrature weight times in-scattering
  float4 ground;  // T(x, P) brightness / pi
  float4 phat;    // P / |P|
};

// Constants of the per-direction lookups of one point-scatter launch.
struct PsConst {
  float hA;            // (NA - 1) / 2: heading coordinate (NA - 1) (1 + dot(omega, l)) / 2 = hA + hA dot(omega, l)
  float na_top;        // NA - 1
  SunCoord se;         // sun-elevation coordinate of dE (size nE[1]): alpha = -3 log2(e), beta = 0
  float mie_k, mie_a, mie_b;  // Mie phase K (1 + mu^2) (a - b mu)^(-3/2): K = 3 (1 - g^2) / (8 pi (2 + g^2))
};

// phase of atmosphere.clj:56-61 with the hardware reciprocal square root (x^-3/2 = rsq(x)^3)
__device__ __forceinline__ float mie_phase(const PsConst& pc, float mu) {
  const float y = __builtin_amdgcn_rsqf(fmaf(-pc.mie_b, mu, pc.mie_a));
  return pc.mie_k * fmaf(mu, mu, 1.0f) * (y * y * y);
}

// the texel at byte_offset and the next one (the pointer is formed once, so the second read takes the
// first one's address register with an immediate offset)
__device__ __forceinline__ void ld4b_pair(const float4* base, unsigned byte_offset, float4& c0, float4& c1) {
  const float4* p = (const float4*)((const char*)base + byte_offset);
  c0 = p[0];
  c1 = p[1];
}

__device__ __forceinline__ float3 lerp_pair(const float4* base, unsigned byte_offset, float s) {
  float4 c0, c1;
  ld4b_pair(base, byte_offset, c0, c1);
  return lerp3(c0, c1, s);
}

// Sum over the quadrature directions of one entry (sun direction l, sun-elevation lerp Ls; row_bytes:
// byte offset of row Ls.i0 in a slab).  Heading lerp: lower column min(floor, NA - 2), so the upper
// column is always the next one; the dE lookup likewise in both dimensions (nested as lookup2).
// kRow: the entry's sun-elevation weight is zero (one slab row)
template <bool kRow, bool kMie>
__device__ __forceinline__ float3 point_scatter_dirs(const PsConst& pc, const DirLds* s_dir, int ndirs,
                                                     const float4* __restrict__ sh, const float4* __restrict__ shm,
                                                     const float4* __restrict__ dE, int NA, int NLA, Lerp Ls,
                                                     float row_bytes, float lx, float ly, float lz) {
  float3 acc = make_float3(0.f, 0.f, 0.f);
  const unsigned d_row = (unsigned)(Ls.i1 - Ls.i0) * (unsigned)NA * (unsigned)sizeof(float4);
  // direction j's heading coordinate: cosine ol, column weight sa, byte offset of the lower column
  auto heading = [&](int j, float& ol, float& sa) {
    const DirLds& dd = s_dir[j];
    ol = fmaf(dd.omega.z, lz, fmaf(dd.omega.y, ly, dd.omega.x * lx));
    const float hc = __builtin_amdgcn_fmed3f(fmaf(pc.hA, ol, pc.hA), 0.0f, pc.na_top);
    const float col = lower_cell(hc, pc.na_top - 1.0f);
    sa = hc - col;
    return (unsigned)fmaf(col, (float)sizeof(float4), row_bytes);
  };
  // the rest of direction j's term given the two texels c0, c1 of its lower slab row
  auto term = [&](int j, float ol, float sa, unsigned off, float4 c0, float4 c1) {
    const DirLds& dd = s_dir[j];
    const float4* row = sh + j * NLA;
    float3 v = lerp3(c0, c1, sa);
    if (!kRow) {
      const float3 u = lerp_pair(row, off + d_row, sa);
      v = make_float3(fmaf(Ls.s, u.x - v.x, v.x), fmaf(Ls.s, u.y - v.y, v.y), fmaf(Ls.s, u.z - v.z, v.z));
    }
    if (kMie) {
      const float4* rowm = shm + j * NLA;
      float3 m = lerp_pair(rowm, off, sa);
      if (!kRow) {
        const float3 u = lerp_pair(rowm, off + d_row, sa);
        m = make_float3(fmaf(Ls.s, u.x - m.x, m.x), fmaf(Ls.s, u.y - m.y, m.y), fmaf(Ls.s, u.z - m.z, m.z));
      }
      const float ph = mie_phase(pc, ol);
      v.x = fmaf(m.x, ph, v.x);
      v.y = fmaf(m.y, ph, v.y);
      v.z = fmaf(m.z, ph, v.z);
    }
    if (dd.omega.w != 0.f) {
      const float spl = fmaf(dd.phat.z, lz, fmaf(dd.phat.y, ly, dd.phat.x * lx));
      const float ce = __builtin_amdgcn_fmed3f(
          fmaf(-pc.se.A, fexp2(fmaf(-3.0f * 1.4426950408889634f, spl, pc.se.c)), pc.se.A), 0.0f, pc.se.top);
      const float ecol = lower_cell(ce, pc.se.top - 1.0f);
      const float se = ce - ecol;
      const unsigned eo = (unsigned)ecol * (unsigned)sizeof(float4);  // height row 0: P is on the ground
      const float3 u = lerp_pair(dE, eo, se);
      v.x = fmaf(dd.ground.x, u.x, v.x);
      v.y = fmaf(dd.ground.y, u.y, v.y);
      v.z = fmaf(dd.ground.z, u.z, v.z);
    }
    acc.x = fmaf(dd.weight.x, v.x, acc.x);
    acc.y = fmaf(dd.weight.y, v.y, acc.y);
    acc.z = fmaf(dd.weight.z, v.z, acc.z);
  };
  // four directions at a time: their eight slab reads are issued before any of their ground terms
  constexpr int kB = 4;
  int j = 0;
  for (; j + kB <= ndirs; j += kB) {
    float ol[kB], sa[kB];
    unsigned off[kB];
    float4 c0[kB], c1[kB];
#pragma unroll
    for (int d = 0; d < kB; ++d) {
      off[d] = heading(j + d, ol[d], sa[d]);
      const float4* row = sh + (j + d) * NLA;
      ld4b_pair(row, off[d], c0[d], c1[d]);
    }
#pragma unroll
    for (int d = 0; d < kB; ++d) term(j + d, ol[d], sa[d], off[d], c0[d], c1[d]);
  }
  for (; j < ndirs; ++j) {
    float ol, sa;
    const unsigned off = heading(j, ol, sa);
    const float4* row = sh + j * NLA;
    float4 c0, c1;
    ld4b_pair(row, off, c0, c1);
    term(j, ol, sa, off, c0, c1);
  }
  return acc;
}

// kELds: the previous order's dE at the ground (height row 0) is staged in LDS after the direction
// table (when it fits 16 KiB).  P, the ground point of a surface direction, is the ray's intersection
// with the planet's sphere (atmosphere.clj:80-85, 213-221), so its height is 0: its FP64 |P| - R is
// rounding noise of ~1e-9 m, whose height coordinate (through the square root of the horizon distance,
// ~1.5e-6 for 16 rows and 100 km) would weight height row 1 by as much.
template <bool kFold>
__device__ __forceinline__ void surface_radiance_entry(const DevParams& P, int ndirs, const float4* __restrict__ dS,
                                                       const float4* __restrict__ mie, float4* __restrict__ out,
                                                       WorkCounters* wc, int entry, float4* s_fold, float3* s_red);

// The iteration's surface radiance (it reads the same dS and Mie strength) rides in the launch: the
// sr.count workgroups after the rays' compute its entries with the LDS-folded kernel body.
template <bool kMie, bool kELds, bool kB>
__global__ void __launch_bounds__(kBlock) point_scatter_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                               const Dir* __restrict__ dirs,
                                                               const DirGeom* __restrict__ geom, int ndirs,
                                                               const float4* __restrict__ slab,
                                                               const float4* __restrict__ slab_mie,
                                                               const float4* __restrict__ dE, float4* __restrict__ out,
                                                               const SrJob sr, WorkCounters* wc) {
  extern __shared__ DirLds s_dir[];
  float4* s_dE = (float4*)(s_dir + ndirs);
  const DevParams& P = params_of<kB>(P0, PB);
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int NLA = NL * NA;
  if ((int)blockIdx.x >= NH * NE) {
    __shared__ float3 s_red[kBlock / 64];
    surface_radiance_entry<true>(P, sr.ndirs, atm_ptr<kB>(sr.dS, P), atm_ptr<kB>(sr.mie, P), atm_ptr<kB>(sr.out, P),
                                 wc, (int)blockIdx.x - NH * NE, (float4*)s_dir, s_red);
    return;
  }
  geom = atm_ptr<kB>(geom, P);
  slab = atm_ptr<kB>(slab, P);
  slab_mie = atm_ptr<kB>(slab_mie, P);
  dE = atm_ptr<kB>(dE, P);
  out = atm_ptr<kB>(out, P);
  const int rayid = xcd_ray(blockIdx.x, NH * NE);
  if (kELds)
    for (int i = threadIdx.x; i < P.nE[1]; i += blockDim.x) s_dE[i] = dE[i];
  const int ih = rayid / NE, ie = rayid - ih * NE;
  const double r0 = index_to_radius(P.g, NH, ih), h0 = index_to_altitude(P.g, NH, ih);
  double se, ce;
  bool above_unused;
  index_to_elevation(P.g, NE, r0, ie, se, ce, above_unused);
  const float bm = (float)exp(-h0 * P.inv_scale_mie), br = (float)exp(-h0 * P.inv_scale_ray);
  const DirGeom* gh = geom + (long long)ih * ndirs;
  const int nE1 = P.nE[1];
  for (int j = threadIdx.x; j < ndirs; j += blockDim.x) {
    Dir d = dirs[j];
    DirGeom g = gh[j];
    float vw = (float)(se * d.x + ce * d.y);
    float pm = phase_f(P.g_mie, vw) * bm, pr = phase_f(P.g_ray, vw) * br;
    float w = (float)d.w;
    DirLds dl;
    dl.omega = make_float4(g.omega.x, g.omega.y, g.omega.z, g.ground.w);
    dl.weight = make_float4((P.sca_mie[0] * pm + P.sca_ray[0] * pr) * w, (P.sca_mie[1] * pm + P.sca_ray[1] * pr) * w,
                            (P.sca_mie[2] * pm + P.sca_ray[2] * pr) * w, 0.f);
    dl.ground = make_float4(g.ground.x, g.ground.y, g.ground.z, 0.f);
    dl.phat = make_float4(g.phat.x, g.phat.y, g.phat.z, 0.f);
    s_dir[j] = dl;
  }
  const float gm = P.g_mie;
  PsConst pc;
  pc.hA = 0.5f * (float)(NA - 1);
  pc.na_top = (float)(NA - 1);
  pc.se = SunCoord{(float)((nE1 - 1) * (double)P.sun_inv_den), (float)(-0.6 * 1.4426950408889634), (float)(nE1 - 1)};
  pc.mie_k = 3.0f * (1.0f - gm * gm) / (8.0f * 3.14159265358979f * (2.0f + gm * gm));
  pc.mie_a = 1.0f + gm * gm;
  pc.mie_b = 2.0f * gm;
  __syncthreads();
  const float4* eTab = kELds ? s_dE : dE;
  const float4* sh = slab + (long long)ih * ndirs * NLA;
  const float4* shm = kMie ? slab_mie + (long long)ih * ndirs * NLA : nullptr;
  const long long base = (long long)rayid * NLA;
  unsigned long long l4 = 0, l2 = 0;
  for (int k = threadIdx.x; k < NLA; k += blockDim.x) {
    float lx, ly, lz;
    Lerp Ls;
    {  // the per-build sun terms (l and the sun-elevation coordinate, rounded once from FP64)
      const float4 m = P.sun_ps[base + k];
      lx = m.x;
      ly = m.y;
      lz = m.z;
      Ls = make_lerp(m.w, NL);
    }
    const float row_bytes = (float)(Ls.i0 * NA * (int)sizeof(float4));
    // a grid entry's own sun-elevation coordinate is an integer, so its weight is exactly zero and only
    // one row of each slab is read; the general lookup stays for safety
    const float3 acc = Ls.s == 0.0f
                           ? point_scatter_dirs<true, kMie>(pc, s_dir, ndirs, sh, shm, eTab, NA, NLA, Ls, row_bytes, lx, ly, lz)
                           : point_scatter_dirs<false, kMie>(pc, s_dir, ndirs, sh, shm, eTab, NA, NLA, Ls, row_bytes, lx, ly, lz);
    l4 += (unsigned long long)ndirs * (kMie ? 2 : 1);
    if (wc)
      for (int j = 0; j < ndirs; ++j) l2 += s_dir[j].omega.w != 0.f;
    out[base + k] = make_float4(acc.x, acc.y, acc.z, 0.f);
  }
  count(wc, 0, 0, l4, l2);
}

// dE' = surface-radiance of dS (atmosphere_lut.clj:89,93; atmosphere.clj:225-230): one workgroup per
// entry of the surface-radiance table, lanes over half-sphere directions, workgroup reduction.
// The half-sphere directions' geometry at every surface-radiance height (once per build).
__device__ __forceinline__ void surface_geometry_body(const DevParams& P, const Dir* __restrict__ dirs,
                                                                  int ndirs, unsigned blk) {
  const int idx = blk * blockDim.x + threadIdx.x;
  if (idx >= P.nE[0] * ndirs) return;
  const int ih = idx / ndirs, j = idx - ih * ndirs;
  const double r0 = index_to_radius(P.g, P.nE[0], ih);
  const Dir d = dirs[j];
  P.sgeo[idx] = make_float4((float)d.x, (float)d.y, (float)elevation_to_index(P.g, P.nS[1], r0, r0 * d.x, true),
                            (float)(d.x * d.w));  // (omega . normal) * quadrature weight, normal = (1 0 0)
}

__global__ void __launch_bounds__(kBlock) surface_geometry_kernel(const DevParams P, const Dir* __restrict__ dirs,
                                                                  int ndirs) {
  surface_geometry_body(P, dirs, ndirs, blockIdx.x);
}

// kFold: the entry's height and sun-elevation coordinates are the same for every direction, so the
// workgroup first folds those two dimensions of dS (and of the Mie strength) into an (elevation,
// heading) slab in LDS, NE x NA float4, and each direction interpolates the other two in it.
// One workgroup computes entry `entry` of the table; s_fold: kFold's LDS slab, s_red: one float3 per
// wavefront.
template <bool kFold>
__device__ __forceinline__ void surface_radiance_entry(const DevParams& P, int ndirs, const float4* __restrict__ dS,
                                                       const float4* __restrict__ mie, float4* __restrict__ out,
                                                       WorkCounters* wc, int entry, float4* s_fold, float3* s_red) {
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const int ih = entry / P.nE[1], il = entry - ih * P.nE[1];
  const double r0 = index_to_radius(P.g, P.nE[0], ih);
  const double s = index_to_sin_sun_elevation(P.nE[1], il);
  const double c = sqrt(fmax(0.0, 1 - s * s));
  const float lx = (float)s, ly = (float)c;
  const Lerp Lh = make_lerp((float)height_to_index(P.g, NH, r0), NH);
  const Lerp Ls = make_lerp((float)sun_elevation_to_index(NL, (r0 * s) / r0), NL);
  const float gm = P.g_mie;
  const float mk = 3.0f * (1.0f - gm * gm) / (8.0f * 3.14159265358979f * (2.0f + gm * gm)), ma = 1.0f + gm * gm,
              mb = 2.0f * gm;
  const float4* __restrict__ geo = P.sgeo + ih * ndirs;
  const int NEA = NE * NA;
  float4* s_m = s_fold + NEA;
  if (kFold) {
    const long long h0 = (long long)Lh.i0 * NE * NL * NA, h1 = (long long)Lh.i1 * NE * NL * NA;
    const int s0 = Ls.i0 * NA, s1 = Ls.i1 * NA;
    for (int c = threadIdx.x; c < NEA; c += blockDim.x) {
      const int e = c / NA, a = c - e * NA;
      const long long eo = (long long)e * NL * NA + a;
      const float3 u = mix3(mix3(ld3(dS, h0 + eo + s0), ld3(dS, h0 + eo + s1), Ls.s),
                            mix3(ld3(dS, h1 + eo + s0), ld3(dS, h1 + eo + s1), Ls.s), Lh.s);
      s_fold[c] = make_float4(u.x, u.y, u.z, 0.f);
      if (mie) {
        const float3 m = mix3(mix3(ld3(mie, h0 + eo + s0), ld3(mie, h0 + eo + s1), Ls.s),
                              mix3(ld3(mie, h1 + eo + s0), ld3(mie, h1 + eo + s1), Ls.s), Lh.s);
        s_m[c] = make_float4(m.x, m.y, m.z, 0.f);
      }
    }
    __syncthreads();
  }
  float3 acc = make_float3(0.f, 0.f, 0.f);
  for (int j = threadIdx.x; j < ndirs; j += blockDim.x) {
    const float4 g = geo[j];
    const float ol = fmaf(g.x, lx, g.y * ly);
    const Lerp Le = make_lerp(g.z, NE);
    const Lerp La = make_lerp((NA - 1) * ((1.0f + ol) * 0.5f), NA);
    float3 v = kFold ? lookup2(s_fold, NA, Le, La) : lookup4(dS, NE, NL, NA, Lh, Le, Ls, La);
    if (mie) {
      const float3 m = kFold ? lookup2(s_m, NA, Le, La) : lookup4(mie, NE, NL, NA, Lh, Le, Ls, La);
      const float y = __builtin_amdgcn_rsqf(fmaf(-mb, ol, ma));  // phase (atmosphere.clj:56-61), x^-3/2 = rsq(x)^3
      const float ph = mk * fmaf(ol, ol, 1.0f) * (y * y * y);
      v.x = fmaf(m.x, ph, v.x);
      v.y = fmaf(m.y, ph, v.y);
      v.z = fmaf(m.z, ph, v.z);
    }
    acc.x = fmaf(v.x, g.w, acc.x);
    acc.y = fmaf(v.y, g.w, acc.y);
    acc.z = fmaf(v.z, g.w, acc.z);
  }
  // wavefront reduction, then across the workgroup's wavefronts
  for (int off = 32; off > 0; off >>= 1) {
    acc.x += __shfl_down(acc.x, off, 64);
    acc.y += __shfl_down(acc.y, off, 64);
    acc.z += __shfl_down(acc.z, off, 64);
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float3 t = s_red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      t.x += s_red[w].x; t.y += s_red[w].y; t.z += s_red[w].z;
    }
    out[entry] = make_float4(t.x, t.y, t.z, 0.f);
  }
  if (threadIdx.x == 0) count(wc, 0, 0, (unsigned long long)ndirs * (mie ? 2 : 1), 0);
}

template <bool kFold, bool kB>
__global__ void __launch_bounds__(kBlock) surface_radiance_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                                  int ndirs, const float4* __restrict__ dS,
                                                                  const float4* __restrict__ mie,
                                                                  float4* __restrict__ out, WorkCounters* wc) {
  __shared__ float3 s_red[kBlock / 64];
  extern __shared__ float4 s_fold[];  // kFold: [NE x NA] dS, then [NE x NA] Mie strength
  const DevParams& P = params_of<kB>(P0, PB);
  surface_radiance_entry<kFold>(P, ndirs, atm_ptr<kB>(dS, P), atm_ptr<kB>(mie, P), atm_ptr<kB>(out, P), wc,
                                blockIdx.x, s_fold, s_red);
}

// f(backward(i)) re-sampling of interpolated tables (interpolate.clj:54-72 with interpolation-table).
// Lerps of forward(backward(i)) for 4D entry i = ((ih NE + ie) NL + is) NA + ia (atmosphere.clj:233-422).
// With ps / rs (nullable) also the entry's sun terms that the point- and ray-scatter kernels use
// (DevParams::sun_ps, sun_rs), from the same FP64 values.
__device__ __forceinline__ void resample4_lerps(const DevParams& P, int ih, int ie, int is, int ia, Lerp& Lh, Lerp& Le,
                                                Lerp& Ls, Lerp& La, float4* ps = nullptr, float4* rs = nullptr,
                                                double2* ss = nullptr) {
  SFSIM_EXACT  // the reference's rounding (see geometry.hpp)
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  double r0 = index_to_radius(P.g, NH, ih);
  double se, ce;
  bool above;
  index_to_elevation(P.g, NE, r0, ie, se, ce, above);
  double sin_s = index_to_sin_sun_elevation(NL, is);
  double l0, l1, l2;
  index_to_sun_direction(NA, se, ce, sin_s, ia, l0, l1, l2);
  const double cs = sun_elevation_to_index(NL, (r0 * l0) / r0), vl = se * l0 + ce * l1;
  const double ca = sun_angle_to_index(NA, vl);
  Lh = make_lerp((float)height_to_index(P.g, NH, r0), NH);
  Le = make_lerp((float)elevation_to_index(P.g, NE, r0, r0 * se, above), NE);
  Ls = make_lerp((float)cs, NL);
  La = make_lerp((float)ca, NA);
  if (ps) *ps = make_float4((float)l0, (float)l1, (float)l2, (float)cs);
  if (rs) *rs = make_float4((float)vl, (float)(r0 * l0), (float)ca, 0.f);
  if (ss) *ss = make_double2(vl, r0 * l0);
}

__device__ __forceinline__ Lerp lerp_of(int i0, float s, int N) {
  Lerp L;
  L.i0 = i0;
  L.i1 = min(i0 + 1, N - 1);
  L.s = s;
  return L;
}

// The 4D re-sampling lookups of a build, once (every accumulation and final table reads them).
__device__ __forceinline__ void resample_map_body(const DevParams& P, unsigned blk) {
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const long long n = (long long)NH * NE * NL * NA;
  const long long idx = (long long)blk * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int ia = (int)(idx % NA);
  long long q = idx / NA;
  const int is = (int)(q % NL);
  q /= NL;
  const int ie = (int)(q % NE), ih = (int)(q / NE);
  Lerp Lh, Le, Ls, La;
  float4 ps, rs;
  double2 ss;
  resample4_lerps(P, ih, ie, is, ia, Lh, Le, Ls, La, &ps, &rs, &ss);
  ((float4*)P.rmap)[idx] = make_float4(Lh.s, Le.s, Ls.s, La.s);
  if (P.sun_ps) {
    P.sun_ps[idx] = ps;
    P.sun_rs[idx] = rs;
    P.sun_ss[idx] = ss;
  }
  ((ushort4*)(P.rmap + n * sizeof(float4)))[idx] =
      make_ushort4((unsigned short)Lh.i0, (unsigned short)Le.i0, (unsigned short)Ls.i0, (unsigned short)La.i0);
}

__global__ void __launch_bounds__(kBlock) resample_map_kernel(const DevParams P) {
  resample_map_body(P, blockIdx.x);
}

// The per-build jobs in one launch, job k over workgroups first[k] .. first[k + 1] - 1, the small
// long-running grids first so that they start at once: direction geometry, transmittance, E0, surface
// geometry, the sun table, ray geometry, the re-sampling map.  An absent job has an empty range.
struct PrologueArgs {
  PrologueJobs j;
  unsigned first[8];
};

// Batched (kB): each atmosphere's own jobs in its z slice; the shared geometry (surface geometry and the
// re-sampling and sun maps, which depend only on the planet and the grid) once, in slice 0.
template <bool kB>
__global__ void __launch_bounds__(kBlock) prologue_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                          const PrologueArgs A) {
  const DevParams& P = params_of<kB>(P0, PB);
  const unsigned b = blockIdx.x;
  const bool shared = !kB || blockIdx.z == 0;
  if (b < A.first[1])
    direction_geometry_body(P, A.j.sdirs, A.j.sphere_n, atm_ptr<kB>((DirGeom*)A.j.geom, P), A.j.wc, b - A.first[0]);
  else if (b < A.first[2])
    transmittance_body(P, atm_ptr<kB>(A.j.T, P), A.j.wc, b - A.first[1]);
  else if (b < A.first[3])
    surface_radiance_base_body(P, atm_ptr<kB>(A.j.E0, P), A.j.wc, b - A.first[2]);
  else if (b < A.first[4]) {
    if (shared) surface_geometry_body(P, A.j.hdirs, A.j.half_n, b - A.first[3]);
  } else if (b < A.first[5])
    sun_depth_body(P, b - A.first[4]);
  else if (b < A.first[6])
    ray_geometry_body(P, A.j.wc, b - A.first[5]);
  else if (shared)
    resample_map_body(P, b - A.first[6]);
}

// Value at 4D entry idx of tab(a + b) (b nullable), or of tab(a + (b - c)) with c (each texel's
// difference formed first): the multilinear lookup at forward(backward(idx)), from the per-build map
// when there is one.
__device__ float3 resample4_value(const DevParams& P, const float4* __restrict__ a, const float4* __restrict__ b,
                                 long long idx, const float4* __restrict__ c) {
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const long long n = (long long)NH * NE * NL * NA;
  Lerp Lh, Le, Ls, La;
  if (P.rmap) {
    const float4 w = ((const float4*)P.rmap)[idx];
    const ushort4 i = ((const ushort4*)(P.rmap + n * sizeof(float4)))[idx];
    Lh = lerp_of(i.x, w.x, NH);
    Le = lerp_of(i.y, w.y, NE);
    Ls = lerp_of(i.z, w.z, NL);
    La = lerp_of(i.w, w.w, NA);
  } else {
    const int ia = (int)(idx % NA);
    long long q = idx / NA;
    const int is = (int)(q % NL);
    q /= NL;
    resample4_lerps(P, (int)(q / NE), (int)(q % NE), is, ia, Lh, Le, Ls, La);
  }
  float3 v;
  if (Lh.s == 0.0f && Le.s == 0.0f && Ls.s == 0.0f && La.s == 0.0f) {
    // the entry maps onto a grid point (every weight zero): the lookup is that point's value, bit for bit
    const long long g = (((long long)Lh.i0 * NE + Le.i0) * NL + Ls.i0) * NA + La.i0;
    v = ld3(a, g);
    if (b) {
      float3 w = ld3(b, g);
      if (c) {
        const float3 m = ld3(c, g);
        w = make_float3(w.x - m.x, w.y - m.y, w.z - m.z);
      }
      v.x += w.x; v.y += w.y; v.z += w.z;
    }
  } else {
    v = lookup4_skip([&](long long j) { return ld3(a, j); }, NE, NL, NA, Lh, Le, Ls, La);
    if (b) {
      const float3 w =
          c ? lookup4_skip(
                  [&](long long j) {
                    const float3 u = ld3(b, j), m = ld3(c, j);
                    return make_float3(u.x - m.x, u.y - m.y, u.z - m.z);
                  },
                  NE, NL, NA, Lh, Le, Ls, La)
            : lookup4_skip([&](long long j) { return ld3(b, j); }, NE, NL, NA, Lh, Le, Ls, La);
      v.x += w.x; v.y += w.y; v.z += w.z;
    }
  }
  return v;
}

// resample4_value(P, a, nullptr, idx) with the per-build map, in the same order of operations, one
// (height, elevation) cell at a time, and without the reads that a zero weight would mix in: mix3(u, v, 0)
// is u exactly, so a dimension whose weight is zero is not read (same result bit for bit).  Over the
// scatter grid of the full workload two thirds of the entries map onto a grid point, a quarter have only
// a heading weight (the clamped sun directions) and the height and sun-elevation weights are always zero:
// 1.4 reads per entry instead of 6.
__device__ float3 resample4_value_lean(const DevParams& P, const float4* __restrict__ a, long long idx) {
  const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
  const long long n = (long long)NH * NE * NL * NA;
  const float4 w = ((const float4*)P.rmap)[idx];
  const ushort4 i = ((const ushort4*)(P.rmap + n * sizeof(float4)))[idx];
  const Lerp Lh = lerp_of(i.x, w.x, NH), Le = lerp_of(i.y, w.y, NE), Ls = lerp_of(i.z, w.z, NL),
             La = lerp_of(i.w, w.w, NA);
  // entry offsets in 32 bits: a 4D table has fewer than 2^29 entries (validated on the host)
  auto row = [&](int h, int e, int s) {  // mixed over the heading
    const float4* c = a + ((unsigned)((h * NE + e) * NL + s)) * (unsigned)NA;
    const float3 u = ld3(c, La.i0);
    return La.s == 0.0f ? u : mix3(u, ld3(c, La.i1), La.s);
  };
  auto cell = [&](int h, int e) {  // mixed over the sun elevation
    const float3 u = row(h, e, Ls.i0);
    const float3 r = Ls.s == 0.0f ? u : mix3(u, row(h, e, Ls.i1), Ls.s);
    __builtin_amdgcn_sched_barrier(0);
    return r;
  };
  auto side = [&](int h) {  // mixed over the elevation
    const float3 u = cell(h, Le.i0);
    return Le.s == 0.0f ? u : mix3(u, cell(h, Le.i1), Le.s);
  };
  const float3 r0 = side(Lh.i0);
  return Lh.s == 0.0f ? r0 : mix3(r0, side(Lh.i1), Lh.s);
}

// Value at 2D entry idx of tab(a + b) (b nullable) for the transmittance (space 2) or
// surface-radiance (space 3) table.
__device__ float3 resample2_value(const DevParams& P, int space, const float4* __restrict__ a,
                                 const float4* __restrict__ b, long long idx) {
  SFSIM_EXACT  // the reference's rounding (see geometry.hpp)
  const int N0 = space == 2 ? P.nT[0] : P.nE[0], N1 = space == 2 ? P.nT[1] : P.nE[1];
  int i0 = (int)(idx / N1), i1 = (int)(idx % N1);
  double r0 = index_to_radius(P.g, N0, i0);
  double c1;
  if (space == 2) {
    double se, ce;
    bool above;
    index_to_elevation(P.g, N1, r0, i1, se, ce, above);
    c1 = elevation_to_index(P.g, N1, r0, r0 * se, above);
  } else {
    double s = index_to_sin_sun_elevation(N1, i1);
    c1 = sun_elevation_to_index(N1, (r0 * s) / r0);
  }
  Lerp L0 = make_lerp((float)height_to_index(P.g, N0, r0), N0);
  Lerp L1 = make_lerp((float)c1, N1);
  float3 v = lookup2_skip(a, N1, L0, L1);
  if (b) {
    float3 w = lookup2_skip(b, N1, L0, L1);
    v.x += w.x; v.y += w.y; v.z += w.z;
  }
  return v;
}

__device__ __forceinline__ void resample_entry(const DevParams& P, int space, const float4* __restrict__ a,
                                               const float4* __restrict__ b, float4* __restrict__ out4,
                                               float* __restrict__ out_rgb, int tiled, long long idx,
                                               const float4* __restrict__ c = nullptr) {
  float3 v;
  long long dst = idx;
  if (space == 4) {
    const int NH = P.nS[0], NE = P.nS[1], NL = P.nS[2], NA = P.nS[3];
    if (idx >= (long long)NH * NE * NL * NA) return;
    v = resample4_value(P, a, b, idx, c);
    if (tiled) {  // image.clj:299-312: row h * NL + s, column e * NA + a
      const int ia = (int)(idx % NA);
      long long q = idx / NA;
      const int is = (int)(q % NL);
      q /= NL;
      const int ie = (int)(q % NE), ih = (int)(q / NE);
      dst = (long long)(ih * NL + is) * (NE * NA) + (long long)ie * NA + ia;
    }
  } else {
    const int N0 = space == 2 ? P.nT[0] : P.nE[0], N1 = space == 2 ? P.nT[1] : P.nE[1];
    if (idx >= (long long)N0 * N1) return;
    v = resample2_value(P, space, a, b, idx);
  }
  if (out4) out4[idx] = make_float4(v.x, v.y, v.z, 0.f);
  if (out_rgb) {
    out_rgb[dst * 3 + 0] = v.x;
    out_rgb[dst * 3 + 1] = v.y;
    out_rgb[dst * 3 + 2] = v.z;
  }
}

template <bool kB>
__global__ void __launch_bounds__(kBlock) resample_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                          int space, const float4* __restrict__ a,
                                                          const float4* __restrict__ b, float4* __restrict__ out4,
                                                          float* __restrict__ out_rgb, int tiled) {
  const DevParams& P = params_of<kB>(P0, PB);
  resample_entry(P, space, atm_ptr<kB>(a, P), atm_ptr<kB>(b, P), atm_ptr<kB>(out4, P), atm_ptr<kB>(out_rgb, P), tiled,
                 (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Several independent re-samplings in one launch: job k takes workgroups first[k] .. first[k + 1] - 1.
struct ResampleJobs {
  ResampleJob job[kMaxResampleJobs];
  unsigned first[kMaxResampleJobs + 1];
  int count;
};

template <bool kB>
__global__ void __launch_bounds__(kBlock) resample_jobs_kernel(const DevParams P0, const DevParams* __restrict__ PB,
                                                               const ResampleJobs J) {
  const DevParams& P = params_of<kB>(P0, PB);
  int k = 0;
  while (k + 1 < J.count && blockIdx.x >= J.first[k + 1]) ++k;
  const ResampleJob& j = J.job[k];
  resample_entry(P, j.space, atm_ptr<kB>(j.a, P), atm_ptr<kB>(j.b, P), atm_ptr<kB>(j.out4, P),
                 atm_ptr<kB>(j.out_rgb, P), j.tiled, (long long)(blockIdx.x - J.first[k]) * blockDim.x + threadIdx.x,
                 atm_ptr<kB>(j.c, P));
}

__global__ void pack_rgb_kernel(const float4* __restrict__ in, float* __restrict__ out, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float4 v = in[i];
  out[i * 3] = v.x;
  out[i * 3 + 1] = v.y;
  out[i * 3 + 2] = v.z;
}

__global__ void unpack_rgb_kernel(const float* __restrict__ in, float4* __restrict__ out, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = make_float4(in[i * 3], in[i * 3 + 1], in[i * 3 + 2], 0.f);
}

__global__ void add4_kernel(const float4* __restrict__ a, const float4* __restrict__ b, float4* __restrict__ c,
                            long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float4 x = a[i], y = b[i];
  c[i] = make_float4(x.x + y.x, x.y + y.y, x.z + y.z, 0.f);
}

// Block maxima of |a + x| per channel: partial[block] (x, y, z).  Two passes without atomics (the
// second is one block over the partials), deterministic and free of memsets.
__device__ __forceinline__ float3 block_max3(float3 m, float3* red) {
  for (int off = 32; off > 0; off >>= 1) {
    m.x = fmaxf(m.x, __shfl_down(m.x, off, 64));
    m.y = fmaxf(m.y, __shfl_down(m.y, off, 64));
    m.z = fmaxf(m.z, __shfl_down(m.z, off, 64));
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      m = make_float3(fmaxf(m.x, red[w].x), fmaxf(m.y, red[w].y), fmaxf(m.z, red[w].z));
  return m;  // valid in thread 0
}

__global__ void __launch_bounds__(kBlock) channel_max_kernel(const float4* __restrict__ a, const float4* __restrict__ b,
                                                             long long count, float4* __restrict__ partial) {
  __shared__ float3 red[kBlock / 64];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float3 m = make_float3(0.f, 0.f, 0.f);
  if (i < count) {
    float4 x = a[i], y = b[i];
    m = make_float3(fabsf(x.x + y.x), fabsf(x.y + y.y), fabsf(x.z + y.z));
  }
  m = block_max3(m, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_float4(m.x, m.y, m.z, 0.f);
}

// out[0] = maximum over the n partials (one block).
__global__ void __launch_bounds__(kBlock) max_partials_kernel(const float4* __restrict__ partial, int n,
                                                              float4* __restrict__ out) {
  __shared__ float3 red[kBlock / 64];
  float3 m = make_float3(0.f, 0.f, 0.f);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float4 v = partial[i];
    m = make_float3(fmaxf(m.x, v.x), fmaxf(m.y, v.y), fmaxf(m.z, v.z));
  }
  m = block_max3(m, red);
  if (threadIdx.x == 0) out[0] = make_float4(m.x, m.y, m.z, 0.f);
}

// Block maxima over entries and channels of |x - y| / max(|a + x|, floor_c), floor_c = fraction *
// channel max (cmax), in partial[block].x.
__global__ void __launch_bounds__(kBlock) residual_kernel(const float4* __restrict__ a, const float4* __restrict__ x,
                                                          const float4* __restrict__ y, long long count,
                                                          const float4* __restrict__ cmax, float fraction,
                                                          float4* __restrict__ partial) {
  __shared__ float3 red[kBlock / 64];
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float r = 0.f;
  if (i < count) {
    const float4 c = cmax[0];
    float4 av = a[i], xv = x[i], yv = y[i];
    float f0 = fraction * c.x, f1 = fraction * c.y, f2 = fraction * c.z;
    r = fmaxf(fabsf(xv.x - yv.x) / fmaxf(fabsf(av.x + xv.x), fmaxf(f0, 1e-30f)),
              fmaxf(fabsf(xv.y - yv.y) / fmaxf(fabsf(av.y + xv.y), fmaxf(f1, 1e-30f)),
                    fabsf(xv.z - yv.z) / fmaxf(fabsf(av.z + xv.z), fmaxf(f2, 1e-30f))));
  }
  float3 m = block_max3(make_float3(r, r, r), red);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_float4(m.x, 0.f, 0.f, 0.f);
}

__global__ void fill_zero_kernel(float* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_transmittance(const DevParams& P, float4* out, WorkCounters* wc, hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  long long n = (long long)P.nT[0] * P.nT[1];
  hipLaunchKernelGGL(transmittance_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, P, out, wc);
  return hipGetLastError();
}

hipError_t launch_surface_radiance_base(const DevParams& P, float4* out, WorkCounters* wc, hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  long long n = (long long)P.nE[0] * P.nE[1];
  hipLaunchKernelGGL(surface_radiance_base_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, P, out, wc);
  return hipGetLastError();
}

hipError_t launch_ray_geometry(const DevParams& P, WorkCounters* wc, hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  if (!P.ray_geo) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ray_geometry_kernel, dim3(blocks_for((long long)ray_geo_samples(P))), dim3(kBlock), 0, stream, P,
                     wc);
  return hipGetLastError();
}

hipError_t launch_single_scatter(const DevParams& P, float4* rayleigh, float4* mie, WorkCounters* wc,
                                 hipStream_t stream, bool sun_table_built) {
  unsigned rays = (unsigned)(P.nS[0] * P.nS[1]);
  // a batch runs the table path with the tables built by its prologue
  if (P.batch > 0 && (P.sun_rows <= 0 || !sun_table_built)) return hipErrorInvalidValue;
  if (P.sun_rows > 0) {
    if (!P.sun_depth || !P.ray_geo || !P.sun_ss || P.sun_rows < 2 || P.sun_cols < 2) return hipErrorInvalidValue;
    if (!sun_table_built)
      hipLaunchKernelGGL(sun_depth_kernel, dim3(blocks_for((long long)P.sun_rows * P.sun_cols)), dim3(kBlock), 0, stream,
                         P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const RayGeoView V = ray_geo_view(P);
    const size_t shmem = (size_t)P.ray_steps * (sizeof(double2) + 2 * sizeof(float4) + sizeof(float2));
    if (P.batch > 0)
      hipLaunchKernelGGL(single_scatter_table_kernel<true>, dim3(rays, 1, P.batch), dim3(kSSTBlock), shmem, stream, P,
                         P.batch_params, V.th, V.T, V.ss, V.row, rayleigh, mie, wc);
    else
      hipLaunchKernelGGL(single_scatter_table_kernel<false>, dim3(rays), dim3(kSSTBlock), shmem, stream, P,
                         P.batch_params, V.th, V.T, V.ss, V.row, rayleigh, mie, wc);
    return hipGetLastError();
  }
  unsigned groups = (unsigned)((P.nS[2] * P.nS[3] + kSSWave - 1) / kSSWave);
  int n = P.ray_steps;
  size_t shmem = (size_t)n * (2 * sizeof(double) + sizeof(float4)) + sizeof(float) * (size_t)((n + 3) & ~3) +
                 (3 * sizeof(float) + sizeof(unsigned short)) * kSSWave * kSSChunk + 2 * sizeof(double) * kSSWave;
  hipLaunchKernelGGL(single_scatter_kernel, dim3(rays * groups), dim3(kSSWave), shmem, stream, P, rayleigh, mie, wc);
  return hipGetLastError();
}

hipError_t launch_direction_geometry(const DevParams& P, const Dir* dirs, int ndirs, void* geom, WorkCounters* wc,
                                    hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  long long n = (long long)P.nS[0] * ndirs;
  hipLaunchKernelGGL(direction_geometry_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, P, dirs, ndirs,
                     (DirGeom*)geom, wc);
  return hipGetLastError();
}

size_t direction_geometry_bytes(const DevParams& P, int ndirs) { return sizeof(DirGeom) * (size_t)P.nS[0] * ndirs; }

hipError_t launch_point_scatter(const DevParams& P, const Dir* dirs, int ndirs, const void* geom, const float4* dS,
                                const float4* mie, const float4* dE, float4* slab, float4* slab_mie, float4* out,
                                WorkCounters* wc, hipStream_t stream, const SrJob* sr) {
  if (!P.sun_ps) return hipErrorInvalidValue;
  unsigned combos = (unsigned)(P.nS[0] * ndirs);
  const unsigned nz = P.batch > 0 ? (unsigned)P.batch : 1u;
  if (P.batch > 0)
    hipLaunchKernelGGL(point_scatter_combine_kernel<true>, dim3(combos, 1, nz), dim3(kBlock), 0, stream, P,
                       P.batch_params, (const DirGeom*)geom, ndirs, dS, mie, slab, mie ? slab_mie : nullptr);
  else
    hipLaunchKernelGGL(point_scatter_combine_kernel<false>, dim3(combos), dim3(kBlock), 0, stream, P, P.batch_params,
                       (const DirGeom*)geom, ndirs, dS, mie, slab, mie ? slab_mie : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t e_bytes = sizeof(float4) * (size_t)P.nE[1];  // dE height row 0, the ground
  const bool e_lds = e_bytes <= 16384;
  size_t shmem = sizeof(DirLds) * (size_t)ndirs + (e_lds ? e_bytes : 0);
  SrJob job{};
  unsigned extra = 0;
  if (sr) {
    if (!P.sgeo || !sr->out || !sr_fits_point_scatter(P, sr->mie != nullptr)) return hipErrorInvalidValue;
    job = *sr;
    job.count = P.nE[0] * P.nE[1];
    extra = (unsigned)job.count;
    shmem = std::max(shmem, sr_fold_bytes(P, sr->mie != nullptr));
  }
  const unsigned blocks = (unsigned)(P.nS[0] * P.nS[1]) + extra;
  const DirGeom* g = (const DirGeom*)geom;
  const dim3 grid(blocks, 1, nz);
  const DevParams* PB = P.batch_params;
  if (P.batch > 0) {
    if (mie && e_lds)
      hipLaunchKernelGGL((point_scatter_kernel<true, true, true>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, slab_mie, dE, out, job, wc);
    else if (mie)
      hipLaunchKernelGGL((point_scatter_kernel<true, false, true>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, slab_mie, dE, out, job, wc);
    else if (e_lds)
      hipLaunchKernelGGL((point_scatter_kernel<false, true, true>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, nullptr, dE, out, job, wc);
    else
      hipLaunchKernelGGL((point_scatter_kernel<false, false, true>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, nullptr, dE, out, job, wc);
  } else {
    if (mie && e_lds)
      hipLaunchKernelGGL((point_scatter_kernel<true, true, false>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, slab_mie, dE, out, job, wc);
    else if (mie)
      hipLaunchKernelGGL((point_scatter_kernel<true, false, false>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, slab_mie, dE, out, job, wc);
    else if (e_lds)
      hipLaunchKernelGGL((point_scatter_kernel<false, true, false>), grid, dim3(kBlock), shmem, stream, P, PB, dirs, g,
                         ndirs, slab, nullptr, dE, out, job, wc);
    else
      hipLaunchKernelGGL((point_scatter_kernel<false, false, false>), grid, dim3(kBlock), shmem, stream, P, PB, dirs,
                         g, ndirs, slab, nullptr, dE, out, job, wc);
  }
  return hipGetLastError();
}

// Samples per folded-slab chunk of the ray-scatter kernel: its two slab halves take 2 x Q x NL x NA float4
// of LDS, about 2 x 1024 slab entries (32 KiB, plus 36 B per ray step of offsets, weights and masks), so
// that four 512-thread workgroups fit a CU; SFSIM_ATMO_RAY_CHUNK (1..8, within the same LDS bound)
// overrides it.
int ray_scatter_chunk(const DevParams& P) {
  int nla = P.nS[2] * P.nS[3];
  int q = 1024 / nla;
  if (const char* env = std::getenv("SFSIM_ATMO_RAY_CHUNK")) {
    int v = std::atoi(env);
    if (v >= 1) q = v;
  }
  const int qmax = std::min(kMaxRayChunk, 1024 / nla);
  return q < 1 ? 1 : (q > qmax ? std::max(qmax, 1) : q);
}

// launch_ray_scatter with the sun-elevation coordinate's form chosen (kExact: sun_coord_exact)
template <bool kExact>
hipError_t launch_ray_scatter_as(const DevParams& P, const float4* dJ, float4* out, WorkCounters* wc,
                                 hipStream_t stream, const RsAccum& acc, unsigned rays, int Q, size_t shmem,
                                 const RayGeoView& V, const DevParams* PB) {
  const int nla = P.nS[2] * P.nS[3];
  if (P.batch > 0) {
    const dim3 grid(rays, 1, (unsigned)P.batch);
    if (nla <= kRSHalf)
      hipLaunchKernelGGL((ray_scatter_kernel<1, true, kExact>), grid, dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
    else if (nla <= 2 * kRSHalf)
      hipLaunchKernelGGL((ray_scatter_kernel<2, true, kExact>), grid, dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
    else
      hipLaunchKernelGGL((ray_scatter_kernel<4, true, kExact>), grid, dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
  } else {
    if (nla <= kRSHalf)
      hipLaunchKernelGGL((ray_scatter_kernel<1, false, kExact>), dim3(rays), dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
    else if (nla <= 2 * kRSHalf)
      hipLaunchKernelGGL((ray_scatter_kernel<2, false, kExact>), dim3(rays), dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
    else
      hipLaunchKernelGGL((ray_scatter_kernel<4, false, kExact>), dim3(rays), dim3(kRSBlock), shmem, stream, P, PB, dJ,
                         V.th, V.T, V.rs, V.fold, out, acc, Q, wc);
  }
  return hipGetLastError();
}

hipError_t launch_ray_scatter(const DevParams& P, const float4* dJ, float4* out, WorkCounters* wc,
                              hipStream_t stream, const RsAccum& acc) {
  if ((acc.S_in == nullptr) != (acc.S_out == nullptr) || (acc.E_out && !acc.dE)) return hipErrorInvalidValue;
  if (!P.ray_geo || !P.sun_rs) return hipErrorInvalidValue;
  const unsigned rays = (unsigned)(P.nS[0] * P.nS[1]) + (acc.E_out ? 1u : 0u);
  int nla = P.nS[2] * P.nS[3];
  int Q = ray_scatter_chunk(P);
  const size_t shmem = 2 * sizeof(float4) * (size_t)Q * nla;
  const RayGeoView V = ray_geo_view(P);
  const DevParams* PB = P.batch_params;
  if (nla > 4 * kRSHalf) return hipErrorInvalidValue;
  if (P.nS[2] < kExactSunRows) {
    if (!P.sun_ss) return hipErrorInvalidValue;
    return launch_ray_scatter_as<true>(P, dJ, out, wc, stream, acc, rays, Q, shmem, V, PB);
  }
  return launch_ray_scatter_as<false>(P, dJ, out, wc, stream, acc, rays, Q, shmem, V, PB);
}

hipError_t launch_surface_geometry(const DevParams& P, const Dir* dirs, int ndirs, hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  if (!P.sgeo) return hipErrorInvalidValue;
  hipLaunchKernelGGL(surface_geometry_kernel, dim3(blocks_for((long long)P.nE[0] * ndirs)), dim3(kBlock), 0, stream, P,
                     dirs, ndirs);
  return hipGetLastError();
}

size_t sr_fold_bytes(const DevParams& P, bool mie) { return sizeof(float4) * (size_t)P.nS[1] * P.nS[3] * (mie ? 2 : 1); }

bool sr_fits_point_scatter(const DevParams& P, bool mie) { return sr_fold_bytes(P, mie) <= 16384; }

hipError_t launch_surface_radiance(const DevParams& P, int ndirs, const float4* dS, const float4* mie, float4* out,
                                   WorkCounters* wc, hipStream_t stream) {
  if (!P.sgeo) return hipErrorInvalidValue;
  unsigned entries = (unsigned)(P.nE[0] * P.nE[1]);
  const size_t fold = sr_fold_bytes(P, mie != nullptr);
  const dim3 grid(entries, 1, P.batch > 0 ? (unsigned)P.batch : 1u);
  const DevParams* PB = P.batch_params;
  if (P.batch > 0) {
    if (fold <= 65536)
      hipLaunchKernelGGL((surface_radiance_kernel<true, true>), grid, dim3(kBlock), fold, stream, P, PB, ndirs, dS, mie,
                         out, wc);
    else
      hipLaunchKernelGGL((surface_radiance_kernel<false, true>), grid, dim3(kBlock), 0, stream, P, PB, ndirs, dS, mie,
                         out, wc);
  } else {
    if (fold <= 65536)
      hipLaunchKernelGGL((surface_radiance_kernel<true, false>), grid, dim3(kBlock), fold, stream, P, PB, ndirs, dS, mie,
                         out, wc);
    else
      hipLaunchKernelGGL((surface_radiance_kernel<false, false>), grid, dim3(kBlock), 0, stream, P, PB, ndirs, dS, mie,
                         out, wc);
  }
  return hipGetLastError();
}

hipError_t launch_prologue(const DevParams& P, const PrologueJobs& J, hipStream_t stream) {
  if (!P.ray_geo || !P.rmap || (P.sun_rows > 0 && !P.sun_depth)) return hipErrorInvalidValue;
  if ((J.sdirs && !J.geom) || (J.hdirs && !P.sgeo) || !J.T || !J.E0) return hipErrorInvalidValue;
  PrologueArgs A;
  A.j = J;
  const long long counts[7] = {J.sdirs ? (long long)P.nS[0] * J.sphere_n : 0,
                               (long long)P.nT[0] * P.nT[1],
                               (long long)P.nE[0] * P.nE[1],
                               J.hdirs ? (long long)P.nE[0] * J.half_n : 0,
                               P.sun_rows > 0 ? (long long)P.sun_rows * P.sun_cols : 0,
                               (long long)ray_geo_samples(P),
                               (long long)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3]};
  A.first[0] = 0;
  for (int k = 0; k < 7; ++k) A.first[k + 1] = A.first[k] + blocks_for(counts[k]);
  if (P.batch > 0)
    hipLaunchKernelGGL(prologue_kernel<true>, dim3(A.first[7], 1, P.batch), dim3(kBlock), 0, stream, P,
                       P.batch_params, A);
  else
    hipLaunchKernelGGL(prologue_kernel<false>, dim3(A.first[7]), dim3(kBlock), 0, stream, P, P.batch_params, A);
  return hipGetLastError();
}

hipError_t launch_resample_map(const DevParams& P, hipStream_t stream) {
  if (P.batch > 0) return hipErrorInvalidValue;  // batches run these jobs in launch_prologue
  if (!P.rmap) return hipErrorInvalidValue;
  const long long n = (long long)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3];
  hipLaunchKernelGGL(resample_map_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, P);
  return hipGetLastError();
}

static long long resample_count(const DevParams& P, int space) {
  return space == 4   ? (long long)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3]
         : space == 2 ? (long long)P.nT[0] * P.nT[1]
                      : (long long)P.nE[0] * P.nE[1];
}

hipError_t launch_resamples(const DevParams& P, const ResampleJob* jobs, int count, hipStream_t stream) {
  if (count < 1 || count > kMaxResampleJobs) return hipErrorInvalidValue;
  ResampleJobs J;
  J.count = count;
  J.first[0] = 0;
  for (int k = 0; k < kMaxResampleJobs; ++k) {
    J.job[k] = jobs[k < count ? k : 0];
    if (k < count) {
      if (jobs[k].space < 2 || jobs[k].space > 4 || !jobs[k].a || (jobs[k].c && (jobs[k].space != 4 || !jobs[k].b)))
        return hipErrorInvalidValue;
      J.first[k + 1] = J.first[k] + blocks_for(resample_count(P, jobs[k].space));
    }
  }
  if (P.batch > 0)
    hipLaunchKernelGGL(resample_jobs_kernel<true>, dim3(J.first[count], 1, P.batch), dim3(kBlock), 0, stream, P,
                       P.batch_params, J);
  else
    hipLaunchKernelGGL(resample_jobs_kernel<false>, dim3(J.first[count]), dim3(kBlock), 0, stream, P, P.batch_params,
                       J);
  return hipGetLastError();
}

hipError_t launch_resample(const DevParams& P, int space, const float4* a, const float4* b, float4* out4,
                           float* out_rgb, int tiled, hipStream_t stream) {
  const long long n = resample_count(P, space);
  if (P.batch > 0)
    hipLaunchKernelGGL(resample_kernel<true>, dim3(blocks_for(n), 1, P.batch), dim3(kBlock), 0, stream, P,
                       P.batch_params, space, a, b, out4, out_rgb, tiled);
  else
    hipLaunchKernelGGL(resample_kernel<false>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, P, P.batch_params, space,
                       a, b, out4, out_rgb, tiled);
  return hipGetLastError();
}

hipError_t launch_add4(const float4* a, const float4* b, float4* c, long long count, hipStream_t stream) {
  hipLaunchKernelGGL(add4_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, stream, a, b, c, count);
  return hipGetLastError();
}

size_t residual_scratch_floats(long long count) { return 4 * (2 + (size_t)blocks_for(count)); }

hipError_t launch_residual(const float4* a, const float4* x, const float4* y, long long count, float fraction,
                           float* scratch, hipStream_t stream) {
  const unsigned nb = blocks_for(count);
  float4* s4 = (float4*)scratch;  // [0] channel maxima, [1] result, [2 ..] block partials
  hipLaunchKernelGGL(channel_max_kernel, dim3(nb), dim3(kBlock), 0, stream, a, x, count, s4 + 2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(kBlock), 0, stream, s4 + 2, (int)nb, s4);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(kBlock), 0, stream, a, x, y, count, s4, fraction, s4 + 2);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(kBlock), 0, stream, s4 + 2, (int)nb, s4 + 1);
  return hipGetLastError();
}

hipError_t launch_fill_zero(float* p, long long n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, p, n);
  return hipGetLastError();
}

hipError_t launch_pack_rgb(const float4* in, float* out, long long count, hipStream_t stream) {
  hipLaunchKernelGGL(pack_rgb_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, stream, in, out, count);
  return hipGetLastError();
}

hipError_t launch_unpack_rgb(const float* in, float4* out, long long count, hipStream_t stream) {
  hipLaunchKernelGGL(unpack_rgb_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, stream, in, out, count);
  return hipGetLastError();
}

}  // namespace sfsim
