// Parameters of one table build as the kernels see them (host fills, passed by value).
#pragma once

#include <cstdint>
#include <type_traits>

#include "geometry.hpp"

namespace sfsim {

constexpr int kMaxSteps = 256;   // ray_steps limit (LDS arrays of the per-ray kernels)
constexpr int kMaxSphereDirs = 1000;  // 64 bytes of LDS per direction in the point-scatter kernel

struct DevParams {
  Geo g;
  // extinction (scatter-base / scatter-quotient) and scattering (scatter-base) per component, FP32
  float ext_mie[3], ext_ray[3];
  float ext_mie_l2[3], ext_ray_l2[3];  // the same times log2(e): T = exp2(-(ext_mie_l2 I_mie + ext_ray_l2 I_ray))
  float sca_mie[3], sca_ray[3];
  // exp(-h / scale) = exp2(h * k): k = -log2(e) / scale
  float k_mie, k_ray;
  // FP32 sums in units of the planet radius: with x = (r^2 - R^2) / R^2 the Mie exponent is
  // h k_mie = x (a0 + x (a1 + x (a2 + x a3))), a cubic fit at Chebyshev nodes on [0, (Rt^2 - R^2) / R^2]
  // of (R / 2) k_mie 2 (sqrt(1 + x) - 1) / x; the Rayleigh exponent is ray_over_mie times it
  float alt_poly[4];
  int alt_deg;           // 2 when the quadratic fit is within 1e-7 relative on the whole range, else 3
  double alt_poly_d[4];  // the cubic in FP64 (FP64 sums)
  float ray_over_mie;
  double inv_R;
  double inv_scale_mie, inv_scale_ray;
  float g_mie, g_ray;
  float sun_inv_den;  // 1 / (1 - exp(-3.6)) of sun-elevation-to-index (atmosphere.clj:322-326)
  float brightness_over_pi[3];
  float intensity[3];
  int nT[2];  // transmittance shape  [height elevation]
  int nE[2];  // surface radiance shape [height sun-elevation]
  int nS[4];  // scatter shape [height elevation sun-elevation heading]
  int ray_steps;
  double inv_steps;  // 1 / ray_steps
  // optical depth of a segment whose lowest point lies strictly inside it and below the altitude
  // graze_height (u = r^2 - R^2 < graze_u there) is accumulated in FP64 (near-horizon / grazing
  // rays); everything else in FP32 with compensated sums
  double graze_height, graze_u;
  // single scatter: the sun segments' optical depths from a sun_rows x sun_cols table of log2 of their
  // midpoint sums (sun_depth, built once per build); sun_rows == 0 sums every sun segment instead
  int sun_rows, sun_cols;
  float2* sun_depth;
  // per-(view ray, sample) geometry of the single-scatter and ray-scatter kernels, computed once per
  // build (ray_geometry_bytes(P) bytes, laid out by ray_geo_view)
  char* ray_geo;
  // per-entry lookup of the 4D re-sampling (forward(backward(i))): NH x NE x NL x NA float4 weights
  // followed by as many ushort4 lower indices (resample_map_bytes); null: computed per launch
  char* rmap;
  // per-entry sun terms, written with the re-sampling map (launch_resample_map) and read by every
  // point- and ray-scatter launch: sun_ps = (l, the sun-elevation coordinate) for point scattering,
  // sun_rs = (v . l, x . l, the heading coordinate, 0) for ray scattering, each rounded once from FP64
  // (NH x NE x NL x NA float4 each); sun_ss = (v . l, x . l) in FP64 for single scattering's horizon
  // tests (the table path)
  float4* sun_ps;
  float4* sun_rs;
  double2* sun_ss;
  // per (surface-radiance height, half-sphere direction): omega.x, omega.y, elevation coordinate of
  // omega at the height's grid point, cosine-weighted quadrature weight (nE[0] x directions float4)
  float4* sgeo;
  // order in which the ray-scatter kernel visits the view rays (NH x NE ray indices, ascending impact
  // parameter |x| sqrt(1 - mu^2)); null: grid order
  const int* ray_order;
  // batched builds (launches over several atmospheres sharing the planet and the grid): batch > 0
  // atmospheres, atmosphere z = blockIdx.z with parameters batch_params[z] (a device array), and every
  // per-atmosphere table (and ray_geo, sun_depth) of atmosphere z lies z atm_stride bytes after
  // atmosphere 0's; the geometry that depends only on the planet and the grid (rmap, sun maps, sgeo,
  // ray_order, quadrature directions) is shared.  batch == 0: one atmosphere, parameters by value.
  int batch;
  size_t atm_stride;
  const DevParams* batch_params;
};

// The parameters of the launch's atmosphere.  Batched: the scalars from the device array, the table
// pointers from the kernel's own argument (a global pointer to the compiler, so that loads through it stay
// global rather than flat loads), offset to the atmosphere's block where the table is per atmosphere (the
// ray geometry and the sun table; the re-sampling and sun maps, the half-sphere geometry and the ray order
// are shared).  Single atmosphere: the kernel argument itself.
template <bool kB>
__device__ __forceinline__ decltype(auto) params_of(const DevParams& P0, const DevParams* __restrict__ PB) {
  if constexpr (kB) {
    DevParams P = PB[blockIdx.z];
    const size_t off = (size_t)blockIdx.z * P0.atm_stride;
    P.sun_depth = P0.sun_depth ? reinterpret_cast<float2*>(reinterpret_cast<char*>(P0.sun_depth) + off) : nullptr;
    P.ray_geo = P0.ray_geo ? P0.ray_geo + off : nullptr;
    P.rmap = P0.rmap;
    P.sun_ps = P0.sun_ps;
    P.sun_rs = P0.sun_rs;
    P.sun_ss = P0.sun_ss;
    P.sgeo = P0.sgeo;
    P.ray_order = P0.ray_order;
    return P;
  } else {
    return (P0);
  }
}

// This atmosphere's copy of a per-atmosphere table.

// (Byte arithmetic on a char pointer, not an integer round trip, so that the compiler still sees a global
// pointer and emits global rather than flat loads.)
template <bool kB, class T>
__device__ __forceinline__ T* atm_ptr(T* p, const DevParams& P) {
  if constexpr (kB) {
    using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
    return p ? reinterpret_cast<T*>(reinterpret_cast<Byte*>(p) + (size_t)blockIdx.z * P.atm_stride) : p;
  } else {
    return p;
  }
}

__host__ __device__ inline size_t resample_map_bytes(const DevParams& P) {
  return (size_t)P.nS[0] * P.nS[1] * P.nS[2] * P.nS[3] * (sizeof(float4) + sizeof(ushort4));
}

// Views of DevParams::ray_geo: m = NH x NE x ray_steps samples, sample i of view ray (ih, ie) at
// (ih NE + ie) ray_steps + i; t_i = (i + 1/2) D / ray_steps along the ray from the grid point.
// The ray-scatter fold of sample p_i: the element offsets in the dJ table of the (height, elevation) nodes
// of its bilinear cell by fold slot (slot 2 (h & 1) + (e & 1); -1 where the slot is unused at the upper
// edge of the grid) and their weights (zero for unused slots).
struct FoldStep {
  int4 off;
  float4 wt;
};

struct RayGeoView {
  double2* th;    // t_i, altitude h_i
  float4* T;      // T(x, p_i) rgb, w: exp(-h_i / H_mie)
  float4* rs;     // height coordinate, elevation coordinate of v, alpha_i, beta_i (sun-elevation coordinate terms)
  float4* ss;     // Rt^2 - r_i^2, D_min, (sun_cols - 1) / (D_max - D_min), exp(-h_i / H_rayleigh) (sun table)
  FoldStep* fold; // the dJ nodes and weights of the ray-scatter fold
  float2* row;    // sun-table row fraction, bits of the int index of the row start
};

__host__ __device__ inline size_t ray_geo_samples(const DevParams& P) {
  return (size_t)P.nS[0] * P.nS[1] * P.ray_steps;
}

__host__ __device__ inline size_t ray_geometry_bytes(const DevParams& P) {
  return ray_geo_samples(P) * (sizeof(double2) + 3 * sizeof(float4) + sizeof(FoldStep) + sizeof(float2));
}

__host__ __device__ inline RayGeoView ray_geo_view(const DevParams& P) {
  const size_t m = ray_geo_samples(P);
  char* b = P.ray_geo;
  RayGeoView v;
  v.th = (double2*)b;
  b += m * sizeof(double2);
  v.T = (float4*)b;
  b += m * sizeof(float4);
  v.rs = (float4*)b;
  b += m * sizeof(float4);
  v.ss = (float4*)b;
  b += m * sizeof(float4);
  v.fold = (FoldStep*)b;
  b += m * sizeof(FoldStep);
  v.row = (float2*)b;
  return v;
}

// Quadrature direction (FP64 unit vector in the frame of a grid point on +x) and its weight.
struct Dir {
  double x, y, z, w;
};

// Optional device counters of the work the kernels do (instrumented builds only).
struct WorkCounters {
  unsigned long long single_samples;  // extinction samples in the single-scatter kernel
  unsigned long long samples;         // extinction samples, all kernels
  unsigned long long lookups4;        // 4D multilinear lookups
  unsigned long long lookups2;        // 2D multilinear lookups
};

}  // namespace sfsim
