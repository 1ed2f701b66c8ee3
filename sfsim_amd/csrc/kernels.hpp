// Launchers of the table kernels (kernels.hip).  All take device pointers and enqueue on `stream`.
// Tables are float4 per entry (RGB + pad) in grid order unless stated otherwise.
#pragma once

#include <hip/hip_runtime.h>

#include "device_params.hpp"

namespace sfsim {

hipError_t launch_transmittance(const DevParams& P, float4* out, WorkCounters* wc, hipStream_t stream);
hipError_t launch_surface_radiance_base(const DevParams& P, float4* out, WorkCounters* wc, hipStream_t stream);
// Per-(view ray, sample) geometry into P.ray_geo (once per build, before single and ray scatter).
hipError_t launch_ray_geometry(const DevParams& P, WorkCounters* wc, hipStream_t stream);
// R1 and M1; with a sun table (P.sun_rows > 0) builds it first (unless sun_table_built: launch_prologue
// did) and reads P.ray_geo and the per-build sun map P.sun_ss (launch_resample_map).
hipError_t launch_single_scatter(const DevParams& P, float4* rayleigh, float4* mie, WorkCounters* wc,
                                 hipStream_t stream, bool sun_table_built = false);
// The per-build launches that read nothing of each other, in one launch: direction geometry of the
// sphere quadrature (when sdirs, into geom), transmittance T, E0, the half-sphere geometry (when hdirs,
// into P.sgeo), the sun table (P.sun_rows > 0), the ray geometry and the re-sampling map.
struct PrologueJobs {
  float4* T;
  float4* E0;
  const Dir* sdirs;
  int sphere_n;
  void* geom;
  const Dir* hdirs;
  int half_n;
  WorkCounters* wc;
};
hipError_t launch_prologue(const DevParams& P, const PrologueJobs& J, hipStream_t stream);
// Per-(height, direction) geometry of the sphere quadrature (once per build); geom holds
// direction_geometry_bytes(P, ndirs) bytes.
hipError_t launch_direction_geometry(const DevParams& P, const Dir* dirs, int ndirs, void* geom, WorkCounters* wc,
                                     hipStream_t stream);
size_t direction_geometry_bytes(const DevParams& P, int ndirs);
// dS: previous-order scatter table; mie (nullable): first-iteration Mie strength table, added with the
// Mie phase of the quadrature direction; dE: previous-order surface-radiance table.  slab / slab_mie
// are scratch tables of NH x ndirs x NL x NA float4 (slab_mie used only when mie is given).  Reads the
// per-build sun map P.sun_ps (launch_resample_map or launch_prologue).
// The surface radiance of the same dS and Mie strength (launch_surface_radiance) computed in the
// point-scatter launch: ndirs half-sphere directions, output table out; count is set by the launcher.
struct SrJob {
  int count;
  int ndirs;
  const float4* dS;
  const float4* mie;
  float4* out;
};
// Whether launch_point_scatter can take the surface radiance along (its LDS slab fits 16 KiB).
bool sr_fits_point_scatter(const DevParams& P, bool mie);
size_t sr_fold_bytes(const DevParams& P, bool mie);
hipError_t launch_point_scatter(const DevParams& P, const Dir* dirs, int ndirs, const void* geom, const float4* dS,
                                const float4* mie, const float4* dE, float4* slab, float4* slab_mie, float4* out,
                                WorkCounters* wc, hipStream_t stream, const SrJob* sr = nullptr);
// The accumulations of atmosphere_lut.clj:95-97 that ride in a ray-scatter launch.  S_out (nullable,
// with S_in): the running sum S + dS at every grid point, S = S_in, or tab(S_in) when S_resample (S_in
// then holds the previous iteration's S + dS).  E_out (nullable): E_out = tab(dE + E_in) over the
// surface-radiance table (E_in nullable), by one extra workgroup.
struct RsAccum {
  const float4* S_in = nullptr;
  float4* S_out = nullptr;
  int S_resample = 0;
  const float4* dE = nullptr;
  const float4* E_in = nullptr;
  float4* E_out = nullptr;
};
// Reads P.ray_geo (launch_ray_geometry) and the per-build sun map P.sun_rs (launch_resample_map or
// launch_prologue).
hipError_t launch_ray_scatter(const DevParams& P, const float4* dJ, float4* out, WorkCounters* wc,
                              hipStream_t stream, const RsAccum& acc = RsAccum());
// Half-sphere direction geometry of every surface-radiance height into P.sgeo (once per build, before
// launch_surface_radiance).
hipError_t launch_surface_geometry(const DevParams& P, const Dir* dirs, int ndirs, hipStream_t stream);
hipError_t launch_surface_radiance(const DevParams& P, int ndirs, const float4* dS, const float4* mie, float4* out,
                                   WorkCounters* wc, hipStream_t stream);
// Re-sample a (+ b when non-null) at forward(backward(i)) of the space (2, 3 or 4).  Output either as
// float4 grid-order table (out4) or as packed RGB floats (out_rgb): grid order, or for space 4 with
// tiled != 0 the 4D->2D tiling of the .scatter files.
// The 4D re-sampling lookups of every entry into P.rmap (once per build, before any space-4 resample
// that runs with P.rmap set).
hipError_t launch_resample_map(const DevParams& P, hipStream_t stream);
hipError_t launch_resample(const DevParams& P, int space, const float4* a, const float4* b, float4* out4,
                           float* out_rgb, int tiled, hipStream_t stream);
// Up to kMaxResampleJobs independent launch_resample jobs in one launch (no job may write what another
// reads).  c (space 4, with b): the job re-samples a + (b - c), each texel's difference formed first.
struct ResampleJob {
  int space;
  const float4* a;
  const float4* b;
  float4* out4;
  float* out_rgb;
  int tiled;
  const float4* c;
};
constexpr int kMaxResampleJobs = 4;
hipError_t launch_resamples(const DevParams& P, const ResampleJob* jobs, int count, hipStream_t stream);
// c = a + b (float4 tables)
hipError_t launch_add4(const float4* a, const float4* b, float4* c, long long count, hipStream_t stream);
// scratch[4] <- max_i max_c |x - y| / max(|a + x|, fraction * max_i |a + x|_c) (scratch[0..2]: the
// channel maxima); scratch holds residual_scratch_floats(count) floats.  Block reductions in two passes,
// no atomics, no memset.
size_t residual_scratch_floats(long long count);
hipError_t launch_residual(const float4* a, const float4* x, const float4* y, long long count, float fraction,
                           float* scratch, hipStream_t stream);
// Pack a float4 grid-order table into RGB floats (grid order), for the stage entry points.
// p[0 .. n) = 0 (a kernel, so that graph captures hold no memset node).
hipError_t launch_fill_zero(float* p, long long n, hipStream_t stream);
hipError_t launch_pack_rgb(const float4* in, float* out, long long count, hipStream_t stream);
hipError_t launch_unpack_rgb(const float* in, float4* out, long long count, hipStream_t stream);

}  // namespace sfsim
