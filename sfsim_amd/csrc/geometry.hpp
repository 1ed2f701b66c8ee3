// Index spaces and ray geometry of sfsim's atmosphere tables, in FP64, usable on host and device.
//
// The table spaces follow src/clj/sfsim/atmosphere.clj:233-422; the intersection distances follow
// the case analysis of ray-sphere-intersection (sphere.clj:46-59) but are evaluated in forms that
// do not cancel at Earth scale (radius 6.4e6 m, heights of metres matter): squared radii are never
// subtracted, c = R_s^2 - r^2 is formed as a product of a height difference and a radius sum.
//
// Every grid point of every table lies on the +x axis (index-to-height returns (r 0 0)), view
// directions lie in the x-y plane and the sphere-quadrature frame of such a point is fixed; the
// kernels rely on that and keep points as (radius, height) plus dot products.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define SFSIM_HD __host__ __device__ inline
// Every function here rounds each product and sum on its own, as the reference's FP64 (JVM) arithmetic
// does: no fused multiply-adds.  Where a formula cancels at a singular point (the square root of
// Delta = (r sin e)^2 - rho^2 at the ground-tangent direction turns rounding noise of 1e-16 into index
// offsets of 1e-8), a fused and an unfused evaluation differ, and on a coarse grid a 1e-8 weight on a
// neighbour thousands of times larger is visible; with the same rounding as the reference the index
// round trips of the re-sampling are the reference's, bit for bit.
#define SFSIM_EXACT _Pragma("clang fp contract(off)")

namespace sfsim {

struct Geo {
  double R;       // ground radius
  double H;       // atmosphere thickness (:sfsim.atmosphere/height)
  double Rt;      // top radius R + H
  double rho_max; // horizon distance at the top, sqrt(Rt^2 - R^2)
};

SFSIM_HD double sqr(double x) {
  SFSIM_EXACT
  return x * x;
}

// util.clj:403-416
SFSIM_HD double limit_quot(double a, double b, double lower, double upper) {
  SFSIM_EXACT
  if (a == 0.0) return a;
  if (b < 0) {
    a = -a;
    b = -b;
  }
  if (a < b * upper) return a > b * lower ? a / b : lower;
  return upper;
}

// atmosphere.clj:233-236
SFSIM_HD double horizon_distance(const Geo& g, double radius) {
  SFSIM_EXACT
  return sqrt(fmax(0.0, sqr(radius) - sqr(g.R)));
}

// atmosphere.clj:282-288: grid point (r 0 0) of height index `index`
SFSIM_HD double index_to_radius(const Geo& g, int size, double index) {
  SFSIM_EXACT
  double horizon_dist = (index / (size - 1)) * g.rho_max;
  return sqrt(sqr(g.R) + sqr(horizon_dist));
}

// Height above ground of the grid point of index_to_radius, without the r - R cancellation.
SFSIM_HD double index_to_altitude(const Geo& g, int size, double index) {
  SFSIM_EXACT
  double horizon_dist = (index / (size - 1)) * g.rho_max;
  double r = sqrt(sqr(g.R) + sqr(horizon_dist));
  return sqr(horizon_dist) / (r + g.R);
}

// atmosphere.clj:273-279 for a point of radius r
SFSIM_HD double height_to_index(const Geo& g, int size, double r) {
  SFSIM_EXACT
  return (size - 1) * (horizon_distance(g, r) / horizon_distance(g, g.R + g.H));
}

// Same for a point given by its altitude h (r = R + h): r^2 - R^2 = h (2R + h).
SFSIM_HD double altitude_to_index(const Geo& g, int size, double h) {
  SFSIM_EXACT
  return (size - 1) * (sqrt(fmax(0.0, h * (2.0 * g.R + h))) / horizon_distance(g, g.R + g.H));
}

// atmosphere.clj:256-270: view direction (sin_e, cos_e, 0) and above-horizon flag of an elevation index
SFSIM_HD void index_to_elevation(const Geo& g, int size, double radius, double index, double& sin_e, double& cos_e,
                                 bool& above) {
  SFSIM_EXACT
  double top_radius = g.Rt;
  double horizon_dist = horizon_distance(g, radius);
  double Hh = sqrt(sqr(top_radius) - sqr(g.R));
  double scaled_index = index / (size - 1);
  if ((scaled_index < 0.5) || ((scaled_index == 0.5) && (2.0 * radius < g.R + top_radius))) {
    double ground_dist = horizon_dist * (1 - 2 * scaled_index);
    sin_e = limit_quot((sqr(g.R) - sqr(radius)) - sqr(ground_dist), 2 * radius * ground_dist, -1.0, 1.0);
    above = false;
  } else {
    double sky_dist = (horizon_dist + Hh) * (2 * scaled_index - 1);
    sin_e = limit_quot((sqr(top_radius) - sqr(radius)) - sqr(sky_dist), 2 * radius * sky_dist, -1.0, 1.0);
    above = true;
  }
  cos_e = sqrt(1 - sqr(sin_e));
}

// atmosphere.clj:239-253 for a point of radius `radius` and dot(point, direction) = pd
SFSIM_HD double elevation_to_index(const Geo& g, int size, double radius, double pd, bool above) {
  SFSIM_EXACT
  double sin_elevation = pd / radius;
  double rho = horizon_distance(g, radius);
  double Delta = sqr(radius * sin_elevation) - sqr(rho);
  double Hh = sqrt(sqr(g.Rt) - sqr(g.R));
  double f;
  if (above)
    f = 0.5 - limit_quot(radius * sin_elevation - sqrt(fmax(0.0, Delta + sqr(Hh))), (rho + rho) + 2 * Hh, -0.5, 0.0);
  else
    f = 0.5 + limit_quot(radius * sin_elevation + sqrt(fmax(0.0, Delta)), rho + rho, -0.5, 0.0);
  return (size - 1) * f;
}

// atmosphere.clj:322-326 with sin_elevation = dot(point, l) / |point|
SFSIM_HD double sun_elevation_to_index(int size, double sin_elevation) {
  SFSIM_EXACT
  return (size - 1) * fmax(0.0, (1 - exp((0 - 3 * sin_elevation) - 0.6)) / (1 - exp(-3.6)));
}

// atmosphere.clj:329-332
SFSIM_HD double index_to_sin_sun_elevation(int size, double index) {
  SFSIM_EXACT
  return (log(1 - (index / (size - 1)) * (1 - exp(-3.6))) + 0.6) / -3;
}

// atmosphere.clj:368-372 with vl = dot(direction, l)
SFSIM_HD double sun_angle_to_index(int size, double vl) {
  SFSIM_EXACT
  return (size - 1) * ((1 + vl) / 2);
}

// atmosphere.clj:375-384 for a view direction (d0, d1, *)
SFSIM_HD void index_to_sun_direction(int size, double d0, double d1, double sin_sun, double index, double& l0,
                                     double& l1, double& l2) {
  SFSIM_EXACT
  double dot_view_sun = 2.0 * (index / (size - 1)) - 1.0;
  double max_sun_1 = sqrt(fmax(0.0, 1.0 - sqr(sin_sun)));
  double sun_1 = limit_quot(dot_view_sun - d0 * sin_sun, d1, -max_sun_1, max_sun_1);
  l0 = sin_sun;
  l1 = sun_1;
  l2 = sqrt(fmax(0.0, (1.0 - sqr(sun_1)) - sqr(sin_sun)));
}

// Exit distance from a point at altitude h (radius r = R + h) along a unit direction with
// dot(point, direction) = pd, through the top sphere: dist + length of ray-sphere-intersection
// (sphere.clj:46-59) for the sphere of radius Rt.  c = Rt^2 - r^2 formed without cancellation.
SFSIM_HD double top_exit_distance(const Geo& g, double h, double pd) {
  SFSIM_EXACT
  double c = (g.H - h) * (2.0 * g.R + g.H + h);
  double disc = pd * pd + c;
  if (disc > 0) {
    double l2 = sqrt(disc);
    double d = pd > 0 ? c / (pd + l2) : l2 - pd;  // -pd + l2 in both cases
    return fmax(0.0, d);
  }
  return fmax(0.0, -pd);
}

// Distance to the ground along a unit direction from altitude h with dot(point, direction) = pd:
// distance of ray-sphere-intersection for the ground sphere (sphere.clj:46-59), i.e. the near
// intersection, or the closest approach when the ray misses.
SFSIM_HD double ground_distance(const Geo& g, double h, double pd) {
  SFSIM_EXACT
  double c = h * (2.0 * g.R + h);  // r^2 - R^2
  double disc = pd * pd - c;
  double middle = -pd;
  if (disc > 0) {
    double l2 = sqrt(disc);
    if (middle < l2) return 0.0;
    return c / (middle + l2);  // middle - l2, rationalised
  }
  return fmax(0.0, middle);
}

// atmosphere.clj:95-102 for a point of altitude h and dot(direction, point) = pd
SFSIM_HD bool is_above_horizon(const Geo& g, double h, double pd) {
  SFSIM_EXACT
  return (pd >= 0) || (sqr(pd) <= h * (2.0 * g.R + h));
}

// Altitude after travelling t along a unit direction from altitude h0 with dot(point, direction) = pd0:
// r(t)^2 - R^2 = h0 (2R + h0) + t (2 pd0 + t),  h = (r^2 - R^2) / (r + R).
SFSIM_HD double altitude_along(const Geo& g, double h0, double pd0, double t) {
  SFSIM_EXACT
  double u = h0 * (2.0 * g.R + h0) + t * (2.0 * pd0 + t);
  return u / (g.R + sqrt(sqr(g.R) + u));
}

}  // namespace sfsim
