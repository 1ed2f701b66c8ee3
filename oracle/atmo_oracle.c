/*
 * FP64 CPU oracle for sfsim's atmospheric lookup-table precompute.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library's
 * results are compared against.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg load it; the product library never links,
 * loads or calls it, and has no CPU fallback through it.
 *
 * It is a literal restatement of the reference's Clojure functions, in IEEE
 * double precision, with the same evaluation order where the order is
 * observable (sums are accumulated in the order the reference reduces them).
 * Every function names the reference source it follows (paths relative to the
 * reference repository root).
 *
 * Pinning (what checks each function against something other than itself):
 *   - scattering/extinction/phase/height/intersections/horizon/transmittance,
 *     integral-ray, integral-sphere, index spaces, limit-quot, interpolation,
 *     4D->2D tiling: the reference's own known-answer tests, transcribed into
 *     tests/golden/reference_goldens.json and checked by tests/test_oracle_*.
 *   - first-order tables (T, R1) and the surface-radiance integral: the
 *     reference's table-level probes at size 12 (t_atmosphere.clj:629-775,
 *     t_planet.clj:312-376), evaluated through the GLSL-consumer restatement
 *     at the end of this file.
 *   - point-scatter (J): no reference-held vector; pinned by a brute-force
 *     physics evaluator that computes the first-order radiance and the ground
 *     irradiance directly (tests/test_oracle_bruteforce.py: the difference is
 *     interpolation error and shrinks with the grid) and by invariants
 *     (closed-form ground and constant-radiance cases, scaling in the
 *     scattering coefficient, the optically thin limit;
 *     tests/test_oracle_invariants.py).
 *   - order 2 of the driver: the same brute-force evaluator carried one order
 *     further (tests/test_oracle_bruteforce.py).
 *   - the multi-order driver (orc_generate_atmosphere_luts): the structure the
 *     truncated series must have when the scattering coefficients (extinction
 *     fixed) and the albedo are scaled (tests/test_oracle_driver_pin.py): the
 *     first iteration adds exactly the brute-force-pinned second order, every
 *     further one a single monomial of the next degree, and the albedo enters
 *     with the degree the ground reflections fix.  Mis-wired drivers (a stage
 *     reading the wrong order's table, E0 kept in E, S accumulating dJ) fail it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } v3;

typedef struct {
  double centre[3];
  double radius;
  double height;        /* :sfsim.atmosphere/height (atmosphere thickness) */
  double brightness[3]; /* :sfsim.atmosphere/brightness (ground albedo) */
} orc_planet;

typedef struct {
  double base[3];  /* :sfsim.atmosphere/scatter-base */
  double scale;    /* :sfsim.atmosphere/scatter-scale */
  double quotient; /* :sfsim.atmosphere/scatter-quotient (1.0 when absent) */
  double g;        /* :sfsim.atmosphere/scatter-g (0.0 when absent) */
} orc_scatter;

/* Parameters of one LUT build (atmosphere_lut.clj:20-62). */
typedef struct {
  orc_planet planet;
  orc_scatter mie;
  orc_scatter rayleigh;
  double intensity[3];
  int transmittance_shape[2];    /* [height elevation] */
  int surface_radiance_shape[2]; /* [height sun-elevation] */
  int scatter_shape[4];          /* [height elevation sun-elevation heading] */
  int ray_steps;
  int sphere_steps;
  int iterations;
} orc_params;

/* ------------------------------------------------------------------ vec3 */
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vemul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vmag(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vnormalize(v3 a) { return vdiv(a, vmag(a)); }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 vexp_neg(v3 a) { return V(exp(-a.x), exp(-a.y), exp(-a.z)); }
static inline v3 arr3(const double* a) { return V(a[0], a[1], a[2]); }
static inline v3 centre_of(const orc_planet* p) { return arr3(p->centre); }

/* util.clj:334-337 */
double orc_sqr(double x) { return x * x; }

/* util.clj:403-416 (4-arity; the 3-arity passes -limit, limit) */
double orc_limit_quot(double a, double b, double lower, double upper) {
  if (a == 0.0) return a;
  if (b < 0) { a = -a; b = -b; }
  if (a < b * upper) {
    if (a > b * lower) return a / b;
    return lower;
  }
  return upper;
}

/* ----------------------------------------------------------- atmosphere */
/* atmosphere.clj:42-47 (component, height arity) */
static v3 scattering_h(const orc_scatter* c, double height) {
  return vmul(arr3(c->base), exp(-(height / c->scale)));
}

/* sphere.clj:28-31 */
double orc_height(const orc_planet* p, v3 point) { return vmag(vsub(point, centre_of(p))) - p->radius; }

/* atmosphere.clj:42-47 (planet, component, x arity) */
static v3 scattering_x(const orc_planet* p, const orc_scatter* c, v3 x) { return scattering_h(c, orc_height(p, x)); }

/* atmosphere.clj:50-53 */
static v3 extinction(const orc_scatter* c, double height) { return vdiv(scattering_h(c, height), c->quotient); }

/* atmosphere.clj:56-61 */
double orc_phase(double g, double mu) {
  double g2 = orc_sqr(g);
  return (3.0 * (1.0 - g2) * (1.0 + orc_sqr(mu))) /
         (8.0 * M_PI * (2.0 + g2) * pow((1.0 + g2) - 2.0 * g * mu, 1.5));
}

/* sphere.clj:34-59: ray-sphere-determinant + ray-sphere-intersection */
void orc_ray_sphere_intersection(v3 centre, double radius, v3 origin, v3 direction, double* distance, double* length) {
  v3 offset = vsub(origin, centre);
  double direction_sqr = vdot(direction, direction);
  double discriminant = orc_sqr(vdot(direction, offset)) - direction_sqr * (vdot(offset, offset) - orc_sqr(radius));
  double middle = -(vdot(direction, offset) / direction_sqr);
  if (discriminant > 0) {
    double length2 = sqrt(discriminant) / direction_sqr;
    if (middle < length2) {
      *distance = 0.0;
      *length = fmax(0.0, middle + length2);
    } else {
      *distance = middle - length2;
      *length = 2 * length2;
    }
  } else {
    *distance = fmax(0.0, middle);
    *length = 0.0;
  }
}

/* atmosphere.clj:70-77 */
static v3 atmosphere_intersection(const orc_planet* p, v3 origin, v3 direction) {
  double d, l;
  orc_ray_sphere_intersection(centre_of(p), p->radius + p->height, origin, direction, &d, &l);
  return vadd(origin, vmul(direction, d + l));
}

/* atmosphere.clj:80-85 */
static v3 surface_intersection(const orc_planet* p, v3 origin, v3 direction) {
  double d, l;
  orc_ray_sphere_intersection(centre_of(p), p->radius, origin, direction, &d, &l);
  return vadd(origin, vmul(direction, d));
}

/* atmosphere.clj:88-92 */
static int surface_point(const orc_planet* p, v3 point) { return 2.0 * orc_height(p, point) < p->height; }

/* atmosphere.clj:95-102 (note: uses |point|, i.e. assumes the centre at the origin) */
static int is_above_horizon(const orc_planet* p, v3 point, v3 direction) {
  double norm_point = vmag(point);
  double sin_elevation_radius = vdot(direction, point);
  double horizon_distance_sqr = orc_sqr(norm_point) - orc_sqr(p->radius);
  return (sin_elevation_radius >= 0) || (orc_sqr(sin_elevation_radius) <= horizon_distance_sqr);
}

/* atmosphere.clj:105-111 */
static v3 ray_extremity(const orc_planet* p, v3 origin, v3 direction) {
  if (is_above_horizon(p, origin, direction)) return atmosphere_intersection(p, origin, direction);
  return surface_intersection(p, origin, direction);
}

/* ray.clj:19-30: midpoint rule, samples at (i+1/2)*distance/steps, weight stepsize*|direction| */
typedef v3 (*ray_fn)(const void* ctx, v3 point);
static v3 integral_ray(v3 origin, v3 direction, int steps, double distance, ray_fn fun, const void* ctx) {
  double stepsize = distance / steps;
  double direction_len = vmag(direction);
  double a = stepsize * direction_len;
  v3 acc = V(0, 0, 0);
  for (int n = 0; n < steps; n++) {
    double s = (0.5 + n) * stepsize;
    v3 point = vadd(origin, vmul(direction, s));
    acc = vadd(acc, vmul(fun(ctx, point), a));
  }
  return acc;
}

/* Overall extinction of atmosphere.clj:114-125: one component, or the sum of the first two. */
typedef struct { const orc_planet* planet; const orc_scatter* scatter; int nscatter; } ext_ctx;
static v3 overall_extinction(const void* vctx, v3 point) {
  const ext_ctx* c = (const ext_ctx*)vctx;
  if (c->nscatter == 1) return extinction(&c->scatter[0], orc_height(c->planet, point));
  return vadd(extinction(&c->scatter[0], orc_height(c->planet, point)),
              extinction(&c->scatter[1], orc_height(c->planet, point)));
}

/* atmosphere.clj:114-125 (planet scatter steps x x0) */
static v3 transmittance_points(const orc_planet* p, const orc_scatter* s, int ns, int steps, v3 x, v3 x0) {
  ext_ctx c = {p, s, ns};
  return vexp_neg(integral_ray(x, vsub(x0, x), steps, 1.0, overall_extinction, &c));
}

/* atmosphere.clj:126-128 (planet scatter steps x v above-horizon) */
static v3 transmittance_dir(const orc_planet* p, const orc_scatter* s, int ns, int steps, v3 x, v3 v, int above) {
  v3 end = above ? atmosphere_intersection(p, x, v) : surface_intersection(p, x, v);
  return transmittance_points(p, s, ns, steps, x, end);
}

/* atmosphere.clj:131-137 (E0) */
static v3 surface_radiance_base(const orc_planet* p, const orc_scatter* s, int ns, int steps, v3 intensity, v3 x, v3 l) {
  v3 normal = vnormalize(vsub(x, centre_of(p)));
  return vmul(vemul(transmittance_dir(p, s, ns, steps, x, l, 1), intensity), fmax(0.0, vdot(normal, l)));
}

/* atmosphere.clj:140-151 */
static v3 overall_in_scattering(const orc_planet* p, const orc_scatter* comps, int ncomps, v3 x, v3 v, v3 l) {
  v3 acc = vmul(scattering_x(p, &comps[0], x), orc_phase(comps[0].g, vdot(v, l)));
  for (int i = 1; i < ncomps; i++)
    acc = vadd(acc, vmul(scattering_x(p, &comps[i], x), orc_phase(comps[i].g, vdot(v, l))));
  return acc;
}

/* atmosphere.clj:154-160 */
static v3 filtered_sun_light(const orc_planet* p, const orc_scatter* s, int ns, int steps, v3 x, v3 l, v3 intensity) {
  if (is_above_horizon(p, x, l)) return vemul(intensity, transmittance_dir(p, s, ns, steps, x, l, 1));
  return V(0, 0, 0);
}

/* ------------------------------------------------- callback signatures */
typedef v3 (*scatter_fn)(const void* ctx, v3 x, v3 v, v3 l, int above);   /* J or S */
typedef v3 (*surface_fn)(const void* ctx, v3 x, v3 l);                    /* E */

/* Single-scatter J of atmosphere.clj:162-189 */
typedef struct {
  const orc_planet* planet;
  const orc_scatter* scatter;
  int nscatter;
  const orc_scatter* component;
  int steps;
  v3 intensity;
  int strength; /* 0: point-scatter-component (with phase), 1: strength-component (no phase) */
} single_ctx;

static v3 single_scatter_J(const void* vctx, v3 x, v3 v, v3 l, int above) {
  (void)above;
  const single_ctx* c = (const single_ctx*)vctx;
  v3 sun = filtered_sun_light(c->planet, c->scatter, c->nscatter, c->steps, x, l, c->intensity);
  if (c->strength) return vemul(scattering_x(c->planet, c->component, x), sun); /* atmosphere.clj:176-181 */
  return vemul(overall_in_scattering(c->planet, c->component, 1, x, v, l), sun); /* atmosphere.clj:169-173 */
}

/* atmosphere.clj:192-200 */
typedef struct {
  const orc_planet* planet;
  const orc_scatter* scatter;
  int nscatter;
  int steps;
  v3 x;
  scatter_fn J;
  const void* J_ctx;
  v3 v, l;
  int above;
} rs_ctx;

static v3 ray_scatter_integrand(const void* vctx, v3 p) {
  const rs_ctx* c = (const rs_ctx*)vctx;
  return vemul(transmittance_points(c->planet, c->scatter, c->nscatter, c->steps, c->x, p),
               c->J(c->J_ctx, p, c->v, c->l, c->above));
}

static v3 ray_scatter(const orc_planet* p, const orc_scatter* s, int ns, int steps, scatter_fn J, const void* J_ctx,
                      v3 x, v3 v, v3 l, int above) {
  v3 point = above ? atmosphere_intersection(p, x, v) : surface_intersection(p, x, v);
  rs_ctx c = {p, s, ns, steps, x, J, J_ctx, v, l, above};
  return integral_ray(x, vsub(point, x), steps, 1.0, ray_scatter_integrand, &c);
}

/* ------------------------------------------------------ sphere quadrature */
/* quaternion.clj:166-171: basis vector with the smallest |n.e| (stable sort -> lowest index on ties) */
static v3 orthogonal(v3 n) {
  v3 rows[3] = {V(1, 0, 0), V(0, 1, 0), V(0, 0, 1)};
  int best = 0;
  double best_val = fabs(vdot(n, rows[0]));
  for (int i = 1; i < 3; i++) {
    double val = fabs(vdot(n, rows[i]));
    if (val < best_val) { best = i; best_val = val; }
  }
  return vnormalize(vcross(n, rows[best]));
}

/* matrix.clj:207-213: rows (n, o1, o2) */
void orc_oriented_matrix(const double* nvec, double* m) {
  v3 n = arr3(nvec);
  v3 o1 = orthogonal(n);
  v3 o2 = vcross(n, o1);
  double r[9] = {n.x, n.y, n.z, o1.x, o1.y, o1.z, o2.x, o2.y, o2.z};
  memcpy(m, r, sizeof(r));
}

typedef v3 (*dir_fn)(const void* ctx, v3 direction);

/* sphere.clj:62-67 (fun of the angle phi, inlined: the caller evaluates the ring point) */
/* sphere.clj:70-93 */
static v3 spherical_integral(int theta_steps, int phi_steps, double theta_range, v3 normal, dir_fn fun, const void* ctx) {
  double delta2 = theta_range / theta_steps / 2;
  double m[9];
  double nv[3] = {normal.x, normal.y, normal.z};
  orc_oriented_matrix(nv, m);
  /* mat = transpose(oriented-matrix): mat * (x y z) = x*n + y*o1 + z*o2 */
  v3 acc = V(0, 0, 0);
  for (int i = 0; i < theta_steps; i++) {
    double theta = theta_range * ((0.5 + i) / theta_steps);
    double factor = cos(theta - delta2) - cos(theta + delta2);
    int ringsteps = (int)ceil(sin(theta) * phi_steps);
    double cos_theta = cos(theta), sin_theta = sin(theta);
    /* integrate-circle (sphere.clj:62-67) */
    double weight = (2 * M_PI) / ringsteps;
    v3 ring = V(0, 0, 0);
    for (int j = 0; j < ringsteps; j++) {
      double phi = 2 * M_PI * ((0.5 + j) / ringsteps);
      double x = cos_theta, y = sin_theta * cos(phi), z = sin_theta * sin(phi);
      v3 d = V(m[0] * x + m[3] * y + m[6] * z, m[1] * x + m[4] * y + m[7] * z, m[2] * x + m[5] * y + m[8] * z);
      ring = vadd(ring, fun(ctx, d));
    }
    acc = vadd(acc, vmul(vmul(ring, weight), factor));
  }
  return acc;
}

/* sphere.clj:96-99 */
static v3 integral_half_sphere(int steps, v3 normal, dir_fn fun, const void* ctx) {
  return spherical_integral(steps >> 2, steps, M_PI / 2, normal, fun, ctx);
}

/* sphere.clj:102-105 */
static v3 integral_sphere(int steps, v3 normal, dir_fn fun, const void* ctx) {
  return spherical_integral(steps >> 1, steps, M_PI, normal, fun, ctx);
}

/* ---------------------------------------------------- point-scatter (J) */
/* atmosphere.clj:203-222 */
typedef struct {
  const orc_planet* planet;
  const orc_scatter* scatter;
  int nscatter;
  scatter_fn S;
  const void* S_ctx;
  surface_fn E;
  const void* E_ctx;
  int ray_steps;
  v3 x, v, l;
} ps_ctx;

static v3 in_scatter_from_direction(const void* vctx, v3 omega) {
  const ps_ctx* c = (const ps_ctx*)vctx;
  v3 point = ray_extremity(c->planet, c->x, omega);
  int surface = surface_point(c->planet, point);
  v3 overall = overall_in_scattering(c->planet, c->scatter, c->nscatter, c->x, c->v, omega);
  v3 ground = V(0, 0, 0);
  if (surface) {
    v3 brightness = vemul(vdiv(arr3(c->planet->brightness), M_PI), c->E(c->E_ctx, point, c->l));
    ground = vemul(transmittance_points(c->planet, c->scatter, c->nscatter, c->ray_steps, c->x, point), brightness);
  }
  return vemul(overall, vadd(c->S(c->S_ctx, c->x, omega, c->l, !surface), ground));
}

static v3 point_scatter(const orc_planet* p, const orc_scatter* s, int ns, scatter_fn S, const void* S_ctx,
                        surface_fn E, const void* E_ctx, int sphere_steps, int ray_steps, v3 x, v3 v, v3 l) {
  v3 normal = vnormalize(vsub(x, centre_of(p)));
  ps_ctx c = {p, s, ns, S, S_ctx, E, E_ctx, ray_steps, x, v, l};
  return integral_sphere(sphere_steps, normal, in_scatter_from_direction, &c);
}

/* atmosphere.clj:225-230 */
typedef struct { scatter_fn S; const void* S_ctx; v3 x, l, normal; } sr_ctx;
static v3 surface_radiance_integrand(const void* vctx, v3 omega) {
  const sr_ctx* c = (const sr_ctx*)vctx;
  return vmul(c->S(c->S_ctx, c->x, omega, c->l, 1), vdot(omega, c->normal));
}

static v3 surface_radiance(const orc_planet* p, scatter_fn S, const void* S_ctx, int steps, v3 x, v3 l) {
  v3 normal = vnormalize(vsub(x, centre_of(p)));
  sr_ctx c = {S, S_ctx, x, l, normal};
  return integral_half_sphere(steps, normal, surface_radiance_integrand, &c);
}

/* --------------------------------------------------------- index spaces */
/* atmosphere.clj:233-236 */
double orc_horizon_distance(const orc_planet* p, double radius) {
  return sqrt(fmax(0.0, orc_sqr(radius) - orc_sqr(p->radius)));
}

/* atmosphere.clj:239-253 */
static double elevation_to_index(const orc_planet* p, int size, v3 point, v3 direction, int above) {
  double radius = vmag(point);
  double ground_radius = p->radius;
  double top_radius = ground_radius + p->height;
  double sin_elevation = vdot(point, direction) / radius;
  double rho = orc_horizon_distance(p, radius);
  double Delta = orc_sqr(radius * sin_elevation) - orc_sqr(rho);
  double H = sqrt(orc_sqr(top_radius) - orc_sqr(ground_radius));
  double f;
  if (above)
    f = 0.5 - orc_limit_quot(radius * sin_elevation - sqrt(fmax(0.0, Delta + orc_sqr(H))), (rho + rho) + 2 * H, -0.5, 0.0);
  else
    f = 0.5 + orc_limit_quot(radius * sin_elevation + sqrt(fmax(0.0, Delta)), rho + rho, -0.5, 0.0);
  return (size - 1) * f;
}

/* atmosphere.clj:256-270 */
static void index_to_elevation(const orc_planet* p, int size, double radius, double index, v3* direction, int* above) {
  double ground_radius = p->radius;
  double top_radius = ground_radius + p->height;
  double horizon_dist = orc_horizon_distance(p, radius);
  double H = sqrt(orc_sqr(top_radius) - orc_sqr(ground_radius));
  double scaled_index = index / (size - 1);
  if ((scaled_index < 0.5) || ((scaled_index == 0.5) && (2.0 * radius < ground_radius + top_radius))) {
    double ground_dist = horizon_dist * (1 - 2 * scaled_index);
    double sin_elevation = orc_limit_quot((orc_sqr(ground_radius) - orc_sqr(radius)) - orc_sqr(ground_dist),
                                          2 * radius * ground_dist, -1.0, 1.0);
    *direction = V(sin_elevation, sqrt(1 - orc_sqr(sin_elevation)), 0);
    *above = 0;
  } else {
    double sky_dist = (horizon_dist + H) * (2 * scaled_index - 1);
    double sin_elevation = orc_limit_quot((orc_sqr(top_radius) - orc_sqr(radius)) - orc_sqr(sky_dist),
                                          2 * radius * sky_dist, -1.0, 1.0);
    *direction = V(sin_elevation, sqrt(1 - orc_sqr(sin_elevation)), 0);
    *above = 1;
  }
}

/* atmosphere.clj:273-279 */
static double height_to_index(const orc_planet* p, int size, v3 point) {
  return (size - 1) * (orc_horizon_distance(p, vmag(point)) / orc_horizon_distance(p, p->radius + p->height));
}

/* atmosphere.clj:282-288 */
static v3 index_to_height(const orc_planet* p, int size, double index) {
  double max_horizon = sqrt(orc_sqr(p->radius + p->height) - orc_sqr(p->radius));
  double horizon_dist = (index / (size - 1)) * max_horizon;
  return V(sqrt(orc_sqr(p->radius) + orc_sqr(horizon_dist)), 0, 0);
}

/* atmosphere.clj:322-326 */
static double sun_elevation_to_index(int size, v3 point, v3 l) {
  double sin_elevation = vdot(point, l) / vmag(point);
  return (size - 1) * fmax(0.0, (1 - exp((0 - 3 * sin_elevation) - 0.6)) / (1 - exp(-3.6)));
}

/* atmosphere.clj:329-332 */
static double index_to_sin_sun_elevation(int size, double index) {
  return (log(1 - (index / (size - 1)) * (1 - exp(-3.6))) + 0.6) / -3;
}

/* atmosphere.clj:368-372 */
static double sun_angle_to_index(int size, v3 direction, v3 l) { return (size - 1) * ((1 + vdot(direction, l)) / 2); }

/* atmosphere.clj:375-384 */
static v3 index_to_sun_direction(int size, v3 direction, double sin_sun_elevation, double index) {
  double dot_view_sun = 2.0 * (index / (size - 1)) - 1.0;
  double max_sun_1 = sqrt(fmax(0.0, 1.0 - orc_sqr(sin_sun_elevation)));
  double sun_1 = orc_limit_quot(dot_view_sun - direction.x * sin_sun_elevation, direction.y, -max_sun_1, max_sun_1);
  double sun_2 = sqrt(fmax(0.0, (1.0 - orc_sqr(sun_1)) - orc_sqr(sin_sun_elevation)));
  return V(sin_sun_elevation, sun_1, sun_2);
}

/* atmosphere.clj:291-319: transmittance-space forward/backward */
static void transmittance_forward(const orc_planet* p, const int* shape, v3 point, v3 direction, int above, double* idx) {
  idx[0] = height_to_index(p, shape[0], point);
  idx[1] = elevation_to_index(p, shape[1], point, direction, above);
}
static void transmittance_backward(const orc_planet* p, const int* shape, const double* idx, v3* point, v3* direction,
                                   int* above) {
  *point = index_to_height(p, shape[0], idx[0]);
  index_to_elevation(p, shape[1], point->x, idx[1], direction, above);
}

/* atmosphere.clj:335-365: surface-radiance-space forward/backward */
static void surface_radiance_forward(const orc_planet* p, const int* shape, v3 point, v3 l, double* idx) {
  idx[0] = height_to_index(p, shape[0], point);
  idx[1] = sun_elevation_to_index(shape[1], point, l);
}
static void surface_radiance_backward(const orc_planet* p, const int* shape, const double* idx, v3* point, v3* l) {
  *point = index_to_height(p, shape[0], idx[0]);
  double s = index_to_sin_sun_elevation(shape[1], idx[1]);
  double c = sqrt(fmax(0.0, 1 - orc_sqr(s)));
  *l = V(s, c, 0);
}

/* atmosphere.clj:387-422: ray-scatter-space (= point-scatter-space) forward/backward */
static void scatter_forward(const orc_planet* p, const int* shape, v3 point, v3 direction, v3 l, int above, double* idx) {
  idx[0] = height_to_index(p, shape[0], point);
  idx[1] = elevation_to_index(p, shape[1], point, direction, above);
  idx[2] = sun_elevation_to_index(shape[2], point, l);
  idx[3] = sun_angle_to_index(shape[3], direction, l);
}
static void scatter_backward(const orc_planet* p, const int* shape, const double* idx, v3* point, v3* direction, v3* l,
                             int* above) {
  *point = index_to_height(p, shape[0], idx[0]);
  index_to_elevation(p, shape[1], point->x, idx[1], direction, above);
  double s = index_to_sin_sun_elevation(shape[2], idx[2]);
  *l = index_to_sun_direction(shape[3], *direction, s, idx[3]);
}

/* -------------------------------------------------------- interpolation */
/* interpolate.clj:75-78 */
static inline double clip(double value, int size) { return fmin(fmax(value, 0.0), size - 1); }

/* interpolate.clj:87-98: recursive multilinear, outer dimension first; mix(a,b,s)=a(1-s)+bs */
static v3 interpolate_value(const double* data, const int* shape, int ndim, const double* point) {
  if (ndim == 0) return V(data[0], data[1], data[2]);
  int stride = 3;
  for (int d = 1; d < ndim; d++) stride *= shape[d];
  int size = shape[0];
  double i = clip(point[0], size);
  double u = floor(i);
  double v = clip(u + 1, size);
  double s = i - u;
  v3 a = interpolate_value(data + (size_t)u * stride, shape + 1, ndim - 1, point + 1);
  v3 b = interpolate_value(data + (size_t)v * stride, shape + 1, ndim - 1, point + 1);
  return vadd(vmul(a, 1 - s), vmul(b, s));
}

/* Interpolated-table functions (interpolate.clj:101-110 interpolation-table with a space's forward) */
typedef struct { const orc_planet* planet; const int* shape; const double* data; } table_ctx;

static v3 interp_scatter(const void* vctx, v3 x, v3 v, v3 l, int above) {
  const table_ctx* c = (const table_ctx*)vctx;
  double idx[4];
  scatter_forward(c->planet, c->shape, x, v, l, above, idx);
  return interpolate_value(c->data, c->shape, 4, idx);
}

static v3 interp_surface(const void* vctx, v3 x, v3 l) {
  const table_ctx* c = (const table_ctx*)vctx;
  double idx[2];
  surface_radiance_forward(c->planet, c->shape, x, l, idx);
  return interpolate_value(c->data, c->shape, 2, idx);
}

/* The first-iteration dS of atmosphere_lut.clj:79-84: R1(q) + M1(q) * phase(mie, v.l) */
typedef struct { table_ctx R1, M1; double mie_g; } first_dS_ctx;
static v3 first_dS(const void* vctx, v3 x, v3 v, v3 l, int above) {
  const first_dS_ctx* c = (const first_dS_ctx*)vctx;
  return vadd(interp_scatter(&c->R1, x, v, l, above),
              vmul(interp_scatter(&c->M1, x, v, l, above), orc_phase(c->mie_g, vdot(v, l))));
}

/* ------------------------------------------------------ exported helpers */
/* Exports used by the golden-vector tests; arguments are flat double arrays. */
void orc_scattering(const orc_scatter* c, double height, double* out) {
  v3 r = scattering_h(c, height); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_extinction(const orc_scatter* c, double height, double* out) {
  v3 r = extinction(c, height); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double orc_height_of(const orc_planet* p, const double* point) { return orc_height(p, arr3(point)); }
void orc_atmosphere_intersection(const orc_planet* p, const double* o, const double* d, double* out) {
  v3 r = atmosphere_intersection(p, arr3(o), arr3(d)); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_surface_intersection(const orc_planet* p, const double* o, const double* d, double* out) {
  v3 r = surface_intersection(p, arr3(o), arr3(d)); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int orc_surface_point(const orc_planet* p, const double* point) { return surface_point(p, arr3(point)); }
int orc_is_above_horizon(const orc_planet* p, const double* point, const double* d) {
  return is_above_horizon(p, arr3(point), arr3(d));
}
void orc_ray_extremity(const orc_planet* p, const double* o, const double* d, double* out) {
  v3 r = ray_extremity(p, arr3(o), arr3(d)); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_transmittance_points(const orc_planet* p, const orc_scatter* s, int ns, int steps, const double* x,
                              const double* x0, double* out) {
  v3 r = transmittance_points(p, s, ns, steps, arr3(x), arr3(x0)); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_transmittance_dir(const orc_planet* p, const orc_scatter* s, int ns, int steps, const double* x,
                           const double* v, int above, double* out) {
  v3 r = transmittance_dir(p, s, ns, steps, arr3(x), arr3(v), above); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_surface_radiance_base(const orc_planet* p, const orc_scatter* s, int ns, int steps, const double* intensity,
                               const double* x, const double* l, double* out) {
  v3 r = surface_radiance_base(p, s, ns, steps, arr3(intensity), arr3(x), arr3(l));
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* point-scatter-component (strength=0) / strength-component (strength=1) of atmosphere.clj:162-189 */
void orc_single_scatter_J(const orc_planet* p, const orc_scatter* s, int ns, const orc_scatter* component, int steps,
                          const double* intensity, int strength, const double* x, const double* v, const double* l,
                          double* out) {
  single_ctx c = {p, s, ns, component, steps, arr3(intensity), strength};
  v3 r = single_scatter_J(&c, arr3(x), arr3(v), arr3(l), 1);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* point-scatter-base of atmosphere.clj:184-189 */
void orc_point_scatter_base(const orc_planet* p, const orc_scatter* s, int ns, int steps, const double* intensity,
                            const double* x, const double* v, const double* l, double* out) {
  v3 sun = filtered_sun_light(p, s, ns, steps, arr3(x), arr3(l), arr3(intensity));
  v3 r = vemul(overall_in_scattering(p, s, ns, arr3(x), arr3(v), arr3(l)), sun);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* ray-scatter of a single-scatter J, for one point (atmosphere.clj:192-200) */
void orc_ray_scatter_single(const orc_planet* p, const orc_scatter* s, int ns, const orc_scatter* component, int steps,
                            const double* intensity, int strength, const double* x, const double* v, const double* l,
                            int above, double* out) {
  single_ctx c = {p, s, ns, component, steps, arr3(intensity), strength};
  v3 r = ray_scatter(p, s, ns, steps, single_scatter_J, &c, arr3(x), arr3(v), arr3(l), above);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* Sphere quadratures with simple test integrands: kind 0 = constant (1,1,1), 1 = direction itself,
 * 2 = (1, d_y, d_z), 3 = (d_x, 1, d_z) (t_sphere.clj:44-62). */
typedef struct { int kind; } test_integrand;
static v3 test_fn(const void* vctx, v3 d) {
  const test_integrand* c = (const test_integrand*)vctx;
  switch (c->kind) {
    case 0: return V(1, 1, 1);
    case 1: return d;
    case 2: return V(1, d.y, d.z);
    default: return V(d.x, 1, d.z);
  }
}
void orc_integral_sphere_test(int half, int steps, const double* normal, int kind, double* out) {
  test_integrand c = {kind};
  v3 r = half ? integral_half_sphere(steps, arr3(normal), test_fn, &c) : integral_sphere(steps, arr3(normal), test_fn, &c);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* Number of directions of a sphere quadrature (sum of ring sizes) */
int orc_sphere_direction_count(int half, int steps) {
  int theta_steps = half ? (steps >> 2) : (steps >> 1);
  double theta_range = half ? M_PI / 2 : M_PI;
  int total = 0;
  for (int i = 0; i < theta_steps; i++) total += (int)ceil(sin(theta_range * ((0.5 + i) / theta_steps)) * steps);
  return total;
}
/* integral-ray with integrand kind 0 = constant (2,2,0), 1 = (x_0, x_0, 0) (t_ray.clj:20-28) */
typedef struct { int kind; } ray_test;
static v3 ray_test_fn(const void* vctx, v3 p) {
  const ray_test* c = (const ray_test*)vctx;
  return c->kind == 0 ? V(2, 2, 0) : V(p.x, p.x, 0);
}
void orc_integral_ray_test(const double* origin, const double* direction, int steps, double distance, int kind,
                           double* out) {
  ray_test c = {kind};
  v3 r = integral_ray(arr3(origin), arr3(direction), steps, distance, ray_test_fn, &c);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

double orc_elevation_to_index(const orc_planet* p, int size, const double* point, const double* d, int above) {
  return elevation_to_index(p, size, arr3(point), arr3(d), above);
}
int orc_index_to_elevation(const orc_planet* p, int size, double radius, double index, double* direction) {
  v3 d; int above;
  index_to_elevation(p, size, radius, index, &d, &above);
  direction[0] = d.x; direction[1] = d.y; direction[2] = d.z;
  return above;
}
double orc_height_to_index(const orc_planet* p, int size, const double* point) {
  return height_to_index(p, size, arr3(point));
}
void orc_index_to_height(const orc_planet* p, int size, double index, double* out) {
  v3 r = index_to_height(p, size, index); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double orc_sun_elevation_to_index(int size, const double* point, const double* l) {
  return sun_elevation_to_index(size, arr3(point), arr3(l));
}
double orc_index_to_sin_sun_elevation(int size, double index) { return index_to_sin_sun_elevation(size, index); }
double orc_sun_angle_to_index(int size, const double* d, const double* l) {
  return sun_angle_to_index(size, arr3(d), arr3(l));
}
void orc_index_to_sun_direction(int size, const double* d, double sin_sun, double index, double* out) {
  v3 r = index_to_sun_direction(size, arr3(d), sin_sun, index); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* space: 2 = transmittance, 3 = surface radiance, 4 = ray/point scatter.
 * forward: in = [point(3) direction(3) light(3) above], out = indices.
 * backward: in = indices, out = [point(3) direction(3) light(3) above]. */
void orc_space_forward(const orc_planet* p, int space, const int* shape, const double* in, double* out) {
  v3 point = arr3(in), d = arr3(in + 3), l = arr3(in + 6);
  int above = in[9] != 0.0;
  if (space == 2) transmittance_forward(p, shape, point, d, above, out);
  else if (space == 3) surface_radiance_forward(p, shape, point, l, out);
  else scatter_forward(p, shape, point, d, l, above, out);
}
void orc_space_backward(const orc_planet* p, int space, const int* shape, const double* idx, double* out) {
  v3 point = V(0, 0, 0), d = V(0, 0, 0), l = V(0, 0, 0);
  int above = 0;
  if (space == 2) transmittance_backward(p, shape, idx, &point, &d, &above);
  else if (space == 3) surface_radiance_backward(p, shape, idx, &point, &l);
  else scatter_backward(p, shape, idx, &point, &d, &l, &above);
  double r[10] = {point.x, point.y, point.z, d.x, d.y, d.z, l.x, l.y, l.z, (double)above};
  memcpy(out, r, sizeof(r));
}
void orc_interpolate(const double* data, const int* shape, int ndim, const double* point, double* out) {
  v3 r = interpolate_value(data, shape, ndim, point); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* image.clj:299-312: 4D [d c b a] -> 2D rows d*b, cols c*a; pixel(y,x) = array[y/b][x/a][y%b][x%a] (3 doubles each) */
void orc_convert_4d_to_2d(const double* in, const int* shape, double* out) {
  int d = shape[0], c = shape[1], b = shape[2], a = shape[3];
  int h = d * b, w = c * a;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t src = ((((size_t)(y / b) * c + x / a) * b + y % b) * a + x % a) * 3;
      size_t dst = ((size_t)y * w + x) * 3;
      out[dst] = in[src]; out[dst + 1] = in[src + 1]; out[dst + 2] = in[src + 2];
    }
}

/* --------------------------------------------------------- table stages */
/* Each stage evaluates f(backward(i)) (interpolate.clj:54-72) for the flat grid indices given
 * (or for every grid point when indices == NULL).  Threads: 0 -> all entries in parallel;
 * otherwise that many OpenMP threads.  structure == 1 parallelises over dimension 0 only, as
 * pmap over the first table dimension does in interpolate.clj:59,72 (given indices must then be
 * sorted ascending). */

typedef struct {
  const orc_params* P;
  orc_scatter scatter[2]; /* [mie rayleigh] (atmosphere_lut.clj:63) */
} build_ctx;

static void make_ctx(build_ctx* b, const orc_params* P) {
  b->P = P;
  b->scatter[0] = P->mie;
  b->scatter[1] = P->rayleigh;
}

static void unflatten(size_t flat, const int* shape, int ndim, double* idx) {
  for (int d = ndim - 1; d >= 0; d--) {
    idx[d] = (double)(flat % (size_t)shape[d]);
    flat /= (size_t)shape[d];
  }
}

typedef void (*entry_fn)(const build_ctx* b, const void* stage, const double* idx, double* out);

static void run_stage(const build_ctx* b, const void* stage, entry_fn f, const int* shape, int ndim,
                      const int64_t* indices, int64_t count, double* out, int threads, int structure) {
  size_t total = 1;
  for (int d = 0; d < ndim; d++) total *= (size_t)shape[d];
#ifdef _OPENMP
  int nthreads = threads > 0 ? threads : omp_get_max_threads();
#else
  int nthreads = 1; (void)threads;
#endif
  if (indices != NULL && structure) {
    /* a sample of the reference structure: the given entries (sorted ascending) grouped by their index
     * along dimension 0, one group per task, at most shape[0] tasks at a time */
    size_t inner = total / (size_t)shape[0];
    int nt = nthreads < shape[0] ? nthreads : shape[0];
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int i0 = 0; i0 < shape[0]; i0++) {
      int64_t lo = 0, hi = count;
      while (lo < hi) { /* first entry with flat index >= i0 * inner */
        int64_t mid = lo + (hi - lo) / 2;
        if ((size_t)indices[mid] < (size_t)i0 * inner) lo = mid + 1; else hi = mid;
      }
      for (int64_t k = lo; k < count && (size_t)indices[k] < (size_t)(i0 + 1) * inner; k++) {
        double idx[4];
        unflatten((size_t)indices[k], shape, ndim, idx);
        f(b, stage, idx, out + (size_t)k * 3);
      }
    }
    return;
  }
  if (indices == NULL && structure) {
    size_t inner = total / (size_t)shape[0];
    int nt = nthreads < shape[0] ? nthreads : shape[0];
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int i0 = 0; i0 < shape[0]; i0++)
      for (size_t k = 0; k < inner; k++) {
        size_t flat = (size_t)i0 * inner + k;
        double idx[4];
        unflatten(flat, shape, ndim, idx);
        f(b, stage, idx, out + flat * 3);
      }
    return;
  }
  int64_t n = indices ? count : (int64_t)total;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t k = 0; k < n; k++) {
    size_t flat = indices ? (size_t)indices[k] : (size_t)k;
    double idx[4];
    unflatten(flat, shape, ndim, idx);
    f(b, stage, idx, out + (size_t)k * 3);
  }
}

static void store(double* out, v3 r) { out[0] = r.x; out[1] = r.y; out[2] = r.z; }

/* T = transmittance over transmittance-space (atmosphere_lut.clj:66-67,74) */
static void entry_transmittance(const build_ctx* b, const void* stage, const double* idx, double* out) {
  (void)stage;
  v3 x, v; int above;
  transmittance_backward(&b->P->planet, b->P->transmittance_shape, idx, &x, &v, &above);
  store(out, transmittance_dir(&b->P->planet, b->scatter, 2, b->P->ray_steps, x, v, above));
}

/* E0 over surface-radiance-space (atmosphere_lut.clj:68-69,75) */
static void entry_surface_radiance_base(const build_ctx* b, const void* stage, const double* idx, double* out) {
  (void)stage;
  v3 x, l;
  surface_radiance_backward(&b->P->planet, b->P->surface_radiance_shape, idx, &x, &l);
  store(out, surface_radiance_base(&b->P->planet, b->scatter, 2, b->P->ray_steps, arr3(b->P->intensity), x, l));
}

/* R1 (strength=0, Rayleigh with phase) / M1 (strength=1, Mie strength) (atmosphere_lut.clj:70-73,77-78) */
typedef struct { int strength; } single_stage;
static void entry_single(const build_ctx* b, const void* stage, const double* idx, double* out) {
  const single_stage* s = (const single_stage*)stage;
  v3 x, v, l; int above;
  scatter_backward(&b->P->planet, b->P->scatter_shape, idx, &x, &v, &l, &above);
  single_ctx c = {&b->P->planet, b->scatter, 2, s->strength ? &b->scatter[0] : &b->scatter[1], b->P->ray_steps,
                  arr3(b->P->intensity), s->strength};
  store(out, ray_scatter(&b->P->planet, b->scatter, 2, b->P->ray_steps, single_scatter_J, &c, x, v, l, above));
}

/* dS as used by point-scatter and surface-radiance: the first-iteration function or a table */
typedef struct {
  int first;        /* 1: R1 + M1*phase(mie); 0: table */
  table_ctx table;  /* dS table (first == 0) */
  first_dS_ctx fd;  /* (first == 1) */
} dS_source;

static void make_dS(dS_source* d, const build_ctx* b, int first, const double* A, const double* B) {
  d->first = first;
  d->table.planet = &b->P->planet; d->table.shape = b->P->scatter_shape; d->table.data = A;
  d->fd.R1 = d->table;
  d->fd.M1.planet = &b->P->planet; d->fd.M1.shape = b->P->scatter_shape; d->fd.M1.data = B;
  d->fd.mie_g = b->P->mie.g;
}
static scatter_fn dS_fn(const dS_source* d) { return d->first ? first_dS : interp_scatter; }
static const void* dS_ctx(const dS_source* d) { return d->first ? (const void*)&d->fd : (const void*)&d->table; }

/* dJ = point-scatter over point-scatter-space (atmosphere_lut.clj:88-91) */
typedef struct { dS_source dS; table_ctx dE; } ps_stage;
static void entry_point_scatter(const build_ctx* b, const void* stage, const double* idx, double* out) {
  const ps_stage* s = (const ps_stage*)stage;
  v3 x, v, l; int above;
  scatter_backward(&b->P->planet, b->P->scatter_shape, idx, &x, &v, &l, &above);
  store(out, point_scatter(&b->P->planet, b->scatter, 2, dS_fn(&s->dS), dS_ctx(&s->dS), interp_surface, &s->dE,
                           b->P->sphere_steps, b->P->ray_steps, x, v, l));
}

/* dS' = ray-scatter of the dJ table (atmosphere_lut.clj:92,94) */
typedef struct { table_ctx dJ; } rs_stage;
static void entry_ray_scatter(const build_ctx* b, const void* stage, const double* idx, double* out) {
  const rs_stage* s = (const rs_stage*)stage;
  v3 x, v, l; int above;
  scatter_backward(&b->P->planet, b->P->scatter_shape, idx, &x, &v, &l, &above);
  store(out, ray_scatter(&b->P->planet, b->scatter, 2, b->P->ray_steps, interp_scatter, &s->dJ, x, v, l, above));
}

/* dE' = surface-radiance of dS (atmosphere_lut.clj:89,93); steps = ray-steps */
typedef struct { dS_source dS; } sr_stage;
static void entry_surface_radiance(const build_ctx* b, const void* stage, const double* idx, double* out) {
  const sr_stage* s = (const sr_stage*)stage;
  v3 x, l;
  surface_radiance_backward(&b->P->planet, b->P->surface_radiance_shape, idx, &x, &l);
  store(out, surface_radiance(&b->P->planet, dS_fn(&s->dS), dS_ctx(&s->dS), b->P->ray_steps, x, l));
}

/* Re-sampling of an interpolated table (or the sum of two on the same space) at f(backward(i)):
 * the accumulations of atmosphere_lut.clj:95-97 and the final make-lookup-table calls of 98-101. */
typedef struct { int space; const int* shape; const double* A; const double* B; } resample_stage;
static void entry_resample(const build_ctx* b, const void* stage, const double* idx, double* out) {
  const resample_stage* s = (const resample_stage*)stage;
  const orc_planet* p = &b->P->planet;
  double fidx[4];
  int ndim;
  if (s->space == 2) {
    v3 x, v; int above;
    transmittance_backward(p, s->shape, idx, &x, &v, &above);
    transmittance_forward(p, s->shape, x, v, above, fidx);
    ndim = 2;
  } else if (s->space == 3) {
    v3 x, l;
    surface_radiance_backward(p, s->shape, idx, &x, &l);
    surface_radiance_forward(p, s->shape, x, l, fidx);
    ndim = 2;
  } else {
    v3 x, v, l; int above;
    scatter_backward(p, s->shape, idx, &x, &v, &l, &above);
    scatter_forward(p, s->shape, x, v, l, above, fidx);
    ndim = 4;
  }
  v3 r = interpolate_value(s->A, s->shape, ndim, fidx);
  if (s->B) r = vadd(r, interpolate_value(s->B, s->shape, ndim, fidx));
  store(out, r);
}

/* Stage exports (flat indices may be NULL for the whole grid). */
void orc_stage_transmittance(const orc_params* P, const int64_t* indices, int64_t count, double* out, int threads,
                             int structure) {
  build_ctx b; make_ctx(&b, P);
  run_stage(&b, NULL, entry_transmittance, P->transmittance_shape, 2, indices, count, out, threads, structure);
}
void orc_stage_surface_radiance_base(const orc_params* P, const int64_t* indices, int64_t count, double* out,
                                     int threads, int structure) {
  build_ctx b; make_ctx(&b, P);
  run_stage(&b, NULL, entry_surface_radiance_base, P->surface_radiance_shape, 2, indices, count, out, threads,
            structure);
}
void orc_stage_single_scatter(const orc_params* P, int strength, const int64_t* indices, int64_t count, double* out,
                              int threads, int structure) {
  build_ctx b; make_ctx(&b, P);
  single_stage s = {strength};
  run_stage(&b, &s, entry_single, P->scatter_shape, 4, indices, count, out, threads, structure);
}
void orc_stage_point_scatter(const orc_params* P, int first, const double* dS_A, const double* dS_B,
                             const double* dE, const int64_t* indices, int64_t count, double* out, int threads,
                             int structure) {
  build_ctx b; make_ctx(&b, P);
  ps_stage s;
  make_dS(&s.dS, &b, first, dS_A, dS_B);
  s.dE.planet = &P->planet; s.dE.shape = P->surface_radiance_shape; s.dE.data = dE;
  run_stage(&b, &s, entry_point_scatter, P->scatter_shape, 4, indices, count, out, threads, structure);
}
void orc_stage_ray_scatter(const orc_params* P, const double* dJ, const int64_t* indices, int64_t count, double* out,
                           int threads, int structure) {
  build_ctx b; make_ctx(&b, P);
  rs_stage s;
  s.dJ.planet = &P->planet; s.dJ.shape = P->scatter_shape; s.dJ.data = dJ;
  run_stage(&b, &s, entry_ray_scatter, P->scatter_shape, 4, indices, count, out, threads, structure);
}
void orc_stage_surface_radiance(const orc_params* P, int first, const double* dS_A, const double* dS_B,
                                const int64_t* indices, int64_t count, double* out, int threads, int structure) {
  build_ctx b; make_ctx(&b, P);
  sr_stage s;
  make_dS(&s.dS, &b, first, dS_A, dS_B);
  run_stage(&b, &s, entry_surface_radiance, P->surface_radiance_shape, 2, indices, count, out, threads, structure);
}
void orc_stage_resample(const orc_params* P, int space, const double* A, const double* B, const int64_t* indices,
                        int64_t count, double* out, int threads, int structure) {
  build_ctx b; make_ctx(&b, P);
  const int* shape = space == 2 ? P->transmittance_shape : space == 3 ? P->surface_radiance_shape : P->scatter_shape;
  resample_stage s = {space, shape, A, B};
  run_stage(&b, &s, entry_resample, shape, space == 4 ? 4 : 2, indices, count, out, threads, structure);
}

static size_t count_of(const int* shape, int ndim) {
  size_t n = 1;
  for (int d = 0; d < ndim; d++) n *= (size_t)shape[d];
  return n;
}

/* atmosphere_lut.clj:43-105: the whole driver.  Outputs are the final make-lookup-table results
 * (grid order, 3 doubles per entry, before the 4D->2D tiling of the file writes).  threads and
 * structure as for the stages (structure == 1: every stage parallel over dimension 0 only). */
int orc_generate_atmosphere_luts(const orc_params* P, double* T_out, double* E_out, double* S_out, double* M_out,
                                 int threads, int structure) {
  size_t nT = count_of(P->transmittance_shape, 2), nE = count_of(P->surface_radiance_shape, 2);
  size_t nS = count_of(P->scatter_shape, 4);
  double* T = malloc(nT * 3 * sizeof(double));
  double* dE = malloc(nE * 3 * sizeof(double));
  double* dE_new = malloc(nE * 3 * sizeof(double));
  double* E = calloc(nE * 3, sizeof(double));
  double* E_new = malloc(nE * 3 * sizeof(double));
  double* R1 = malloc(nS * 3 * sizeof(double));
  double* M1 = malloc(nS * 3 * sizeof(double));
  double* dJ = malloc(nS * 3 * sizeof(double));
  double* dS = malloc(nS * 3 * sizeof(double));
  double* dS_new = malloc(nS * 3 * sizeof(double));
  double* S = malloc(nS * 3 * sizeof(double));
  double* S_new = malloc(nS * 3 * sizeof(double));
  if (!T || !dE || !dE_new || !E || !E_new || !R1 || !M1 || !dJ || !dS || !dS_new || !S || !S_new) {
    free(T); free(dE); free(dE_new); free(E); free(E_new); free(R1); free(M1); free(dJ); free(dS); free(dS_new);
    free(S); free(S_new);
    return -1;
  }
  orc_stage_transmittance(P, NULL, 0, T, threads, structure);
  orc_stage_surface_radiance_base(P, NULL, 0, dE, threads, structure);
  orc_stage_single_scatter(P, 0, NULL, 0, R1, threads, structure);
  orc_stage_single_scatter(P, 1, NULL, 0, M1, threads, structure);
  memcpy(S, R1, nS * 3 * sizeof(double)); /* S <- first-order-rayleigh */
  int first = 1;
  int e_is_zero = 1; /* E <- (constantly 0) */
  for (int it = 0; it < P->iterations; it++) {
    orc_stage_point_scatter(P, first, first ? R1 : dS, first ? M1 : NULL, dE, NULL, 0, dJ, threads, structure);
    orc_stage_surface_radiance(P, first, first ? R1 : dS, first ? M1 : NULL, NULL, 0, dE_new, threads, structure);
    orc_stage_ray_scatter(P, dJ, NULL, 0, dS_new, threads, structure);
    { double* t = dE; dE = dE_new; dE_new = t; }
    { double* t = dS; dS = dS_new; dS_new = t; }
    first = 0;
    /* E <- table(E + dE); the constant-zero E contributes exactly 0 */
    orc_stage_resample(P, 3, dE, e_is_zero ? NULL : E, NULL, 0, E_new, threads, structure);
    { double* t = E; E = E_new; E_new = t; }
    e_is_zero = 0;
    /* S <- table(S + dS) */
    orc_stage_resample(P, 4, S, dS, NULL, 0, S_new, threads, structure);
    { double* t = S; S = S_new; S_new = t; }
  }
  orc_stage_resample(P, 2, T, NULL, NULL, 0, T_out, threads, structure);
  if (e_is_zero) memset(E_out, 0, nE * 3 * sizeof(double));
  else orc_stage_resample(P, 3, E, NULL, NULL, 0, E_out, threads, structure);
  orc_stage_resample(P, 4, S, NULL, NULL, 0, S_out, threads, structure);
  orc_stage_resample(P, 4, M1, NULL, NULL, 0, M_out, threads, structure);
  free(T); free(dE); free(dE_new); free(E); free(E_new); free(R1); free(M1); free(dJ); free(dS); free(dS_new);
  free(S); free(S_new);
  return 0;
}

/* ------------------------------------- GLSL consumer restatement (FP64) */
/* The renderer reads the tables through these shader functions; the reference's table-level
 * golden vectors are probes of them.  GL_LINEAR + clamp with the index->texel mapping of
 * convert-2d-index.glsl / make-2d-index-from-4d.glsl is multilinear interpolation on the
 * (size-1)-scaled indices, which is what interpolate_value computes. */

/* atmosphere/transmittance-outer.glsl */
void orc_glsl_transmittance_outer(const orc_planet* p, const int* shape, const double* T, const double* point,
                                  const double* direction, double* out) {
  double idx[2];
  transmittance_forward(p, shape, arr3(point), arr3(direction), 1, idx);
  store(out, interpolate_value(T, shape, 2, idx));
}

/* atmosphere/transmittance-track.glsl */
static v3 glsl_transmittance_track(const orc_planet* p, const int* shape, const double* T, v3 pp, v3 q) {
  double dist = vmag(vsub(q, pp));
  if (dist > 0) {
    v3 direction = vdiv(vsub(q, pp), dist);
    int above = is_above_horizon(p, pp, direction);
    double uvp[2], uvq[2];
    transmittance_forward(p, shape, pp, direction, above, uvp);
    transmittance_forward(p, shape, q, direction, above, uvq);
    v3 t1 = interpolate_value(T, shape, 2, uvp), t2 = interpolate_value(T, shape, 2, uvq);
    return V(t1.x / t2.x, t1.y / t2.y, t1.z / t2.z);
  }
  return V(1, 1, 1);
}
void orc_glsl_transmittance_track(const orc_planet* p, const int* shape, const double* T, const double* pp,
                                  const double* q, double* out) {
  store(out, glsl_transmittance_track(p, shape, T, arr3(pp), arr3(q)));
}

/* atmosphere/ray-scatter-track.glsl (Mie phase g given; the shader hard-codes 0.76) */
void orc_glsl_ray_scatter_track(const orc_planet* p, const int* shapeS, const double* S, const double* M,
                                const int* shapeT, const double* T, double g, const double* light, const double* pp,
                                const double* q, double* out) {
  v3 P0 = arr3(pp), Q = arr3(q), l = arr3(light);
  double dist = vmag(vsub(Q, P0));
  if (!(dist > 0)) { store(out, V(0, 0, 0)); return; }
  v3 direction = vdiv(vsub(Q, P0), dist);
  int above = is_above_horizon(p, P0, direction);
  double mu = vdot(direction, l);
  double ip[4], iq[4];
  scatter_forward(p, shapeS, P0, direction, l, above, ip);
  scatter_forward(p, shapeS, Q, direction, l, above, iq);
  v3 sp = vadd(interpolate_value(S, shapeS, 4, ip), vmul(interpolate_value(M, shapeS, 4, ip), orc_phase(g, mu)));
  v3 sq = vadd(interpolate_value(S, shapeS, 4, iq), vmul(interpolate_value(M, shapeS, 4, iq), orc_phase(g, mu)));
  v3 tpq = glsl_transmittance_track(p, shapeT, T, P0, Q);
  store(out, vsub(sp, vemul(tpq, sq)));
}

/* planet/surface-radiance.glsl */
void orc_glsl_surface_radiance(const orc_planet* p, const int* shape, const double* E, const double* point,
                               const double* light, double* out) {
  double idx[2];
  surface_radiance_forward(p, shape, arr3(point), arr3(light), idx);
  store(out, interpolate_value(E, shape, 2, idx));
}

/* ---------------------- GLSL texture path over the tiled .scatter tables (FP64) */
/* The renderer uploads each file as one 2D RGB texture with GL_LINEAR filtering and CLAMP_TO_EDGE
 * (atmosphere.clj:587-619; texture.clj:238-245, 380-398: the 4D tables as a (heading x elevation) wide,
 * (sun elevation x height) high image) and reads it through these shader functions.  texture() is
 * restated with exact bilinear weights (the hardware's are fixed point, hence the reference tests'
 * 1e-4 tolerances on GL results). */

/* texture(sampler2D, vec2(s, t)) of a width x height texture, texel (i, j) at tex[(j width + i) 3] */
static v3 gl_texture(const double* tex, int width, int height, double s, double t) {
  double u = s * width - 0.5, v = t * height - 0.5;
  double fu = floor(u), fv = floor(v);
  double a = u - fu, b = v - fv;
  int i0 = (int)fu, j0 = (int)fv, i1 = i0 + 1, j1 = j0 + 1;
  i0 = i0 < 0 ? 0 : (i0 > width - 1 ? width - 1 : i0);
  i1 = i1 < 0 ? 0 : (i1 > width - 1 ? width - 1 : i1);
  j0 = j0 < 0 ? 0 : (j0 > height - 1 ? height - 1 : j0);
  j1 = j1 < 0 ? 0 : (j1 > height - 1 ? height - 1 : j1);
  v3 t00 = arr3(tex + ((size_t)j0 * width + i0) * 3), t10 = arr3(tex + ((size_t)j0 * width + i1) * 3);
  v3 t01 = arr3(tex + ((size_t)j1 * width + i0) * 3), t11 = arr3(tex + ((size_t)j1 * width + i1) * 3);
  v3 lo = vadd(vmul(t00, 1 - a), vmul(t10, a)), hi = vadd(vmul(t01, 1 - a), vmul(t11, a));
  return vadd(vmul(lo, 1 - b), vmul(hi, b));
}

/* core/convert-2d-index.glsl: normalised index (x, y) -> texture coordinates of texel centres */
void orc_glsl_convert_2d_index(const double* idx, int size_y, int size_x, double* out) {
  out[0] = (idx[0] * (size_x - 1) + 0.5) / size_x;
  out[1] = (idx[1] * (size_y - 1) + 0.5) / size_y;
}

/* core/interpolate-2d.glsl */
static v3 glsl_interpolate_2d(const double* tex, int size_y, int size_x, const double* idx) {
  double c[2];
  orc_glsl_convert_2d_index(idx, size_y, size_x, c);
  return gl_texture(tex, size_x, size_y, c[0], c[1]);
}
void orc_glsl_interpolate_2d(const double* tex, int size_y, int size_x, const double* idx, double* out) {
  store(out, glsl_interpolate_2d(tex, size_y, size_x, idx));
}

/* core/make-2d-index-from-4d.glsl: pixel index (x heading, y sun elevation, z elevation, w height) ->
 * (s, t, p, q): texture columns of elevation tiles floor(z) and floor(z) + 1 (clamped), rows of height
 * tiles floor(w) and floor(w) + 1 (clamped) */
void orc_glsl_make_2d_index_from_4d(const double* idx, int size_w, int size_z, int size_y, int size_x, double* out) {
  double z_floor = floor(idx[2]), w_floor = floor(idx[3]);
  double dx = (double)size_x * size_z, dy = (double)size_y * size_w;
  out[0] = (0.5 + idx[0] + z_floor * size_x) / dx;
  out[1] = (0.5 + idx[0] + fmin(z_floor + 1, size_z - 1) * size_x) / dx;
  out[2] = (0.5 + idx[1] + w_floor * size_y) / dy;
  out[3] = (0.5 + idx[1] + fmin(w_floor + 1, size_w - 1) * size_y) / dy;
}

/* core/interpolate-4d.glsl: heading and sun elevation filtered by the texture unit, elevation and height
 * mixed explicitly */
static v3 glsl_interpolate_4d(const double* tex, int size_w, int size_z, int size_y, int size_x, const double* idx) {
  double pixel[4] = {idx[0] * (size_x - 1), idx[1] * (size_y - 1), idx[2] * (size_z - 1), idx[3] * (size_w - 1)};
  double fs = pixel[2] - floor(pixel[2]), ft = pixel[3] - floor(pixel[3]);
  double c[4];
  orc_glsl_make_2d_index_from_4d(pixel, size_w, size_z, size_y, size_x, c);
  int width = size_x * size_z, height = size_y * size_w;
  v3 sp = gl_texture(tex, width, height, c[0], c[2]), tp = gl_texture(tex, width, height, c[1], c[2]);
  v3 sq = gl_texture(tex, width, height, c[0], c[3]), tq = gl_texture(tex, width, height, c[1], c[3]);
  return vadd(vadd(vmul(sp, (1 - fs) * (1 - ft)), vmul(tp, fs * (1 - ft))),
              vadd(vmul(sq, (1 - fs) * ft), vmul(tq, fs * ft)));
}
void orc_glsl_interpolate_4d(const double* tex, int size_w, int size_z, int size_y, int size_x, const double* idx,
                             double* out) {
  store(out, glsl_interpolate_4d(tex, size_w, size_z, size_y, size_x, idx));
}

/* core/ray-scatter-forward.glsl (normalised, reversed order) */
static void glsl_ray_scatter_forward(const orc_planet* p, const int* shape, v3 point, v3 direction, v3 l, int above,
                                     double* idx) {
  double c[4];
  scatter_forward(p, shape, point, direction, l, above, c);
  for (int d = 0; d < 4; d++) idx[3 - d] = c[d] / (shape[d] - 1);
}

/* atmosphere/transmittance-outer.glsl over the transmittance file's texture (64 x 255 -> 255 wide) */
void orc_glsl_tiled_transmittance_outer(const orc_planet* p, const int* shape, const double* T, const double* point,
                                        const double* direction, double* out) {
  double c[2], idx[2];
  transmittance_forward(p, shape, arr3(point), arr3(direction), 1, c);
  idx[0] = c[1] / (shape[1] - 1);
  idx[1] = c[0] / (shape[0] - 1);
  store(out, glsl_interpolate_2d(T, shape[0], shape[1], idx));
}

/* atmosphere/ray-scatter-outer.glsl over the tiled ray-scatter and Mie-strength textures (g: the Mie
 * phase asymmetry, hard-coded 0.76 in the shader) */
void orc_glsl_tiled_ray_scatter_outer(const orc_planet* p, const int* shape, const double* S, const double* M, double g,
                                      const double* light, const double* point, const double* direction, double* out) {
  double idx[4];
  v3 l = arr3(light), d = arr3(direction);
  glsl_ray_scatter_forward(p, shape, arr3(point), d, l, 1, idx);
  v3 s = glsl_interpolate_4d(S, shape[0], shape[1], shape[2], shape[3], idx);
  v3 m = glsl_interpolate_4d(M, shape[0], shape[1], shape[2], shape[3], idx);
  store(out, vadd(s, vmul(m, orc_phase(g, vdot(d, l)))));
}

/* planet/surface-radiance.glsl over the surface-radiance file's texture */
void orc_glsl_tiled_surface_radiance(const orc_planet* p, const int* shape, const double* E, const double* point,
                                     const double* light, double* out) {
  double c[2], idx[2];
  surface_radiance_forward(p, shape, arr3(point), arr3(light), c);
  idx[0] = c[1] / (shape[1] - 1);
  idx[1] = c[0] / (shape[0] - 1);
  store(out, glsl_interpolate_2d(E, shape[0], shape[1], idx));
}

/* The analytic ray-scatter of t_planet.clj:312-315, integrated by surface-radiance over a
 * surface-radiance-space table (t_planet.clj:318-320). */
typedef struct { const orc_planet* planet; } analytic_ctx;
static v3 analytic_ray_scatter(const void* vctx, v3 x, v3 v, v3 l, int above) {
  (void)above;
  const analytic_ctx* c = (const analytic_ctx*)vctx;
  double value = pow(fmax(vdot(v, l), 0.0), 10) * exp((c->planet->radius - vmag(x)) / 5500.0);
  return V(value, value, value);
}
void orc_surface_radiance_analytic_table(const orc_planet* p, const int* shape, int steps, double* out) {
  analytic_ctx c = {p};
  for (int i = 0; i < shape[0]; i++)
    for (int j = 0; j < shape[1]; j++) {
      double idx[2] = {i, j};
      v3 x, l;
      surface_radiance_backward(p, shape, idx, &x, &l);
      store(out + ((size_t)i * shape[1] + j) * 3, surface_radiance(p, analytic_ray_scatter, &c, steps, x, l));
    }
}

/* Point-scatter with constant S (sky) and E (ground) fields, for closed-form invariant tests. */
typedef struct { v3 value; } const_ctx;
static v3 const_scatter(const void* vctx, v3 x, v3 v, v3 l, int above) {
  (void)x; (void)v; (void)l;
  const const_ctx* c = (const const_ctx*)vctx;
  return above ? c->value : V(0, 0, 0);
}
static v3 const_surface(const void* vctx, v3 x, v3 l) {
  (void)x; (void)l;
  return ((const const_ctx*)vctx)->value;
}
void orc_point_scatter_constant(const orc_planet* p, const orc_scatter* s, int ns, const double* sky,
                                const double* ground, int sphere_steps, int ray_steps, const double* x,
                                const double* v, const double* l, double* out) {
  const_ctx cs = {arr3(sky)}, ce = {arr3(ground)};
  store(out, point_scatter(p, s, ns, const_scatter, &cs, const_surface, &ce, sphere_steps, ray_steps, arr3(x), arr3(v),
                           arr3(l)));
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
