"""ctypes wrapper of the FP64 CPU oracle (oracle/atmo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the CPU-baseline leg of ``bench.py``.  The product library never uses it.

Parameter dictionaries use the reference's keys without namespaces
(``centre``, ``radius``, ``height``, ``brightness`` for the planet;
``scatter-base``, ``scatter-scale``, ``scatter-quotient``, ``scatter-g`` for a
scattering component), as in atmosphere.clj:35-39, 64-67 and sphere.clj:25.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_atmo.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)


class Planet(ctypes.Structure):
    _fields_ = [("centre", ctypes.c_double * 3), ("radius", ctypes.c_double), ("height", ctypes.c_double),
                ("brightness", ctypes.c_double * 3)]


class Scatter(ctypes.Structure):
    _fields_ = [("base", ctypes.c_double * 3), ("scale", ctypes.c_double), ("quotient", ctypes.c_double),
                ("g", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("planet", Planet), ("mie", Scatter), ("rayleigh", Scatter), ("intensity", ctypes.c_double * 3),
                ("transmittance_shape", ctypes.c_int * 2), ("surface_radiance_shape", ctypes.c_int * 2),
                ("scatter_shape", ctypes.c_int * 4), ("ray_steps", ctypes.c_int), ("sphere_steps", ctypes.c_int),
                ("iterations", ctypes.c_int)]


_lib = None


def lib():
    """Load the oracle library, building it with make when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            import subprocess
            subprocess.run(["make", "-C", _HERE, "-s"], check=True)
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_phase.restype = ctypes.c_double
        _lib.orc_height_of.restype = ctypes.c_double
        _lib.orc_limit_quot.restype = ctypes.c_double
        _lib.orc_horizon_distance.restype = ctypes.c_double
        _lib.orc_elevation_to_index.restype = ctypes.c_double
        _lib.orc_height_to_index.restype = ctypes.c_double
        _lib.orc_sun_elevation_to_index.restype = ctypes.c_double
        _lib.orc_index_to_sin_sun_elevation.restype = ctypes.c_double
        _lib.orc_sun_angle_to_index.restype = ctypes.c_double
        for name in ("orc_limit_quot",):
            getattr(_lib, name).argtypes = [ctypes.c_double] * 4
        _lib.orc_phase.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_index_to_sin_sun_elevation.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib.orc_horizon_distance.argtypes = [ctypes.POINTER(Planet), ctypes.c_double]
        _lib.orc_index_to_height.argtypes = [ctypes.POINTER(Planet), ctypes.c_int, ctypes.c_double, _dp]
        _lib.orc_index_to_elevation.argtypes = [ctypes.POINTER(Planet), ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double, _dp]
        _lib.orc_index_to_sun_direction.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp]
        _lib.orc_integral_ray_test.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        _lib.orc_scattering.argtypes = [ctypes.POINTER(Scatter), ctypes.c_double, _dp]
        _lib.orc_extinction.argtypes = [ctypes.POINTER(Scatter), ctypes.c_double, _dp]
        _lib.orc_max_threads.restype = ctypes.c_int
        PP, i32, i64 = ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int64
        _lib.orc_stage_transmittance.argtypes = [PP, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_surface_radiance_base.argtypes = [PP, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_single_scatter.argtypes = [PP, i32, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_point_scatter.argtypes = [PP, i32, _dp, _dp, _dp, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_ray_scatter.argtypes = [PP, _dp, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_surface_radiance.argtypes = [PP, i32, _dp, _dp, _lp, i64, _dp, i32, i32]
        _lib.orc_stage_resample.argtypes = [PP, i32, _dp, _dp, _lp, i64, _dp, i32, i32]
        _lib.orc_generate_atmosphere_luts.argtypes = [PP, _dp, _dp, _dp, _dp, i32, i32]
    return _lib


def _vec(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a


def _ptr(a):
    return a.ctypes.data_as(_dp)


def planet(radius=6378000.0, height=35000.0, brightness=(0.3, 0.3, 0.3), centre=(0.0, 0.0, 0.0)):
    p = Planet()
    p.centre[:] = list(centre)
    p.radius = radius
    p.height = height
    p.brightness[:] = list(brightness)
    return p


def scatter(base, scale, quotient=1.0, g=0.0):
    s = Scatter()
    s.base[:] = list(base)
    s.scale = scale
    s.quotient = quotient
    s.g = g
    return s


def scatter_array(components):
    arr = (Scatter * len(components))()
    for i, c in enumerate(components):
        arr[i] = c
    return arr


# Parameters of atmosphere_lut.clj:20-40 and of the reference's tests (t_atmosphere.clj:43-67).
def mie(g=0.76, quotient=0.9, base=(2e-5, 2e-5, 2e-5), scale=1200.0):
    return scatter(base, scale, quotient, g)


def rayleigh(base=(5.8e-6, 13.5e-6, 33.1e-6), scale=8000.0, quotient=1.0, g=0.0):
    return scatter(base, scale, quotient, g)


def params(planet_, mie_, rayleigh_, transmittance_shape, surface_radiance_shape, scatter_shape, ray_steps,
           sphere_steps, iterations, intensity=(1.0, 1.0, 1.0)):
    P = Params()
    P.planet = planet_
    P.mie = mie_
    P.rayleigh = rayleigh_
    P.intensity[:] = list(intensity)
    P.transmittance_shape[:] = list(transmittance_shape)
    P.surface_radiance_shape[:] = list(surface_radiance_shape)
    P.scatter_shape[:] = list(scatter_shape)
    P.ray_steps = ray_steps
    P.sphere_steps = sphere_steps
    P.iterations = iterations
    return P


def _out3():
    return np.zeros(3)


# ---------------------------------------------------------------- scalar functions
def phase(g, mu):
    return lib().orc_phase(g, mu)


def limit_quot(a, b, lower, upper=None):
    if upper is None:
        lower, upper = -lower, lower
    return lib().orc_limit_quot(a, b, lower, upper)


def scattering(component, height):
    out = _out3()
    lib().orc_scattering(ctypes.byref(component), height, _ptr(out))
    return out


def extinction(component, height):
    out = _out3()
    lib().orc_extinction(ctypes.byref(component), height, _ptr(out))
    return out


def height(p, point):
    lib().orc_height_of.argtypes = [ctypes.POINTER(Planet), _dp]
    return lib().orc_height_of(ctypes.byref(p), _ptr(_vec(point)))


def ray_sphere_intersection(centre, radius, origin, direction):
    d, ln = ctypes.c_double(), ctypes.c_double()
    f = lib().orc_ray_sphere_intersection
    f.argtypes = [Vec3, ctypes.c_double, Vec3, Vec3, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f(Vec3(*centre), radius, Vec3(*origin), Vec3(*direction), ctypes.byref(d), ctypes.byref(ln))
    return d.value, ln.value


class Vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


def _pt3(name, p, a, b):
    out = _out3()
    getattr(lib(), name)(ctypes.byref(p), _ptr(_vec(a)), _ptr(_vec(b)), _ptr(out))
    return out


def atmosphere_intersection(p, origin, direction):
    return _pt3("orc_atmosphere_intersection", p, origin, direction)


def surface_intersection(p, origin, direction):
    return _pt3("orc_surface_intersection", p, origin, direction)


def ray_extremity(p, origin, direction):
    return _pt3("orc_ray_extremity", p, origin, direction)


def surface_point(p, point):
    return bool(lib().orc_surface_point(ctypes.byref(p), _ptr(_vec(point))))


def is_above_horizon(p, point, direction):
    return bool(lib().orc_is_above_horizon(ctypes.byref(p), _ptr(_vec(point)), _ptr(_vec(direction))))


def transmittance(p, components, steps, x, x0_or_v, above=None):
    """atmosphere.clj:114-128: two-point arity when ``above`` is None, else direction arity."""
    arr = scatter_array(components)
    out = _out3()
    if above is None:
        lib().orc_transmittance_points(ctypes.byref(p), arr, len(components), steps, _ptr(_vec(x)),
                                       _ptr(_vec(x0_or_v)), _ptr(out))
    else:
        lib().orc_transmittance_dir(ctypes.byref(p), arr, len(components), steps, _ptr(_vec(x)),
                                    _ptr(_vec(x0_or_v)), int(above), _ptr(out))
    return out


def surface_radiance_base(p, components, steps, intensity, x, l):
    arr = scatter_array(components)
    out = _out3()
    lib().orc_surface_radiance_base(ctypes.byref(p), arr, len(components), steps, _ptr(_vec(intensity)),
                                    _ptr(_vec(x)), _ptr(_vec(l)), _ptr(out))
    return out


def single_scatter_J(p, components, component, steps, intensity, strength, x, v, l):
    arr = scatter_array(components)
    out = _out3()
    lib().orc_single_scatter_J(ctypes.byref(p), arr, len(components), ctypes.byref(component), steps,
                               _ptr(_vec(intensity)), int(strength), _ptr(_vec(x)), _ptr(_vec(v)), _ptr(_vec(l)),
                               _ptr(out))
    return out


def point_scatter_base(p, components, steps, intensity, x, v, l):
    arr = scatter_array(components)
    out = _out3()
    lib().orc_point_scatter_base(ctypes.byref(p), arr, len(components), steps, _ptr(_vec(intensity)), _ptr(_vec(x)),
                                 _ptr(_vec(v)), _ptr(_vec(l)), _ptr(out))
    return out


def ray_scatter_single(p, components, component, steps, intensity, strength, x, v, l, above):
    arr = scatter_array(components)
    out = _out3()
    lib().orc_ray_scatter_single(ctypes.byref(p), arr, len(components), ctypes.byref(component), steps,
                                 _ptr(_vec(intensity)), int(strength), _ptr(_vec(x)), _ptr(_vec(v)), _ptr(_vec(l)),
                                 int(above), _ptr(out))
    return out


def point_scatter_constant(p, components, sky, ground, sphere_steps, ray_steps, x, v, l):
    arr = scatter_array(components)
    out = _out3()
    lib().orc_point_scatter_constant(ctypes.byref(p), arr, len(components), _ptr(_vec(sky)), _ptr(_vec(ground)),
                                     sphere_steps, ray_steps, _ptr(_vec(x)), _ptr(_vec(v)), _ptr(_vec(l)), _ptr(out))
    return out


def integral_sphere_test(half, steps, normal, kind):
    out = _out3()
    lib().orc_integral_sphere_test(int(half), steps, _ptr(_vec(normal)), kind, _ptr(out))
    return out


def sphere_direction_count(half, steps):
    return lib().orc_sphere_direction_count(int(half), steps)


def integral_ray_test(origin, direction, steps, distance, kind):
    out = _out3()
    lib().orc_integral_ray_test(_ptr(_vec(origin)), _ptr(_vec(direction)), steps, distance, kind, _ptr(out))
    return out


def oriented_matrix(n):
    out = np.zeros(9)
    lib().orc_oriented_matrix(_ptr(_vec(n)), _ptr(out))
    return out.reshape(3, 3)


# ---------------------------------------------------------------- index spaces
def horizon_distance(p, radius):
    return lib().orc_horizon_distance(ctypes.byref(p), radius)


def elevation_to_index(p, size, point, direction, above):
    return lib().orc_elevation_to_index(ctypes.byref(p), size, _ptr(_vec(point)), _ptr(_vec(direction)), int(above))


def index_to_elevation(p, size, radius, index):
    out = _out3()
    above = lib().orc_index_to_elevation(ctypes.byref(p), size, radius, index, _ptr(out))
    return out, bool(above)


def height_to_index(p, size, point):
    return lib().orc_height_to_index(ctypes.byref(p), size, _ptr(_vec(point)))


def index_to_height(p, size, index):
    out = _out3()
    lib().orc_index_to_height(ctypes.byref(p), size, index, _ptr(out))
    return out


def sun_elevation_to_index(size, point, l):
    return lib().orc_sun_elevation_to_index(size, _ptr(_vec(point)), _ptr(_vec(l)))


def index_to_sin_sun_elevation(size, index):
    return lib().orc_index_to_sin_sun_elevation(size, index)


def sun_angle_to_index(size, direction, l):
    return lib().orc_sun_angle_to_index(size, _ptr(_vec(direction)), _ptr(_vec(l)))


def index_to_sun_direction(size, direction, sin_sun_elevation, index):
    out = _out3()
    lib().orc_index_to_sun_direction(size, _ptr(_vec(direction)), sin_sun_elevation, index, _ptr(out))
    return out


SPACE_TRANSMITTANCE, SPACE_SURFACE_RADIANCE, SPACE_SCATTER = 2, 3, 4


def space_forward(p, space, shape, point, direction=(0, 0, 0), light=(0, 0, 0), above=False):
    shp = (ctypes.c_int * len(shape))(*shape)
    inp = _vec(list(point) + list(direction) + list(light) + [1.0 if above else 0.0])
    out = np.zeros(4)
    lib().orc_space_forward(ctypes.byref(p), space, shp, _ptr(inp), _ptr(out))
    return out[:len(shape)]


def space_backward(p, space, shape, idx):
    shp = (ctypes.c_int * len(shape))(*shape)
    inp = _vec(list(idx) + [0.0] * (4 - len(idx)))
    out = np.zeros(10)
    lib().orc_space_backward(ctypes.byref(p), space, shp, _ptr(inp), _ptr(out))
    return out[0:3], out[3:6], out[6:9], bool(out[9])


def interpolate(data, shape, point):
    data = np.ascontiguousarray(data, dtype=np.float64)
    shp = (ctypes.c_int * len(shape))(*shape)
    out = _out3()
    lib().orc_interpolate(_ptr(data), shp, len(shape), _ptr(_vec(point)), _ptr(out))
    return out


def convert_4d_to_2d(table):
    """image.clj:299-312 on a [d c b a 3] array -> [d*b, c*a, 3]."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    d, c, b, a = t.shape[:4]
    out = np.zeros((d * b, c * a, 3))
    shp = (ctypes.c_int * 4)(d, c, b, a)
    lib().orc_convert_4d_to_2d(_ptr(t), shp, _ptr(out))
    return out


# ---------------------------------------------------------------- table stages
def _idx(indices):
    if indices is None:
        return None, 0
    arr = np.ascontiguousarray(indices, dtype=np.int64)
    return arr, len(arr)


def _lptr(arr):
    return None if arr is None else arr.ctypes.data_as(_lp)


def _tab(a):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.float64)
    return arr, _ptr(arr)


def _count(shape):
    n = 1
    for s in shape:
        n *= s
    return n


def _out_for(shape, count):
    n = count if count else _count(shape)
    return np.zeros((n, 3))


def stage_transmittance(P, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.transmittance_shape, n)
    lib().orc_stage_transmittance(ctypes.byref(P), _lptr(arr), n, _ptr(out), threads, structure)
    return out


def stage_surface_radiance_base(P, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.surface_radiance_shape, n)
    lib().orc_stage_surface_radiance_base(ctypes.byref(P), _lptr(arr), n, _ptr(out), threads, structure)
    return out


def stage_single_scatter(P, strength, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.scatter_shape, n)
    lib().orc_stage_single_scatter(ctypes.byref(P), int(strength), _lptr(arr), n, _ptr(out), threads, structure)
    return out


def stage_point_scatter(P, first, dS_A, dS_B, dE, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.scatter_shape, n)
    a, pa = _tab(dS_A)
    b, pb = _tab(dS_B)
    e, pe = _tab(dE)
    lib().orc_stage_point_scatter(ctypes.byref(P), int(first), pa, pb, pe, _lptr(arr), n, _ptr(out), threads,
                                  structure)
    return out


def stage_ray_scatter(P, dJ, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.scatter_shape, n)
    j, pj = _tab(dJ)
    lib().orc_stage_ray_scatter(ctypes.byref(P), pj, _lptr(arr), n, _ptr(out), threads, structure)
    return out


def stage_surface_radiance(P, first, dS_A, dS_B, indices=None, threads=0, structure=0):
    arr, n = _idx(indices)
    out = _out_for(P.surface_radiance_shape, n)
    a, pa = _tab(dS_A)
    b, pb = _tab(dS_B)
    lib().orc_stage_surface_radiance(ctypes.byref(P), int(first), pa, pb, _lptr(arr), n, _ptr(out), threads,
                                     structure)
    return out


def stage_resample(P, space, A, B=None, indices=None, threads=0, structure=0):
    shape = {2: P.transmittance_shape, 3: P.surface_radiance_shape, 4: P.scatter_shape}[space]
    arr, n = _idx(indices)
    out = _out_for(shape, n)
    a, pa = _tab(A)
    b, pb = _tab(B)
    lib().orc_stage_resample(ctypes.byref(P), space, pa, pb, _lptr(arr), n, _ptr(out), threads, structure)
    return out


def generate_atmosphere_luts(P, threads=0, structure=0):
    """atmosphere_lut.clj:43-105 -> (T, E, S, M) grid-order tables of shape [..., 3] (before 4D->2D tiling)."""
    T = np.zeros((_count(P.transmittance_shape), 3))
    E = np.zeros((_count(P.surface_radiance_shape), 3))
    S = np.zeros((_count(P.scatter_shape), 3))
    M = np.zeros((_count(P.scatter_shape), 3))
    rc = lib().orc_generate_atmosphere_luts(ctypes.byref(P), _ptr(T), _ptr(E), _ptr(S), _ptr(M), threads, structure)
    if rc != 0:
        raise MemoryError("oracle could not allocate its tables")
    return (T.reshape(tuple(P.transmittance_shape) + (3,)), E.reshape(tuple(P.surface_radiance_shape) + (3,)),
            S.reshape(tuple(P.scatter_shape) + (3,)), M.reshape(tuple(P.scatter_shape) + (3,)))


# ---------------------------------------------------------------- GLSL consumer
def glsl_transmittance_outer(p, shape, T, point, direction):
    shp = (ctypes.c_int * 2)(*shape)
    t, pt = _tab(T)
    out = _out3()
    lib().orc_glsl_transmittance_outer(ctypes.byref(p), shp, pt, _ptr(_vec(point)), _ptr(_vec(direction)), _ptr(out))
    return out


def glsl_transmittance_track(p, shape, T, pp, q):
    shp = (ctypes.c_int * 2)(*shape)
    t, pt = _tab(T)
    out = _out3()
    lib().orc_glsl_transmittance_track(ctypes.byref(p), shp, pt, _ptr(_vec(pp)), _ptr(_vec(q)), _ptr(out))
    return out


def glsl_ray_scatter_track(p, shapeS, S, M, shapeT, T, g, light, pp, q):
    sS = (ctypes.c_int * 4)(*shapeS)
    sT = (ctypes.c_int * 2)(*shapeT)
    s, ps = _tab(S)
    m, pm = _tab(M)
    t, pt = _tab(T)
    out = _out3()
    f = lib().orc_glsl_ray_scatter_track
    f.argtypes = [ctypes.POINTER(Planet), ctypes.POINTER(ctypes.c_int), _dp, _dp, ctypes.POINTER(ctypes.c_int), _dp,
                  ctypes.c_double, _dp, _dp, _dp, _dp]
    f(ctypes.byref(p), sS, ps, pm, sT, pt, g, _ptr(_vec(light)), _ptr(_vec(pp)), _ptr(_vec(q)), _ptr(out))
    return out


def glsl_surface_radiance(p, shape, E, point, light):
    shp = (ctypes.c_int * 2)(*shape)
    e, pe = _tab(E)
    out = _out3()
    lib().orc_glsl_surface_radiance(ctypes.byref(p), shp, pe, _ptr(_vec(point)), _ptr(_vec(light)), _ptr(out))
    return out


# ---------------------------------------------------------------- GLSL texture path over tiled tables
_ci = ctypes.c_int


def glsl_convert_2d_index(idx, size_y, size_x):
    out = np.zeros(2)
    f = lib().orc_glsl_convert_2d_index
    f.argtypes = [_dp, _ci, _ci, _dp]
    f(_ptr(_vec(idx)), size_y, size_x, _ptr(out))
    return out


def glsl_make_2d_index_from_4d(idx, size_w, size_z, size_y, size_x):
    out = np.zeros(4)
    f = lib().orc_glsl_make_2d_index_from_4d
    f.argtypes = [_dp, _ci, _ci, _ci, _ci, _dp]
    f(_ptr(_vec(idx)), size_w, size_z, size_y, size_x, _ptr(out))
    return out


def glsl_interpolate_2d(tex, size_y, size_x, idx):
    """interpolate-2d.glsl on a texture of size_y rows and size_x columns (RGB per texel, row-major)."""
    t, pt = _tab(tex)
    out = _out3()
    f = lib().orc_glsl_interpolate_2d
    f.argtypes = [_dp, _ci, _ci, _dp, _dp]
    f(pt, size_y, size_x, _ptr(_vec(idx)), _ptr(out))
    return out


def glsl_interpolate_4d(tex, size_w, size_z, size_y, size_x, idx):
    """interpolate-4d.glsl on the tiled texture (size_y size_w rows, size_x size_z columns)."""
    t, pt = _tab(tex)
    out = _out3()
    f = lib().orc_glsl_interpolate_4d
    f.argtypes = [_dp, _ci, _ci, _ci, _ci, _dp, _dp]
    f(pt, size_w, size_z, size_y, size_x, _ptr(_vec(idx)), _ptr(out))
    return out


def glsl_tiled_transmittance_outer(p, shape, T, point, direction):
    t, pt = _tab(T)
    out = _out3()
    f = lib().orc_glsl_tiled_transmittance_outer
    f.argtypes = [ctypes.POINTER(Planet), ctypes.POINTER(_ci), _dp, _dp, _dp, _dp]
    f(ctypes.byref(p), (_ci * 2)(*shape), pt, _ptr(_vec(point)), _ptr(_vec(direction)), _ptr(out))
    return out


def glsl_tiled_ray_scatter_outer(p, shape, S, M, g, light, point, direction):
    s, ps = _tab(S)
    m, pm = _tab(M)
    out = _out3()
    f = lib().orc_glsl_tiled_ray_scatter_outer
    f.argtypes = [ctypes.POINTER(Planet), ctypes.POINTER(_ci), _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
    f(ctypes.byref(p), (_ci * 4)(*shape), ps, pm, g, _ptr(_vec(light)), _ptr(_vec(point)), _ptr(_vec(direction)),
      _ptr(out))
    return out


def glsl_tiled_surface_radiance(p, shape, E, point, light):
    e, pe = _tab(E)
    out = _out3()
    f = lib().orc_glsl_tiled_surface_radiance
    f.argtypes = [ctypes.POINTER(Planet), ctypes.POINTER(_ci), _dp, _dp, _dp, _dp]
    f(ctypes.byref(p), (_ci * 2)(*shape), pe, _ptr(_vec(point)), _ptr(_vec(light)), _ptr(out))
    return out


def surface_radiance_analytic_table(p, shape, steps):
    shp = (ctypes.c_int * 2)(*shape)
    out = np.zeros((shape[0] * shape[1], 3))
    lib().orc_surface_radiance_analytic_table(ctypes.byref(p), shp, steps, _ptr(out))
    return out


def max_threads():
    return lib().orc_max_threads()
