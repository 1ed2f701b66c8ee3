"""Host-side mirror of sfsim's LUT precompute interface over libsfsim_atmo (C ABI, ctypes).

Names and argument meaning follow the reference: ``generate_atmosphere_luts`` is
src/clj/sfsim/atmosphere_lut.clj:43-105, the table functions are the ``interpolate-function``
calls of atmosphere_lut.clj:74-101, parameter dictionaries use the keys of atmosphere.clj:35-39,
64-67 (planet: ``centre``, ``radius``, ``height``, ``brightness``; component: ``scatter-base``,
``scatter-scale``, ``scatter-quotient``, ``scatter-g``).

Every call runs on the GPU through the HIP library; there is no CPU fallback.  A missing library
or device raises ``AtmoError``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfsim_atmo.so")

OK, ERR_ARGUMENT, ERR_DEVICE, ERR_ALLOC, ERR_NONFINITE, ERR_IO = 0, -1, -2, -3, -4, -5
SPACE_TRANSMITTANCE, SPACE_SURFACE_RADIANCE, SPACE_SCATTER = 2, 3, 4

# Every symbol include/sfsim_atmo.h declares.
EXPORTS = (
    "sfsim_atmo_create", "sfsim_atmo_destroy", "sfsim_atmo_last_error", "sfsim_atmo_transmittance_table",
    "sfsim_atmo_surface_radiance_base_table", "sfsim_atmo_single_scatter_tables", "sfsim_atmo_point_scatter_table",
    "sfsim_atmo_ray_scatter_table", "sfsim_atmo_surface_radiance_table", "sfsim_atmo_resample_table",
    "sfsim_atmo_generate_luts", "sfsim_atmo_build", "sfsim_atmo_build_timed", "sfsim_atmo_fetch", "sfsim_atmo_stage_timings",
    "sfsim_atmo_stage_name", "sfsim_atmo_count_work", "sfsim_atmo_write_scatter", "sfsim_atmo_read_scatter",
    "sfsim_atmo_sphere_directions", "sfsim_atmo_surrogate_size", "sfsim_atmo_surrogate_reset",
    "sfsim_atmo_surrogate_weights", "sfsim_atmo_surrogate_set_weights", "sfsim_atmo_surrogate_eval",
    "sfsim_atmo_surrogate_train", "sfsim_atmo_build_seeded", "sfsim_atmo_build_from_sums", "sfsim_atmo_order_sums",
    "sfsim_atmo_surrogate_predict", "sfsim_atmo_set_sun_table", "sfsim_atmo_get_sun_table",
    "sfsim_atmo_generate_luts_noalloc", "sfsim_atmo_build_batch", "sfsim_atmo_fetch_batch",
    "sfsim_atmo_standard_atmosphere", "sfsim_atmo_speed_of_sound", "sfsim_atmo_set_progress",
)

# sfsim_atmo_set_progress stages and the callback type
PROGRESS_FIRST_ORDER, PROGRESS_ITERATION, PROGRESS_FINAL = 0, 1, 2
ProgressFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int)


class AtmoError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"libsfsim_atmo error {code}: {message}")
        self.code = code


class Planet(ctypes.Structure):
    _fields_ = [("centre", ctypes.c_double * 3), ("radius", ctypes.c_double), ("height", ctypes.c_double),
                ("brightness", ctypes.c_double * 3)]


class Scatter(ctypes.Structure):
    _fields_ = [("scatter_base", ctypes.c_double * 3), ("scatter_scale", ctypes.c_double),
                ("scatter_quotient", ctypes.c_double), ("scatter_g", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("planet", Planet), ("mie", Scatter), ("rayleigh", Scatter), ("intensity", ctypes.c_double * 3),
                ("transmittance_shape", ctypes.c_int * 2), ("surface_radiance_shape", ctypes.c_int * 2),
                ("scatter_shape", ctypes.c_int * 4), ("ray_steps", ctypes.c_int), ("sphere_steps", ctypes.c_int),
                ("iterations", ctypes.c_int), ("reserved", ctypes.c_int)]


# atmosphere_lut.clj:20-40
EARTH = {"centre": (0.0, 0.0, 0.0), "radius": 6378000.0, "height": 35000.0, "brightness": (0.3, 0.3, 0.3)}
MIE = {"scatter-base": (2e-5, 2e-5, 2e-5), "scatter-scale": 1200.0, "scatter-g": 0.76, "scatter-quotient": 0.9}
RAYLEIGH = {"scatter-base": (5.8e-6, 13.5e-6, 33.1e-6), "scatter-scale": 8000.0}


def make_params(planet=None, mie=None, rayleigh=None, transmittance_shape=(64, 255), surface_radiance_shape=(16, 63),
                scatter_shape=(32, 127, 32, 8), ray_steps=100, sphere_steps=15, iterations=5,
                intensity=(1.0, 1.0, 1.0)):
    """Parameters of one LUT build; defaults are the reference's production values
    (atmosphere_lut.clj:20-62)."""
    planet = dict(EARTH, **(planet or {}))
    mie = dict(MIE, **(mie or {})) if mie is not None else dict(MIE)
    rayleigh = dict(RAYLEIGH, **(rayleigh or {})) if rayleigh is not None else dict(RAYLEIGH)
    P = Params()
    P.planet.centre[:] = list(planet["centre"])
    P.planet.radius = planet["radius"]
    P.planet.height = planet["height"]
    P.planet.brightness[:] = list(planet["brightness"])
    for dst, src in ((P.mie, mie), (P.rayleigh, rayleigh)):
        dst.scatter_base[:] = list(src["scatter-base"])
        dst.scatter_scale = src["scatter-scale"]
        dst.scatter_quotient = src.get("scatter-quotient", 1.0)
        dst.scatter_g = src.get("scatter-g", 0.0)
    P.intensity[:] = list(intensity)
    P.transmittance_shape[:] = list(transmittance_shape)
    P.surface_radiance_shape[:] = list(surface_radiance_shape)
    P.scatter_shape[:] = list(scatter_shape)
    P.ray_steps = ray_steps
    P.sphere_steps = sphere_steps
    P.iterations = iterations
    return P


def workload(name):
    """Named workloads.  ``production``: the reference's own parameters (atmosphere_lut.clj:20-62).
    ``earth100-64``: the 100 km / 64-step transmittance grid; ``full-5-order``: the 5-order set at
    T 64x256, E 16x64, S 32x64x32x8, 64 ray steps, 14 sphere steps (66 directions), g = 0.76;
    ``large-10-order``: S 64x128x64x16, T 128x512, E 16x64 (no size is given for it beside the entry count of
    16,843,776 per set, which E 16x64 makes), 128 ray steps, 28 sphere steps (256 directions), 10 orders."""
    earth100 = {"height": 100000.0}
    if name == "production":
        return make_params()
    if name == "earth100-64":
        return make_params(planet=earth100, transmittance_shape=(64, 64), surface_radiance_shape=(16, 64),
                           scatter_shape=(8, 16, 8, 4), ray_steps=64, sphere_steps=14, iterations=0)
    if name == "full-5-order":
        return make_params(planet=earth100, transmittance_shape=(64, 256), surface_radiance_shape=(16, 64),
                           scatter_shape=(32, 64, 32, 8), ray_steps=64, sphere_steps=14, iterations=5)
    if name == "large-10-order":
        return make_params(planet=earth100, transmittance_shape=(128, 512), surface_radiance_shape=(16, 64),
                           scatter_shape=(64, 128, 64, 16), ray_steps=128, sphere_steps=28, iterations=10)
    raise KeyError(name)


def entries(P):
    """LUT entries of a full set: T + E + S + M."""
    s = P.scatter_shape
    return (P.transmittance_shape[0] * P.transmittance_shape[1] + P.surface_radiance_shape[0] *
            P.surface_radiance_shape[1] + 2 * s[0] * s[1] * s[2] * s[3])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AtmoError(ERR_DEVICE, f"{LIB_PATH} is not built (python -m sfsim_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp, fp, pp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(Params)
        L.sfsim_atmo_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.sfsim_atmo_destroy.argtypes = [vp]
        L.sfsim_atmo_last_error.argtypes = [vp]
        L.sfsim_atmo_last_error.restype = ctypes.c_char_p
        L.sfsim_atmo_transmittance_table.argtypes = [vp, pp, fp]
        L.sfsim_atmo_surface_radiance_base_table.argtypes = [vp, pp, fp]
        L.sfsim_atmo_single_scatter_tables.argtypes = [vp, pp, fp, fp]
        L.sfsim_atmo_point_scatter_table.argtypes = [vp, pp, fp, fp, fp, fp]
        L.sfsim_atmo_ray_scatter_table.argtypes = [vp, pp, fp, fp]
        L.sfsim_atmo_surface_radiance_table.argtypes = [vp, pp, fp, fp, fp]
        L.sfsim_atmo_resample_table.argtypes = [vp, pp, ctypes.c_int, fp, fp, fp]
        L.sfsim_atmo_generate_luts.argtypes = [vp, pp, fp, fp, fp, fp]
        L.sfsim_atmo_generate_luts_noalloc.argtypes = [vp, pp, vp, vp, vp, vp]
        L.sfsim_atmo_build.argtypes = [vp, pp]
        L.sfsim_atmo_build_timed.argtypes = [vp, pp]
        L.sfsim_atmo_fetch.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t]
        L.sfsim_atmo_build_batch.argtypes = [vp, pp, ctypes.c_int]
        L.sfsim_atmo_set_progress.argtypes = [vp, ProgressFn, vp]
        L.sfsim_atmo_fetch_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_size_t]
        L.sfsim_atmo_stage_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_int]
        L.sfsim_atmo_stage_name.argtypes = [ctypes.c_int]
        L.sfsim_atmo_stage_name.restype = ctypes.c_char_p
        L.sfsim_atmo_count_work.argtypes = [vp, pp, ctypes.POINTER(ctypes.c_ulonglong)]
        L.sfsim_atmo_write_scatter.argtypes = [ctypes.c_char_p, fp, ctypes.c_size_t]
        L.sfsim_atmo_read_scatter.argtypes = [ctypes.c_char_p, fp, ctypes.c_size_t]
        L.sfsim_atmo_sphere_directions.argtypes = [ctypes.c_int, ctypes.c_int]
        dp = ctypes.POINTER(ctypes.c_double)
        L.sfsim_atmo_standard_atmosphere.argtypes = [ctypes.c_double, dp, dp, dp]
        L.sfsim_atmo_speed_of_sound.argtypes = [ctypes.c_double]
        L.sfsim_atmo_speed_of_sound.restype = ctypes.c_double
        L.sfsim_atmo_surrogate_size.argtypes = []
        L.sfsim_atmo_surrogate_size.restype = ctypes.c_size_t
        L.sfsim_atmo_surrogate_reset.argtypes = [vp, ctypes.c_ulonglong, ctypes.c_int]
        L.sfsim_atmo_surrogate_weights.argtypes = [vp, fp, ctypes.c_size_t]
        L.sfsim_atmo_surrogate_set_weights.argtypes = [vp, fp, ctypes.c_size_t, ctypes.c_int]
        L.sfsim_atmo_surrogate_eval.argtypes = [vp, fp, fp, ctypes.c_int, fp, fp, fp]
        L.sfsim_atmo_surrogate_train.argtypes = [vp, pp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong,
                                                 fp]
        L.sfsim_atmo_build_seeded.argtypes = [vp, pp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int)]
        L.sfsim_atmo_build_from_sums.argtypes = [vp, pp, ctypes.c_int, fp, ctypes.c_double,
                                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        L.sfsim_atmo_order_sums.argtypes = [vp, pp, fp]
        L.sfsim_atmo_surrogate_predict.argtypes = [vp, pp, fp]
        L.sfsim_atmo_set_sun_table.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.sfsim_atmo_get_sun_table.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _fp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, entries_):
    if a is None:
        return None
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(-1)
    if arr.size != entries_ * 3:
        raise AtmoError(ERR_ARGUMENT, f"table has {arr.size} floats, expected {entries_ * 3}")
    return arr


def sphere_directions(half, steps):
    return lib().sfsim_atmo_sphere_directions(int(half), steps)


def standard_atmosphere(height):
    """(temperature K, pressure Pa, density kg/m^3) at a height in m: temperature-at-height,
    pressure-at-height and density-at-height of atmosphere.clj:425-506."""
    t, p, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().sfsim_atmo_standard_atmosphere(float(height), ctypes.byref(t), ctypes.byref(p), ctypes.byref(d))
    if rc:
        raise AtmoError(rc, f"height {height} is not finite")
    return t.value, p.value, d.value


def speed_of_sound(temperature):
    """speed-of-sound of atmosphere.clj:508-511 (m/s at a temperature in K)."""
    return lib().sfsim_atmo_speed_of_sound(float(temperature))


def write_scatter(path, data):
    arr = np.ascontiguousarray(np.asarray(data, dtype=np.float32)).reshape(-1)
    rc = lib().sfsim_atmo_write_scatter(os.fsencode(path), _fp(arr), arr.size)
    if rc:
        raise AtmoError(rc, f"could not write {path}")


def read_scatter(path, count):
    out = np.zeros(count, dtype=np.float32)
    rc = lib().sfsim_atmo_read_scatter(os.fsencode(path), _fp(out), count)
    if rc:
        raise AtmoError(rc, f"could not read {count} floats from {path}")
    return out


class Atmosphere:
    """One library context on one GPU (``sfsim_atmo_create``)."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        rc = lib().sfsim_atmo_create(device, ctypes.byref(self._ctx))
        if rc:
            raise AtmoError(rc, f"no usable gfx950 device {device}")

    def close(self):
        if self._ctx:
            lib().sfsim_atmo_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise AtmoError(rc, lib().sfsim_atmo_last_error(self._ctx).decode())

    @property
    def sun_table(self):
        """(rows, cols) of the single-scatter stage's sun optical-depth table; (0, 0): every sun
        segment summed node by node (``sfsim_atmo_set_sun_table``)."""
        r, c = ctypes.c_int(), ctypes.c_int()
        self._check(lib().sfsim_atmo_get_sun_table(self._ctx, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    @sun_table.setter
    def sun_table(self, shape):
        rows, cols = shape
        self._check(lib().sfsim_atmo_set_sun_table(self._ctx, int(rows), int(cols)))

    # ---------------------------------------------------------------- tables
    def transmittance_table(self, P):
        out = np.zeros(tuple(P.transmittance_shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_transmittance_table(self._ctx, ctypes.byref(P), _fp(out)))
        return out

    def surface_radiance_base_table(self, P):
        out = np.zeros(tuple(P.surface_radiance_shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_surface_radiance_base_table(self._ctx, ctypes.byref(P), _fp(out)))
        return out

    def single_scatter_tables(self, P):
        """(Rayleigh single scatter R1, Mie strength M1)."""
        shape = tuple(P.scatter_shape) + (3,)
        r1, m1 = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
        self._check(lib().sfsim_atmo_single_scatter_tables(self._ctx, ctypes.byref(P), _fp(r1), _fp(m1)))
        return r1, m1

    def point_scatter_table(self, P, dS, dE, mie_strength=None):
        n = int(np.prod(P.scatter_shape))
        ne = P.surface_radiance_shape[0] * P.surface_radiance_shape[1]
        a, m, e = _f32(dS, n), _f32(mie_strength, n), _f32(dE, ne)
        out = np.zeros(tuple(P.scatter_shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_point_scatter_table(self._ctx, ctypes.byref(P), _fp(a), _fp(m), _fp(e), _fp(out)))
        return out

    def ray_scatter_table(self, P, dJ):
        n = int(np.prod(P.scatter_shape))
        j = _f32(dJ, n)
        out = np.zeros(tuple(P.scatter_shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_ray_scatter_table(self._ctx, ctypes.byref(P), _fp(j), _fp(out)))
        return out

    def surface_radiance_table(self, P, dS, mie_strength=None):
        n = int(np.prod(P.scatter_shape))
        a, m = _f32(dS, n), _f32(mie_strength, n)
        out = np.zeros(tuple(P.surface_radiance_shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_surface_radiance_table(self._ctx, ctypes.byref(P), _fp(a), _fp(m), _fp(out)))
        return out

    def resample_table(self, P, space, a, b=None):
        shape = {2: P.transmittance_shape, 3: P.surface_radiance_shape, 4: P.scatter_shape}.get(space)
        if shape is None:
            raise AtmoError(ERR_ARGUMENT, f"unknown space {space}")
        n = int(np.prod(shape))
        aa, bb = _f32(a, n), _f32(b, n)
        out = np.zeros(tuple(shape) + (3,), np.float32)
        self._check(lib().sfsim_atmo_resample_table(self._ctx, ctypes.byref(P), space, _fp(aa), _fp(bb), _fp(out)))
        return out

    # ---------------------------------------------------------------- whole build
    def generate_atmosphere_luts(self, P):
        """atmosphere_lut.clj:43-105: the four tables in .scatter file layout, as float32 arrays
        shaped like the 2D images the files hold."""
        T = np.zeros((P.transmittance_shape[0], P.transmittance_shape[1], 3), np.float32)
        E = np.zeros((P.surface_radiance_shape[0], P.surface_radiance_shape[1], 3), np.float32)
        s = P.scatter_shape
        S = np.zeros((s[0] * s[2], s[1] * s[3], 3), np.float32)
        M = np.zeros_like(S)
        self._check(lib().sfsim_atmo_generate_luts(self._ctx, ctypes.byref(P), _fp(T), _fp(E), _fp(S), _fp(M)))
        return {"transmittance": T, "surface-radiance": E, "ray-scatter": S, "mie-strength": M}

    def generate_luts_noalloc(self, P, outputs):
        """sfsim_atmo_generate_luts_noalloc: the four tables of generate_atmosphere_luts written to
        caller-owned device buffers (``outputs``: their four device addresses, in the order transmittance,
        surface radiance, ray scatter, Mie strength); the intermediate tables live in the library's
        static device scratch array, so the call allocates no device memory."""
        ptrs = [ctypes.c_void_p(int(a)) if a else None for a in outputs]
        self._check(lib().sfsim_atmo_generate_luts_noalloc(self._ctx, ctypes.byref(P), *ptrs))

    def build(self, P):
        """sfsim_atmo_build: graph-replayed device-resident build."""
        self._check(lib().sfsim_atmo_build(self._ctx, ctypes.byref(P)))

    def build_timed(self, P):
        """sfsim_atmo_build_timed: kernel-by-kernel build with per-kernel HIP-event timings."""
        self._check(lib().sfsim_atmo_build_timed(self._ctx, ctypes.byref(P)))

    def fetch(self, P, which):
        s = P.scatter_shape
        shapes = [(P.transmittance_shape[0], P.transmittance_shape[1], 3),
                  (P.surface_radiance_shape[0], P.surface_radiance_shape[1], 3),
                  (s[0] * s[2], s[1] * s[3], 3), (s[0] * s[2], s[1] * s[3], 3)]
        out = np.zeros(shapes[which], np.float32)
        self._check(lib().sfsim_atmo_fetch(self._ctx, which, _fp(out), out.size))
        return out

    def build_batch(self, atmospheres):
        """sfsim_atmo_build_batch: the LUT sets of several atmospheres (same planet radius and height, shapes,
        steps and iterations) in one launch sequence, kept on the device; read with fetch_batch."""
        arr = (Params * len(atmospheres))(*atmospheres)
        self._check(lib().sfsim_atmo_build_batch(self._ctx, arr, len(atmospheres)))

    def fetch_batch(self, P, atmosphere, which):
        """Final table `which` (0 transmittance, 1 surface radiance, 2 ray scatter, 3 Mie strength) of
        atmosphere `atmosphere` of the last build_batch, shaped as fetch's."""
        s = P.scatter_shape
        shapes = [(P.transmittance_shape[0], P.transmittance_shape[1], 3),
                  (P.surface_radiance_shape[0], P.surface_radiance_shape[1], 3),
                  (s[0] * s[2], s[1] * s[3], 3), (s[0] * s[2], s[1] * s[3], 3)]
        out = np.zeros(shapes[which], np.float32)
        self._check(lib().sfsim_atmo_fetch_batch(self._ctx, atmosphere, which, _fp(out), out.size))
        return out

    def set_progress(self, fn):
        """sfsim_atmo_set_progress: fn(stage, iteration, iterations) after each stage of later builds (called
        on a thread of the HIP runtime; keep it short); None removes it."""
        self._progress = None if fn is None else ProgressFn(lambda user, stage, it, n: fn(stage, it, n))
        self._check(lib().sfsim_atmo_set_progress(self._ctx, self._progress or ProgressFn(), None))

    def stage_timings(self):
        """{kernel kind: (milliseconds, launches)} of the last build_timed() or seeded build."""
        ms = (ctypes.c_double * 16)()
        n = (ctypes.c_int * 16)()
        kinds = lib().sfsim_atmo_stage_timings(self._ctx, ms, n, 16)
        return {lib().sfsim_atmo_stage_name(k).decode(): (ms[k], n[k]) for k in range(kinds)}

    def count_work(self, P):
        c = (ctypes.c_ulonglong * 4)()
        self._check(lib().sfsim_atmo_count_work(self._ctx, ctypes.byref(P), c))
        return {"single_scatter_samples": c[0], "extinction_samples": c[1], "lookups4": c[2], "lookups2": c[3]}


def surrogate_size():
    return lib().sfsim_atmo_surrogate_size()


def random_atmosphere(rng, base=None):
    """An atmosphere drawn from the surrogate's training ranges (Rayleigh scale height 6..10 km and
    coefficient x0.5..2, Mie scale height 0.8..2 km, coefficient 5e-6..5e-5 /m, g 0.6..0.85);
    ``base`` supplies grid shapes and step counts."""
    P = base if base is not None else workload("full-5-order")
    Q = Params()
    ctypes.pointer(Q)[0] = P
    s = rng.uniform(0.5, 2.0)
    Q.rayleigh.scatter_base[:] = [5.8e-6 * s, 13.5e-6 * s, 33.1e-6 * s]
    Q.rayleigh.scatter_scale = rng.uniform(6000.0, 10000.0)
    b = rng.uniform(5e-6, 5e-5)
    Q.mie.scatter_base[:] = [b, b, b]
    Q.mie.scatter_scale = rng.uniform(800.0, 2000.0)
    Q.mie.scatter_g = rng.uniform(0.6, 0.85)
    return Q


class Surrogate:
    """The learned stand-in for scattering orders >= 2 living on a context (sfsim_atmo_surrogate_*)."""

    def __init__(self, ctx, seed=0, batch=8192):
        self.ctx = ctx
        ctx._check(lib().sfsim_atmo_surrogate_reset(ctx._ctx, seed, batch))
        self.batch = batch

    def weights(self):
        out = np.zeros(surrogate_size(), np.float32)
        self.ctx._check(lib().sfsim_atmo_surrogate_weights(self.ctx._ctx, _fp(out), out.size))
        return out

    def set_weights(self, w, trained=True):
        w = np.ascontiguousarray(w, np.float32)
        self.ctx._check(lib().sfsim_atmo_surrogate_set_weights(self.ctx._ctx, _fp(w), w.size, int(trained)))

    def eval(self, inputs, targets=None):
        """(outputs [rows, 16], gradient or None, loss or None) for input rows of 10 values in [0, 1]."""
        x = np.ascontiguousarray(inputs, np.float32)
        rows = x.shape[0]
        out = np.zeros((rows, 16), np.float32)
        t = None if targets is None else np.ascontiguousarray(targets, np.float32)
        g = np.zeros(surrogate_size(), np.float32) if t is not None else None
        loss = np.zeros(1, np.float32)
        self.ctx._check(lib().sfsim_atmo_surrogate_eval(self.ctx._ctx, _fp(x), _fp(t), rows, _fp(out), _fp(g),
                                                        _fp(loss) if t is not None else None))
        return out, g, (float(loss[0]) if t is not None else None)

    def train(self, atmospheres, steps, lr=2e-3, seed=1):
        arr = (Params * len(atmospheres))(*atmospheres)
        losses = np.zeros(max(steps, 1), np.float32)
        self.ctx._check(lib().sfsim_atmo_surrogate_train(self.ctx._ctx, arr, len(atmospheres), steps, lr, seed,
                                                         _fp(losses)))
        return losses[:steps]

    def predict(self, P):
        """The network's order sums P_1..P_5 over the scatter grid of P: array [5, *scatter_shape, 3]."""
        out = np.zeros((ORDERS,) + tuple(P.scatter_shape) + (3,), np.float32)
        self.ctx._check(lib().sfsim_atmo_surrogate_predict(self.ctx._ctx, ctypes.byref(P), _fp(out)))
        return out


ORDERS = 5


def order_sums(ctx, P):
    """P_k = dS_1 + ... + dS_k (k = 1..5) of the plain iteration (P.iterations >= 5; the first five
    orders run): array [5, *scatter_shape, 3]."""
    out = np.zeros((ORDERS,) + tuple(P.scatter_shape) + (3,), np.float32)
    ctx._check(lib().sfsim_atmo_order_sums(ctx._ctx, ctypes.byref(P), _fp(out)))
    return out


def build_seeded(ctx, P, applications=1, tolerance=1e-3):
    """sfsim_atmo_build_seeded: the set of P.iterations (5..32) orders from the network's partial sum of
    P.iterations - applications orders (1..4) and `applications` exact steps of the iteration; returns
    (residual, fell_back)."""
    r = ctypes.c_double()
    f = ctypes.c_int()
    ctx._check(lib().sfsim_atmo_build_seeded(ctx._ctx, ctypes.byref(P), applications, tolerance, ctypes.byref(r),
                                             ctypes.byref(f)))
    return r.value, bool(f.value)


def build_from_sums(ctx, P, sums, applications=1, tolerance=1e-3):
    """sfsim_atmo_build_from_sums: as build_seeded with caller-supplied order sums [5, *scatter_shape, 3]."""
    arr = np.ascontiguousarray(np.asarray(sums, np.float32))
    if arr.size != ORDERS * int(np.prod(P.scatter_shape)) * 3:
        raise AtmoError(ERR_ARGUMENT, "sums must hold 5 scatter tables")
    r = ctypes.c_double()
    f = ctypes.c_int()
    ctx._check(lib().sfsim_atmo_build_from_sums(ctx._ctx, ctypes.byref(P), applications, _fp(arr), tolerance,
                                                ctypes.byref(r), ctypes.byref(f)))
    return r.value, bool(f.value)


def write_scatter_files(directory, luts):
    """atmosphere_lut.clj:102-105: data/atmosphere/<name>.scatter"""
    os.makedirs(directory, exist_ok=True)
    for name, table in luts.items():
        write_scatter(os.path.join(directory, f"{name}.scatter"), table)
