"""GPU test of sfsim_atmo_generate_luts_noalloc, the entry point that allocates no device memory (its
intermediate tables live in the library's static scratch array), against sfsim_atmo_generate_luts on
the same parameters: two calls in a row, the second on a smaller grid, both equal bit for bit (the same
kernels run on the same inputs; the scratch array keeps the first call's contents, which the second must
not depend on)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLES = ["transmittance", "surface-radiance", "ray-scatter", "mie-strength"]


class DeviceBuffers:
    """Device buffers from the HIP runtime the library itself runs on (libamdhip64.so.7, already loaded
    with it), filled with 0xFF bytes (NaN) so that an entry the entry point does not write shows."""

    def __init__(self, sizes):
        from sfsim_amd import atmosphere as atm
        atm.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs, self.sizes = [], list(sizes)
        for n in self.sizes:
            p = ctypes.c_void_p()
            assert self.hip.hipMalloc(ctypes.byref(p), n * 4) == 0
            assert self.hip.hipMemset(p, 0xFF, n * 4) == 0
            self.ptrs.append(p.value)

    def read(self, i):
        out = np.zeros(self.sizes[i], np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptrs[i], out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _small(atm):
    return atm.make_params(planet={"height": 100000.0, "brightness": (0.2, 0.2, 0.2)}, transmittance_shape=(10, 17),
                           surface_radiance_shape=(6, 11), scatter_shape=(6, 11, 7, 4), ray_steps=16, sphere_steps=10,
                           iterations=3)


def test_noalloc_build_equals_generate_luts():
    from sfsim_amd import atmosphere as atm
    with atm.Atmosphere(0) as ctx:
        for P in (atm.workload("full-5-order"), _small(atm)):
            print("noalloc build", tuple(P.scatter_shape), P.iterations, flush=True)
            ref = ctx.generate_atmosphere_luts(P)
            bufs = DeviceBuffers([ref[k].size for k in TABLES])
            try:
                ctx.generate_luts_noalloc(P, bufs.ptrs)
                for i, k in enumerate(TABLES):
                    assert np.array_equal(bufs.read(i), ref[k].reshape(-1)), (tuple(P.scatter_shape), k)
            finally:
                bufs.free()


def test_noalloc_sweep_shapes():
    """The same equality at ten of the shape sweep's configurations, called one after another on one
    context (each call on a grid of a different shape than the last)."""
    from sfsim_amd import atmosphere as atm
    from tests.test_gpu_shape_sweep import sweep_cases
    with atm.Atmosphere(0) as ctx:
        for case in sweep_cases()[:10]:
            c = dict(case)
            P = atm.make_params(planet={"height": c.pop("height"), "brightness": (0.3,) * 3}, **c)
            print("noalloc sweep", tuple(P.scatter_shape), P.ray_steps, P.iterations, flush=True)
            ref = ctx.generate_atmosphere_luts(P)
            bufs = DeviceBuffers([ref[k].size for k in TABLES])
            try:
                ctx.generate_luts_noalloc(P, bufs.ptrs)
                for i, k in enumerate(TABLES):
                    assert np.array_equal(bufs.read(i), ref[k].reshape(-1)), (tuple(P.scatter_shape), k)
            finally:
                bufs.free()


def test_noalloc_rejects_what_does_not_fit():
    from sfsim_amd import atmosphere as atm
    with atm.Atmosphere(0) as ctx:
        P = atm.workload("large-10-order")  # its intermediate tables need 2.3 GiB
        bufs = DeviceBuffers([16] * 4)
        try:
            with pytest.raises(atm.AtmoError) as e:
                ctx.generate_luts_noalloc(P, bufs.ptrs)
            assert e.value.code == atm.ERR_ARGUMENT and "scratch" in str(e.value)
            with pytest.raises(atm.AtmoError) as e:
                ctx.generate_luts_noalloc(_small(atm), [bufs.ptrs[0], None, bufs.ptrs[2], bufs.ptrs[3]])
            assert e.value.code == atm.ERR_ARGUMENT
        finally:
            bufs.free()
