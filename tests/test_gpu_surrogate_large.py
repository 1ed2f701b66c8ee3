"""The learned part on the large workload (S 64x128x64x16, T 128x512, E 16x64, 128 ray steps, 256 sphere
directions, 10 orders): 64 atmospheres drawn from the seeded splitmix64 stream over the training ranges,
built plain in batches of 16 and seeded one by one.

The network is trained on the GPU on the first five orders of all 64 atmospheres (536 M entries).  A
10-order seeded build takes the network's partial sum of 10 - m orders (m = 6..9) and m exact
applications of the iteration, so it saves 10 - m of the 10 point/ray-scatter passes.

What is asserted: training converges; every seeded build is finite and within the same bound of its plain
build; for the base atmosphere each extra application shrinks the error (the iteration contracts the
seed's error), the seeded build beats the same applications from a zero seed, and at the 1e-3 bound the
build falls back to the plain iteration, whose tables it then returns bit for bit.  Timings of the plain
batches and the seeded builds are printed (entries/s per atmosphere).
"""
import time

import numpy as np
import pytest

from tests.parity import rel_errors
from tests.test_atmosphere_batch import atmospheres

pytestmark = pytest.mark.gpu

TABLES = ["transmittance", "surface-radiance", "ray-scatter", "mie-strength"]


@pytest.fixture(scope="module")
def atm():
    from sfsim_amd import atmosphere
    return atmosphere


@pytest.fixture(scope="module")
def gpu(atm):
    ctx = atm.Atmosphere(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def family(atm):
    base = atm.workload("large-10-order")
    shape = dict(transmittance_shape=tuple(base.transmittance_shape),
                 surface_radiance_shape=tuple(base.surface_radiance_shape), scatter_shape=tuple(base.scatter_shape),
                 ray_steps=base.ray_steps, sphere_steps=base.sphere_steps, iterations=base.iterations)
    return atmospheres(atm, 64, seed=5, **shape)


@pytest.fixture(scope="module")
def trained(atm, gpu, family):
    sg = atm.Surrogate(gpu, seed=1, batch=8192)
    print("training on the first five orders of 64 large atmospheres", flush=True)
    t0 = time.perf_counter()
    losses = sg.train(family, steps=30000, lr=3e-3, seed=2)
    print(f"trained in {time.perf_counter() - t0:.1f} s", flush=True)
    return sg, losses


def _stats(got, ref):
    e = np.concatenate([rel_errors(g, r).reshape(-1) for g, r in zip(got, ref)])
    return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}


def test_large_training_converges(trained):
    _, losses = trained
    first, last = float(losses[:10].mean()), float(losses[-100:].mean())
    print(f"loss {first:.4g} -> {last:.4g}")
    assert np.all(np.isfinite(losses)) and last < 2e-3 * first


def test_large_seeded_base_atmosphere(atm, gpu, family, trained):
    sg, _ = trained
    P = family[0]
    gpu.build(P)
    plain = [gpu.fetch(P, i) for i in range(4)]
    errs = {}
    for m in (6, 7, 8, 9):
        r, fell_back = atm.build_seeded(gpu, P, m, tolerance=np.inf)
        assert not fell_back
        seeded = [gpu.fetch(P, i) for i in range(4)]
        assert np.array_equal(seeded[0], plain[0]) and np.array_equal(seeded[3], plain[3])  # T, M1: exact stages
        errs[m] = _stats(seeded[1:3], plain[1:3])
        print(f"m={m}: surface radiance + ray scatter vs plain {errs[m]}, residual {r:.3e}", flush=True)
    for m in (6, 7, 8):
        assert errs[m + 1]["rms"] < errs[m]["rms"]
    # the same six applications from a zero seed (6 orders instead of 10)
    w = sg.weights()
    zero = np.zeros_like(w)
    zero[-16:-1] = 10.0 + np.log(1e-8)  # decodes to P = 0
    sg.set_weights(zero, trained=True)
    try:
        atm.build_seeded(gpu, P, 6, tolerance=np.inf)
        trunc = _stats([gpu.fetch(P, i) for i in (1, 2)], plain[1:3])
    finally:
        sg.set_weights(w, trained=True)
    print(f"m=6 without the network: {trunc}")
    assert errs[6]["rms"] < 0.5 * trunc["rms"]
    r, fell_back = atm.build_seeded(gpu, P, 9, tolerance=1e-3)
    assert fell_back and r > 1e-3
    for i in range(4):
        assert np.array_equal(gpu.fetch(P, i), plain[i])


def test_large_64_atmospheres_plain_and_seeded(atm, gpu, family, trained):
    """All 64: plain in batches of 16 (timed), seeded with m = 9 one by one (timed, residual of each);
    four of them compared table by table with their plain builds."""
    entries = atm.entries(family[0])
    t0 = time.perf_counter()
    for b in range(4):
        gpu.build_batch(family[16 * b:16 * (b + 1)])
    plain_ms = (time.perf_counter() - t0) * 1e3 / 64
    check = (0, 21, 42, 63)
    kept = {}
    residuals = []
    seeded_ms = 0.0
    for a, P in enumerate(family):
        atm.build_seeded(gpu, P, 9, tolerance=np.inf)  # graph captured on the first call with these parameters
        t0 = time.perf_counter()
        r, fell_back = atm.build_seeded(gpu, P, 9, tolerance=np.inf)
        seeded_ms += (time.perf_counter() - t0) * 1e3
        assert not fell_back and np.isfinite(r)
        residuals.append(r)
        if a in check:
            kept[a] = [gpu.fetch(P, i) for i in (1, 2)]
    seeded_ms /= len(family)
    print(f"plain (batches of 16): {plain_ms:.2f} ms per atmosphere ({entries / plain_ms * 1e3:.3e} entries/s); "
          f"seeded m=9: {seeded_ms:.2f} ms ({entries / seeded_ms * 1e3:.3e} entries/s); residual median "
          f"{np.median(residuals):.3e}, max {max(residuals):.3e}", flush=True)
    for a, got in kept.items():
        gpu.build(family[a])
        s = _stats(got, [gpu.fetch(family[a], i) for i in (1, 2)])
        print(f"atmosphere {a}: seeded m=9 vs plain {s}", flush=True)
        assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1]))
        assert s["max"] < 0.1
