"""GPU tests of the batched build (sfsim_atmo_build_batch): one launch sequence over the entries of many
atmospheres (the reference builds one atmosphere at a time, parallel over dimension 0 only,
interpolate.clj:72).

* every atmosphere of a batch equals its own single-atmosphere build (sfsim_atmo_generate_luts) bit for
  bit, with 0, 1 and 3 scattering orders;
* 256 atmospheres drawn from the seeded splitmix64 stream on the 100 km, 64 x 64 transmittance grid (64 ray
  steps) built as one batch: every atmosphere's transmittance table and its single-scatter tables on a small
  4D grid against the FP64 oracle (max and RMS relative error per table per atmosphere);
* 64 atmospheres on the large grid (S 64x128x64x16, T 128x512, 128 steps, 256 directions, 10 orders) in
  batches of 16: all tables finite and physical, transmittance of every atmosphere against the oracle, and
  a few atmospheres equal to their single builds bit for bit.
"""
import numpy as np
import pytest

from tests.parity import TOL_SCATTER, TOL_TRANSMITTANCE, oracle_params, report
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


def _small(atm, iterations):
    return dict(transmittance_shape=(10, 17), surface_radiance_shape=(6, 11), scatter_shape=(6, 11, 7, 4),
                ray_steps=16, sphere_steps=10, iterations=iterations)


@pytest.mark.parametrize("iterations,slab", [(0, (7, 4)), (1, (7, 4)), (3, (7, 4)), (2, (24, 16)), (2, (32, 24))])
def test_batch_equals_single_builds(atm, gpu, iterations, slab):
    """Slabs of 28, 384 and 768 (sun elevation, heading) entries: one, two and four entries per ray-scatter
    thread."""
    shape = _small(atm, iterations)
    shape["scatter_shape"] = (6, 11) + slab
    Ps = atmospheres(atm, 5, seed=7 + iterations, **shape)
    for i, P in enumerate(Ps):  # albedo and intensity may differ between the atmospheres of a batch
        P.planet.brightness[:] = [0.1 + 0.1 * i] * 3
        P.intensity[:] = [1.0, 0.9 + 0.05 * i, 1.1]
    print("batch", len(Ps), "iterations", iterations, flush=True)
    gpu.build_batch(Ps)
    for a, P in enumerate(Ps):
        single = gpu.generate_atmosphere_luts(P)
        for which, name in enumerate(TABLES):
            assert np.array_equal(gpu.fetch_batch(P, a, which), single[name]), (a, name)


def test_batch_rejects_mixed_grids(atm, gpu):
    Ps = atmospheres(atm, 2, **_small(atm, 1))
    Ps[1].ray_steps = 20
    with pytest.raises(atm.AtmoError) as e:
        gpu.build_batch(Ps)
    assert e.value.code == atm.ERR_ARGUMENT
    Ps = atmospheres(atm, 2, **_small(atm, 1))
    Ps[1].planet.height = 90000.0
    with pytest.raises(atm.AtmoError):
        gpu.build_batch(Ps)
    with pytest.raises(atm.AtmoError):
        gpu.fetch_batch(Ps[0], 5, 0)


def test_batch_of_256_atmospheres_against_the_oracle(atm, gpu, oracle):
    """256 atmospheres, one batched build: transmittance (64 x 64, the whole table) and the first-order
    scatter tables of a small 4D grid against the FP64 oracle, per atmosphere."""
    Ps = atmospheres(atm, 256, transmittance_shape=(64, 64), surface_radiance_shape=(8, 16),
                     scatter_shape=(6, 16, 8, 4), ray_steps=64, sphere_steps=14, iterations=0)
    print("batch of 256 atmospheres", flush=True)
    gpu.build_batch(Ps)
    t_max, t_rms, s_max = [], [], []
    for a, P in enumerate(Ps):
        OP = oracle_params(oracle, P)
        T = gpu.fetch_batch(P, a, 0)
        ref = oracle.stage_resample(OP, 2, oracle.stage_transmittance(OP))
        r = report("T", T, ref)
        t_max.append(r["max_rel"])
        t_rms.append(r["rms_rel"])
        if a % 16 == 0:  # first-order tables of every 16th atmosphere (the final S and M at 0 orders)
            S, M = oracle.generate_atmosphere_luts(OP)[2:]
            s_max.append(max(report("S", gpu.fetch_batch(P, a, 2), oracle.convert_4d_to_2d(S))["max_rel"],
                             report("M", gpu.fetch_batch(P, a, 3), oracle.convert_4d_to_2d(M))["max_rel"]))
    print(f"T: max rel {max(t_max):.3e}, mean RMS {np.mean(t_rms):.3e}; S/M (16 atmospheres): max rel {max(s_max):.3e}")
    assert max(t_max) <= TOL_TRANSMITTANCE
    assert max(s_max) <= TOL_SCATTER


def test_large_batches_of_64_atmospheres(atm, gpu, oracle):
    """The large workload (10 orders) for 64 atmospheres in batches of 16."""
    base = atm.workload("large-10-order")
    shape = dict(transmittance_shape=tuple(base.transmittance_shape),
                 surface_radiance_shape=tuple(base.surface_radiance_shape), scatter_shape=tuple(base.scatter_shape),
                 ray_steps=base.ray_steps, sphere_steps=base.sphere_steps, iterations=base.iterations)
    Ps = atmospheres(atm, 64, seed=5, **shape)
    worst_t = 0.0
    for b in range(4):
        batch = Ps[16 * b:16 * (b + 1)]
        print("large batch", b, flush=True)
        gpu.build_batch(batch)
        for a, P in enumerate(batch):
            tabs = [gpu.fetch_batch(P, a, w) for w in range(4)]
            for name, t in zip(TABLES, tabs):
                assert np.all(np.isfinite(t)) and np.all(t >= 0), (b, a, name)
            assert np.all(tabs[0] <= 1.0)
            OP = oracle_params(oracle, P)
            worst_t = max(worst_t, report("T", tabs[0], oracle.stage_resample(OP, 2, oracle.stage_transmittance(OP)))
                          ["max_rel"])
        if b == 0:  # the first and the last atmosphere of the first batch alone
            kept = {a: [gpu.fetch_batch(batch[a], a, w) for w in range(4)] for a in (0, 15)}
            for a, tabs in kept.items():
                gpu.build(batch[a])
                for w in range(4):
                    assert np.array_equal(gpu.fetch(batch[a], w), tabs[w]), (a, TABLES[w])
    print(f"64 large atmospheres: transmittance max rel {worst_t:.3e}")
    assert worst_t <= TOL_TRANSMITTANCE
