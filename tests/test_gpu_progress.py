"""GPU tests of the progress reports (sfsim_atmo_set_progress; util.clj:441-449 progress-wrap and the
"Iteration k/N" line of atmosphere_lut.clj:80): one report after the first-order tables, one per scattering
order and one after the final tables, in order, for generate_luts, build and build_batch; a build with
reports returns the same tables bit for bit as one without."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def atm():
    from sfsim_amd import atmosphere
    return atmosphere


@pytest.fixture(scope="module")
def gpu(atm):
    ctx = atm.Atmosphere(0)
    yield ctx
    ctx.close()


def _small(atm, iterations=3):
    return atm.make_params(transmittance_shape=(10, 17), surface_radiance_shape=(6, 11), scatter_shape=(6, 11, 7, 4),
                           ray_steps=16, sphere_steps=10, iterations=iterations)


def _expected(atm, n):
    return ([(atm.PROGRESS_FIRST_ORDER, 0, n)] + [(atm.PROGRESS_ITERATION, k, n) for k in range(1, n + 1)]
            + [(atm.PROGRESS_FINAL, 0, n)])


def test_progress_reports_of_a_build(atm, gpu):
    P = _small(atm)
    gpu.build(P)
    plain = [gpu.fetch(P, i) for i in range(4)]
    seen = []
    gpu.set_progress(lambda stage, it, n: seen.append((stage, it, n)))
    try:
        luts = gpu.generate_atmosphere_luts(P)
        assert seen == _expected(atm, 3)
        seen.clear()
        gpu.build(P)
        assert seen == _expected(atm, 3)
        for i in range(4):
            assert np.array_equal(gpu.fetch(P, i), plain[i]), i
        assert np.array_equal(luts["transmittance"].reshape(-1), plain[0].reshape(-1))
    finally:
        gpu.set_progress(None)
    seen.clear()
    gpu.build(P)
    assert seen == []


def test_progress_reports_of_a_batch(atm, gpu):
    P = _small(atm, iterations=2)
    Q = _small(atm, iterations=2)
    Q.planet.brightness[:] = [0.5, 0.5, 0.5]
    seen = []
    gpu.set_progress(lambda stage, it, n: seen.append((stage, it, n)))
    try:
        gpu.build_batch([P, Q])
    finally:
        gpu.set_progress(None)
    assert seen == _expected(atm, 2)
