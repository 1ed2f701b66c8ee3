"""Mixed-precision parity over a batch of atmospheres: 256 atmospheres drawn from a splitmix64 stream
(seed 20261202; Rayleigh scale height U[6, 10] km and coefficient scale U[0.5, 2]; Mie scale height
U[0.8, 2] km, scattering coefficient U[5e-6, 5e-5] /m, g U[0.6, 0.85]) on the 100 km, 64 x 64
transmittance grid with 64 ray steps.  Every table's max and RMS relative error against the FP64
oracle is reported; the bound is the transmittance tolerance, 1e-4.  A subset of the atmospheres also
gets its single-scatter tables checked on a small 4D grid."""
import numpy as np
import pytest

from tests.parity import TOL_SCATTER, TOL_TRANSMITTANCE, oracle_params, report

MASK64 = (1 << 64) - 1


def splitmix64_uniforms(seed):
    state = seed & MASK64
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z ^= z >> 31
        yield (z >> 11) * (1.0 / (1 << 53))


def atmospheres(atm, count, seed=20261202, **shape):
    u = splitmix64_uniforms(seed)
    out = []
    for _ in range(count):
        hr, sr, hm, bm, g = (next(u) for _ in range(5))
        s = 0.5 + 1.5 * sr
        b = 5e-6 + 4.5e-5 * bm
        out.append(atm.make_params(
            planet={"height": 100000.0},
            rayleigh={"scatter-base": (5.8e-6 * s, 13.5e-6 * s, 33.1e-6 * s), "scatter-scale": 6000.0 + 4000.0 * hr},
            mie={"scatter-base": (b, b, b), "scatter-scale": 800.0 + 1200.0 * hm, "scatter-g": 0.6 + 0.25 * g},
            **shape))
    return out


def test_splitmix64_stream_is_the_published_one():
    """First outputs of splitmix64 from seed 0 (Vigna's reference sequence)."""
    state, outs = 0, []
    for _ in range(3):
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        outs.append(z ^ (z >> 31))
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.fixture(scope="module")
def gpu():
    from sfsim_amd import atmosphere as atm
    ctx = atm.Atmosphere(0)
    yield atm, ctx
    ctx.close()


@pytest.mark.gpu
def test_transmittance_batch_of_256_atmospheres(gpu, oracle):
    atm, ctx = gpu
    worst_max, rms = 0.0, []
    for P in atmospheres(atm, 256, transmittance_shape=(64, 64), ray_steps=64, iterations=0):
        r = report("T", ctx.transmittance_table(P), oracle.stage_transmittance(oracle_params(oracle, P)))
        worst_max = max(worst_max, r["max_rel"])
        rms.append(r["rms_rel"])
    print(f"256 atmospheres: max rel {worst_max:.3e}, mean RMS rel {np.mean(rms):.3e}, worst RMS {max(rms):.3e}")
    assert worst_max <= TOL_TRANSMITTANCE


@pytest.mark.gpu
def test_single_scatter_batch_subset(gpu, oracle):
    atm, ctx = gpu
    worst = 0.0
    for P in atmospheres(atm, 16, transmittance_shape=(32, 64), scatter_shape=(8, 32, 8, 4), ray_steps=64,
                         iterations=0):
        OP = oracle_params(oracle, P)
        R1, M1 = ctx.single_scatter_tables(P)
        worst = max(worst, report("R1", R1, oracle.stage_single_scatter(OP, 0))["max_rel"],
                    report("M1", M1, oracle.stage_single_scatter(OP, 1))["max_rel"])
    print(f"16 atmospheres, single scatter: max rel {worst:.3e}")
    assert worst <= TOL_SCATTER
