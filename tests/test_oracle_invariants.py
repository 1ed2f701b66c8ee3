"""Invariant checks that pin the parts of the oracle no reference-held vector reaches:
point-scatter (J), the surface-radiance integral of a table, and the multi-order driver.

Each check is chosen so that a plausible restatement mistake fails it: a dropped albedo/pi factor or
transmittance (ground closed form), a wrong horizon split or ring weight (sky/ground symmetry), a
scattering coefficient applied twice or not at all (homogeneity per scattering order), a shifted
quadrature node (thin-atmosphere closed form of the midpoint rule), and the driver's bookkeeping
(one iteration rebuilt from the stage functions).
"""
import math

import numpy as np
import pytest

R = 6378000.0


def test_point_scatter_ground_and_sky_split(oracle):
    """x on the ground, horizontal view, Rayleigh only: with an even number of latitude rings the
    sky and ground hemispheres are mirror images, so J from a unit sky radiance equals J from a unit
    ground irradiance divided by albedo/pi (ground directions end at x itself, T = 1), and both
    are half of the phase function's full-sphere quadrature (~1)."""
    earth = oracle.planet(R, 100000.0, brightness=(0.3, 0.3, 0.3))
    ray = oracle.rayleigh()
    none = oracle.scatter((0, 0, 0), 1200.0, 1.0, 0.76)
    beta = np.array(ray.base[:])
    x, v, l = (R, 0, 0), (0, 1, 0), (0.6, 0.8, 0)
    js = oracle.point_scatter_constant(earth, [none, ray], (1, 1, 1), (0, 0, 0), 28, 10, x, v, l) / beta
    jg = oracle.point_scatter_constant(earth, [none, ray], (0, 0, 0), (1, 1, 1), 28, 10, x, v, l) / beta
    jg = jg / (0.3 / math.pi)
    assert np.allclose(js, jg, rtol=1e-12, atol=0)
    assert np.allclose(js, 0.5, atol=1e-3)
    # odd ring count (sphere_steps 14): the equator ring counts as sky (cos(pi/2) > 0 in FP64)
    js14 = oracle.point_scatter_constant(earth, [none, ray], (1, 1, 1), (0, 0, 0), 14, 10, x, v, l) / beta
    jg14 = oracle.point_scatter_constant(earth, [none, ray], (0, 0, 0), (1, 1, 1), 14, 10, x, v, l) / beta
    assert js14[0] > jg14[0] / (0.3 / math.pi)
    assert abs(js14[0] + jg14[0] / (0.3 / math.pi) - 1.0) < 3e-3


def test_point_scatter_elevated_ground_transmittance(oracle):
    """Above the ground the ground directions carry T(x, P) < 1 and the sky directions reach the top."""
    earth = oracle.planet(R, 100000.0, brightness=(0.3, 0.3, 0.3))
    mie, ray = oracle.mie(), oracle.rayleigh()
    x, v, l = (R + 5000.0, 0, 0), (0, 1, 0), (0.6, 0.8, 0)
    jg_low = oracle.point_scatter_constant(earth, [mie, ray], (0, 0, 0), (1, 1, 1), 28, 32, (R, 0, 0), v, l)
    jg_high = oracle.point_scatter_constant(earth, [mie, ray], (0, 0, 0), (1, 1, 1), 28, 32, x, v, l)
    # scattering coefficients fall with height and the ground is further away: strictly less
    assert np.all(jg_high < jg_low)


def _scaled(oracle, lam):
    """Scattering scaled by lam with extinction (base/quotient) unchanged."""
    mie = oracle.scatter(np.array([2e-5] * 3) * lam, 1200.0, 0.9 * lam, 0.76)
    ray = oracle.scatter(np.array([5.8e-6, 13.5e-6, 33.1e-6]) * lam, 8000.0, lam, 0.0)
    return mie, ray


def _small_params(oracle, lam, brightness, iterations=1):
    earth = oracle.planet(R, 60000.0, brightness=(brightness,) * 3)
    mie, ray = _scaled(oracle, lam)
    return oracle.params(earth, mie, ray, (6, 9), (4, 7), (4, 7, 5, 3), 8, 8, iterations)


def test_single_scatter_is_first_order_in_scattering(oracle):
    P1, P2 = _small_params(oracle, 1.0, 0.0), _small_params(oracle, 0.5, 0.0)
    assert np.allclose(oracle.stage_transmittance(P1), oracle.stage_transmittance(P2), rtol=1e-14, atol=0)
    for strength in (0, 1):
        a = oracle.stage_single_scatter(P1, strength)
        b = oracle.stage_single_scatter(P2, strength)
        assert np.allclose(b, 0.5 * a, rtol=1e-12, atol=1e-300)


def test_point_scatter_orders(oracle):
    """J[dS, dE]: the sky part carries one more scattering factor than dS (second order), the ground
    part one scattering factor times the direct irradiance E0, which does not scale."""
    for brightness in (0.0, 0.3):
        parts = []
        for lam in (1.0, 0.5):
            P = _small_params(oracle, lam, brightness)
            R1 = oracle.stage_single_scatter(P, 0)
            M1 = oracle.stage_single_scatter(P, 1)
            E0 = oracle.stage_surface_radiance_base(P)
            parts.append(oracle.stage_point_scatter(P, 1, R1, M1, E0))
        J1, J05 = parts
        if brightness == 0.0:
            assert np.allclose(J05, 0.25 * J1, rtol=1e-12, atol=1e-300)
        else:
            # J(lam) = lam^2 A + lam B  ->  J(1) - 2 J(1/2) = A/2 with A the sky part at lam = 1
            P = _small_params(oracle, 1.0, 0.0)
            A = oracle.stage_point_scatter(P, 1, oracle.stage_single_scatter(P, 0), oracle.stage_single_scatter(P, 1),
                                           oracle.stage_surface_radiance_base(P))
            assert np.allclose(J1 - 2 * J05, 0.5 * A, rtol=1e-9, atol=1e-20)


def test_thin_atmosphere_ray_scatter_midpoint_closed_form(oracle):
    """Vertical ray from the ground, sun overhead, vanishing extinction: R1 equals the midpoint-rule
    sum of beta_R(h) * phase_R(1) over the ray (closed form of a geometric series)."""
    lam = 1e-9
    earth = oracle.planet(R, 100000.0)
    mie, ray = _scaled(oracle, lam)
    ray_ext = oracle.scatter(np.array([5.8e-6, 13.5e-6, 33.1e-6]) * lam, 8000.0, 1.0, 0.0)
    mie_ext = oracle.scatter(np.array([2e-5] * 3) * lam, 1200.0, 0.9, 0.76)
    n = 16
    got = oracle.ray_scatter_single(earth, [mie_ext, ray_ext], ray_ext, n, (1, 1, 1), 0, (R, 0, 0), (1, 0, 0),
                                    (1, 0, 0), True)
    step = 100000.0 / n
    q = math.exp(-step / 8000.0)
    series = step * math.exp(-step / 2 / 8000.0) * (1 - q ** n) / (1 - q)
    expected = np.array(ray_ext.base[:]) * oracle.phase(0.0, 1.0) * series
    assert np.allclose(got, expected, rtol=1e-6)


def test_vanishing_scattering(oracle):
    earth = oracle.planet(R, 60000.0)
    mie, ray = _scaled(oracle, 1e-30)
    mie = oracle.scatter(mie.base[:], 1200.0, 1.0, 0.76)
    ray = oracle.scatter(ray.base[:], 8000.0, 1.0, 0.0)
    P = oracle.params(earth, mie, ray, (6, 9), (4, 7), (4, 7, 5, 3), 8, 8, 2)
    T, E, S, M = oracle.generate_atmosphere_luts(P)
    assert np.allclose(T, 1.0, atol=1e-20)
    assert np.max(np.abs(S)) < 1e-25 and np.max(np.abs(M)) < 1e-25


@pytest.mark.parametrize("iterations", [0, 1, 2])
def test_driver_bookkeeping(oracle, iterations):
    """atmosphere_lut.clj:74-105 rebuilt from the stage functions: which dS/dE each stage reads, the
    accumulations, and the final re-sampling."""
    P = _small_params(oracle, 1.0, 0.3, iterations)
    T, E, S, M = oracle.generate_atmosphere_luts(P)
    Tt = oracle.stage_transmittance(P)
    dE = oracle.stage_surface_radiance_base(P)
    R1 = oracle.stage_single_scatter(P, 0)
    M1 = oracle.stage_single_scatter(P, 1)
    Sacc, Eacc, dS, first = R1, None, None, True
    for _ in range(iterations):
        dJ = oracle.stage_point_scatter(P, first, R1 if first else dS, M1 if first else None, dE)
        dE = oracle.stage_surface_radiance(P, first, R1 if first else dS, M1 if first else None)
        dS = oracle.stage_ray_scatter(P, dJ)
        first = False
        Eacc = oracle.stage_resample(P, 3, dE, Eacc)
        Sacc = oracle.stage_resample(P, 4, Sacc, dS)
    assert np.array_equal(T.reshape(-1, 3), oracle.stage_resample(P, 2, Tt))
    assert np.array_equal(M.reshape(-1, 3), oracle.stage_resample(P, 4, M1))
    assert np.array_equal(S.reshape(-1, 3), oracle.stage_resample(P, 4, Sacc))
    if iterations == 0:
        assert not E.any()
    else:
        assert np.array_equal(E.reshape(-1, 3), oracle.stage_resample(P, 3, Eacc))
