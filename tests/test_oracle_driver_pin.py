"""Pins the oracle's multi-order driver (orc_generate_atmosphere_luts, atmosphere_lut.clj:43-105) with
properties that the mathematics of the iteration fixes, whatever code computes it.

Scale every scattering coefficient by t with the extinction held fixed (the quotient scaled with it,
atmosphere.clj:50-53) and the ground albedo by t as well.  Transmittance and the direct irradiance E0 do
not change; single scattering R1 and M1 grow as t; in each iteration point-scatter multiplies by one
scattering coefficient (atmosphere.clj:203-222, the in-scattering factor) and its ground term by the
albedo, and every other step (ray-scatter, surface-radiance, the re-samplings of atmosphere_lut.clj:95-101)
is linear in its table.  So order k of the truncated series, dS_k, is homogeneous of degree k + 1 in t and
dE_k of degree k, and with tab() the re-sampling f(backward(i)) of a table:

1. with 0 iterations the final S is tab(R1) and E is zero (E0 is excluded, atmosphere_lut.clj:64);
   R1 itself is pinned by the reference's first-order goldens;
2. the first iteration adds tab(tab(dS_1)) to S and tab(tab(dE_1)) to E, with dS_1 = ray-scatter of
   point-scatter(R1 + M1 phase, E0) and dE_1 = surface-radiance(R1 + M1 phase): the stage compositions
   that tests/test_oracle_bruteforce.py checks against a table-free evaluator and t_planet.clj:312-376 pins;
3. every further iteration adds a single monomial: S_{K+1}(t) - tab(S_K(t)) is proportional to t^(K+2)
   and E_{K+1}(t) - tab(E_K(t)) to t^(K+1);
4. in the albedo alone (t = 1) S_K is a polynomial of degree ceil(K/2) and E_K of degree ceil((K-1)/2)
   (an order gains a ground reflection only through the dE of the order before).

With t a power of two every scaling is exact in FP64, so the checks hold to rounding of the sums.  Each
plausible wiring mistake breaks at least one of them: the mis-wired drivers at the end, built from the
same oracle stages, fail.
"""
import math

import numpy as np
import pytest

R = 6378000.0
T_VALUES = (0.5, 1.0, 2.0)
ALBEDO = 0.3
MAX_K = 5


def _params(oracle, t, albedo, iterations):
    earth = oracle.planet(R, 60000.0, brightness=(albedo,) * 3)
    mie = oracle.scatter(np.array([2e-5] * 3) * t, 1200.0, 0.9 * t, 0.76)
    ray = oracle.scatter(np.array([5.8e-6, 13.5e-6, 33.1e-6]) * t, 8000.0, t, 0.0)
    return oracle.params(earth, mie, ray, (5, 9), (4, 7), (4, 7, 5, 3), 8, 8, iterations)


def _oracle_driver(oracle, P):
    T, E, S, M = oracle.generate_atmosphere_luts(P)
    return T.reshape(-1, 3), E.reshape(-1, 3), S.reshape(-1, 3), M.reshape(-1, 3)


def _mis_wired(mistake):
    """The driver of atmosphere_lut.clj:74-105 rebuilt from the oracle's stages with one wiring mistake
    (None: as the reference wires it)."""
    def run(oracle, P):
        R1 = oracle.stage_single_scatter(P, 0)
        M1 = oracle.stage_single_scatter(P, 1)
        E0 = oracle.stage_surface_radiance_base(P)
        T = oracle.stage_transmittance(P)
        dE, dS, first = E0, None, True
        Sacc = R1
        Eacc = E0 if mistake == "E0 in E" else None
        for _ in range(P.iterations):
            dS_src = R1 if first else (Sacc if mistake == "point scatter reads S" else dS)
            mie = M1 if first else None
            if mistake == "point scatter reads the new dE":
                dE = oracle.stage_surface_radiance(P, first, dS_src, mie)
                dJ = oracle.stage_point_scatter(P, first, dS_src, mie, dE)
                dS = oracle.stage_ray_scatter(P, dJ)
            elif mistake == "surface radiance reads the new dS":
                dJ = oracle.stage_point_scatter(P, first, dS_src, mie, dE)
                dS = oracle.stage_ray_scatter(P, dJ)
                dE = oracle.stage_surface_radiance(P, False, dS, None)
            else:
                dJ = oracle.stage_point_scatter(P, first, dS_src, mie, dE)
                dE = oracle.stage_surface_radiance(P, first, dS_src, mie)
                dS = oracle.stage_ray_scatter(P, dJ)
            first = False
            Eacc = oracle.stage_resample(P, 3, dE, Eacc)
            Sacc = oracle.stage_resample(P, 4, Sacc, dJ if mistake == "S accumulates dJ" else dS)
        E = np.zeros_like(E0) if Eacc is None else oracle.stage_resample(P, 3, Eacc)
        return (oracle.stage_resample(P, 2, T), E, oracle.stage_resample(P, 4, Sacc),
                oracle.stage_resample(P, 4, M1))
    return run


MISTAKES = ("E0 in E", "point scatter reads S", "point scatter reads the new dE", "surface radiance reads the new dS",
            "S accumulates dJ")


def _close(a, b, rel, scale=None):
    """max |a - b| within rel of the larger table's magnitude (per channel)."""
    scale = np.maximum(np.max(np.abs(a), axis=0), np.max(np.abs(b), axis=0)) if scale is None else scale
    return bool(np.all(np.abs(a - b) <= rel * np.maximum(scale, 1e-300)))


def _runs(oracle, driver):
    """{(t, K): (T, E, S, M)} for every t and iteration count, and the re-sampling of each space."""
    out = {}
    for t in T_VALUES:
        for k in range(MAX_K + 1):
            P = _params(oracle, t, ALBEDO * t, k)
            out[t, k] = driver(oracle, P)
    return out


def _tab(oracle, space, table):
    return oracle.stage_resample(_params(oracle, 1.0, ALBEDO, 0), space, table)


def structure_violations(oracle, driver):
    """Names of the checks 1-4 of the module docstring that `driver` fails (empty: all hold)."""
    bad = []
    runs = _runs(oracle, driver)
    P1 = _params(oracle, 1.0, ALBEDO, 1)
    R1 = oracle.stage_single_scatter(P1, 0)
    M1 = oracle.stage_single_scatter(P1, 1)
    E0 = oracle.stage_surface_radiance_base(P1)
    # 1. zero iterations
    for t in T_VALUES:
        T, E, S, M = runs[t, 0]
        if not (_close(S, _tab(oracle, 4, R1) * t, 1e-14) and not E.any()):
            bad.append("first order")
            break
    # 2. the first iteration's increments are the pinned stage compositions
    dS1 = oracle.stage_ray_scatter(P1, oracle.stage_point_scatter(P1, 1, R1, M1, E0))
    dE1 = oracle.stage_surface_radiance(P1, 1, R1, M1)
    S0, S1 = runs[1.0, 0][2], runs[1.0, 1][2]
    E1 = runs[1.0, 1][1]
    if not _close(S1 - _tab(oracle, 4, S0), _tab(oracle, 4, _tab(oracle, 4, dS1)), 1e-12, np.max(np.abs(S1), axis=0)):
        bad.append("second order")
    if not _close(E1, _tab(oracle, 3, _tab(oracle, 3, dE1)), 1e-12):
        bad.append("first surface radiance")
    # 3. each further iteration adds one monomial in t
    for k in range(MAX_K):
        for i, space, deg in ((2, 4, k + 2), (1, 3, k + 1)):
            inc = {t: runs[t, k + 1][i] - _tab(oracle, space, runs[t, k][i]) for t in T_VALUES}
            ref = inc[1.0]
            for t in T_VALUES:
                # rounding of the accumulated sums: relative to the whole table at t, scaled like the increment
                scale = np.max(np.abs(runs[t, k + 1][i]), axis=0) / t ** deg
                if not _close(inc[t] / t ** deg, ref, 1e-11, scale):
                    bad.append(f"{'S' if space == 4 else 'E'} increment {k + 1} not of degree {deg}")
                    break
    # 4. polynomial degree in the albedo alone
    rhos = np.arange(6) * 0.25
    for k in range(1, MAX_K + 1):
        tabs = [driver(oracle, _params(oracle, 1.0, r, k)) for r in rhos]
        for i, name, deg in ((2, "S", math.ceil(k / 2)), (1, "E", math.ceil((k - 1) / 2))):
            vals = np.stack([tb[i] for tb in tabs])
            scale = np.max(np.abs(vals), axis=(0, 1))
            top = np.diff(vals, n=deg, axis=0)
            over = np.diff(vals, n=deg + 1, axis=0)
            if np.max(np.abs(over) / scale) > 1e-12 or np.max(np.abs(top) / scale) < 1e-8:
                bad.append(f"{name}_{k} not of degree {deg} in the albedo")
    return bad


@pytest.mark.slow
def test_oracle_driver_has_the_structure_of_the_truncated_series(oracle):
    assert structure_violations(oracle, _oracle_driver) == []


def test_reference_wiring_rebuilt_from_stages_equals_the_oracle_driver(oracle):
    for k in (0, 1, 3):
        P = _params(oracle, 1.0, ALBEDO, k)
        for a, b in zip(_oracle_driver(oracle, P), _mis_wired(None)(oracle, P)):
            assert np.array_equal(a, b)


@pytest.mark.slow
@pytest.mark.parametrize("mistake", MISTAKES)
def test_each_wiring_mistake_breaks_the_structure(oracle, mistake):
    assert structure_violations(oracle, _mis_wired(mistake)) != []
