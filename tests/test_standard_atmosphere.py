"""The standard atmosphere of sfsim's aerodynamics (atmosphere.clj:425-511, after Hull, "Fundamentals of
Airplane Flight Mechanics") in libsfsim_atmo (host code, no GPU), against the reference's own known answers
t_atmosphere.clj:1011-1042 with their tolerances, in the reference's units (units.clj:14-19: degrees
Rankine, feet, pounds-force, slugs)."""
import math

import pytest

from sfsim_amd import atmosphere as atm

RANKINE, FOOT = 5.0 / 9.0, 0.3048
POUND_FORCE = 0.45359237 * 9.8066500286389
SLUG = 14.5939029372

# (height in feet, expected, tolerance), t_atmosphere.clj:1011-1018 (temperature in degrees Rankine)
TEMPERATURE = [(0, 518.69, 1e-2), (10000, 483.03, 1e-2), (36089, 389.99, 1e-2), (50000, 389.99, 1e-2),
               (65617, 389.99, 1e-2), (80000, 397.88, 1e-2), (104990, 411.59, 1e-2)]
# t_atmosphere.clj:1021-1027 (pressure in lbf / ft^2; tolerance in Pa)
PRESSURE = [(0, 2116.1, 1e1), (36089, 472.68, 1e0), (50000, 242.2, 1e0), (65617, 114.35, 1e0), (80000, 57.7, 1e1),
            (104990, 18.12, 1e0)]
# t_atmosphere.clj:1030-1036 (density in slugs / ft^3; tolerance in kg / m^3)
DENSITY = [(0, 2.3769e-3, 1e-3), (36089, 7.0613e-4, 1e-4), (50000, 3.6184e-4, 1e-4), (65617, 1.7083e-4, 1e-5),
           (80000, 8.4459e-5, 1e-5), (104990, 2.5660e-5, 1e-5)]
# t_atmosphere.clj:1039-1042 (temperature K, speed m/s)
SOUND = [(273.15, 331.3), (293.15, 343.2), (223.15, 299.4)]


@pytest.mark.parametrize("feet,expected,tol", TEMPERATURE)
def test_temperature(feet, expected, tol):
    assert abs(atm.standard_atmosphere(feet * FOOT)[0] - expected * RANKINE) <= tol


@pytest.mark.parametrize("feet,expected,tol", PRESSURE)
def test_pressure(feet, expected, tol):
    assert abs(atm.standard_atmosphere(feet * FOOT)[1] - expected * POUND_FORCE / FOOT ** 2) <= tol


@pytest.mark.parametrize("feet,expected,tol", DENSITY)
def test_density(feet, expected, tol):
    assert abs(atm.standard_atmosphere(feet * FOOT)[2] - expected * SLUG / FOOT ** 3) <= tol


@pytest.mark.parametrize("temperature,expected", SOUND)
def test_speed_of_sound(temperature, expected):
    assert abs(atm.speed_of_sound(temperature) - expected) <= 1e-1


def test_layers_meet_and_state_equation_holds():
    """The three layers meet at 36089 ft and 65617 ft (temperature, pressure and density to the model's 4-5
    digit coefficients), and p = rho R T with the specific gas constant of dry air (287.05 J/kg/K)
    to the same accuracy, a check that a mis-typed coefficient or exponent fails."""
    for ft in (36089.0, 65617.0):
        lo, hi = atm.standard_atmosphere(ft * FOOT), atm.standard_atmosphere(ft * FOOT * (1 + 1e-12))
        assert lo[0] == pytest.approx(hi[0], rel=1e-5)
        assert lo[1] == pytest.approx(hi[1], rel=2e-3)
        assert lo[2] == pytest.approx(hi[2], rel=2e-3)
    for ft in (0, 5000, 20000, 36089, 45000, 65617, 75000, 100000, 150000):
        t, p, d = atm.standard_atmosphere(ft * FOOT)
        assert p / (d * t) == pytest.approx(287.05, rel=2e-3), ft


def test_nonfinite_height_is_an_error():
    with pytest.raises(atm.AtmoError) as e:
        atm.standard_atmosphere(math.nan)
    assert e.value.code == atm.ERR_NONFINITE
