"""Pins the FP64 oracle against the known answers held in the reference's own tests.

Every case of tests/golden/reference_goldens.json names the reference test file and line it was
transcribed from.  The table-level cases (size-12 transmittance / first-order scatter tables probed
through the renderer's shader functions, and the surface-radiance integral of an analytic scatter
field) rebuild the tables exactly as the reference tests do (t_atmosphere.clj:569-578,
t_planet.clj:312-320) and read them through the oracle's restatement of the GLSL consumer.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
CASES = GOLDEN["cases"]


def _value(expr):
    if isinstance(expr, str):
        return eval(expr, {"e": math.e, "pi": math.pi, "exp": math.exp})
    return expr


def _planet(oracle, name):
    spec = GOLDEN["planets"][name]
    return oracle.planet(spec["radius"], spec["height"], centre=spec.get("centre", (0, 0, 0)))


def _component(oracle, name):
    return {"RAYLEIGH": oracle.rayleigh(), "MIE": oracle.mie()}[name]


_TABLES = {}


def _size12_tables(oracle):
    """t_atmosphere.clj:43-67, 569-578: size 12, 10 ray steps, 100 km, scatter [mie rayleigh]."""
    if "T" not in _TABLES:
        earth = oracle.planet(6378000.0, 100000.0)
        P = oracle.params(earth, oracle.mie(), oracle.rayleigh(), (12, 12), (12, 12), (12, 12, 12, 12), 10, 8, 0)
        _TABLES["earth"] = earth
        _TABLES["T"] = oracle.stage_transmittance(P)
        # S: ray-scatter of the Rayleigh point-scatter-component; the test's M is built with the same
        # Rayleigh J (t_atmosphere.clj:573-574), so M == S.
        _TABLES["S"] = oracle.stage_single_scatter(P, 0)
        _TABLES["E"] = oracle.surface_radiance_analytic_table(earth, (12, 12), 10)
    return _TABLES


def _linear_index(x, minima, maxima, shape):
    """linear-forward of interpolate.clj:25-33."""
    return [(xi - a) / (b - a) * (n - 1) for xi, a, b, n in zip(x, minima, maxima, shape)]


def _ids(case):
    return f"{case['fn']}@{case['src'].split('/')[-1]}"


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_reference_golden(oracle, case):
    fn = case["fn"]
    tol = case.get("tol", 0.0)
    if fn in ("scattering", "extinction"):
        comp = oracle.scatter(case["base"], case["scale"], case.get("quotient", 1.0))
        got = getattr(oracle, fn)(comp, case["height"])[case["channel"]]
        assert abs(got - _value(case["expected"])) <= tol
    elif fn == "phase":
        assert abs(oracle.phase(case["g"], case["mu"]) - _value(case["expected"])) <= tol
    elif fn in ("atmosphere_intersection", "surface_intersection", "ray_extremity"):
        got = getattr(oracle, fn)(_planet(oracle, case["planet"]), case["origin"], case["direction"])
        assert np.linalg.norm(got - np.array(case["expected"], float)) <= tol
    elif fn == "surface_point":
        assert oracle.surface_point(_planet(oracle, case["planet"]), case["point"]) == case["expected"]
    elif fn == "is_above_horizon":
        got = oracle.is_above_horizon(_planet(oracle, case["planet"]), case["point"], case["direction"])
        assert got == case["expected"]
    elif fn == "horizon_distance":
        assert abs(oracle.horizon_distance(_planet(oracle, case["planet"]), case["radius"]) - case["expected"]) <= tol
    elif fn == "transmittance_points":
        comps = [_component(oracle, c) for c in case["components"]]
        got = oracle.transmittance(_planet(oracle, case["planet"]), comps, case["steps"], case["x"], case["x0"])
        assert abs(got[case["channel"]] - _value(case["expected"])) <= tol
    elif fn == "transmittance_dir":
        comps = [_component(oracle, c) for c in case["components"]]
        got = oracle.transmittance(_planet(oracle, case["planet"]), comps, case["steps"], case["x"], case["v"],
                                   case["above"])
        assert abs(got[case["channel"]] - _value(case["expected"])) <= tol
    elif fn == "elevation_to_index":
        got = oracle.elevation_to_index(_planet(oracle, case["planet"]), case["size"], case["point"],
                                        case["direction"], case["above"])
        assert abs(got - case["expected"]) <= tol
    elif fn == "index_to_elevation":
        d, above = oracle.index_to_elevation(_planet(oracle, case["planet"]), case["size"], case["radius"],
                                             case["index"])
        if "expected_direction" in case:
            assert np.linalg.norm(d - np.array(case["expected_direction"], float)) <= tol
        if "expected_above" in case:
            assert above == case["expected_above"]
    elif fn == "height_to_index":
        got = oracle.height_to_index(_planet(oracle, case["planet"]), case["size"], case["point"])
        assert abs(got - case["expected"]) <= tol
    elif fn == "index_to_height":
        got = oracle.index_to_height(_planet(oracle, case["planet"]), case["size"], case["index"])
        assert np.linalg.norm(got - np.array(case["expected"], float)) <= tol
    elif fn == "sun_elevation_to_index":
        assert abs(oracle.sun_elevation_to_index(case["size"], case["point"], case["light"]) - case["expected"]) <= tol
    elif fn == "index_to_sin_sun_elevation":
        assert abs(oracle.index_to_sin_sun_elevation(case["size"], case["index"]) - case["expected"]) <= tol
    elif fn == "sun_angle_to_index":
        assert abs(oracle.sun_angle_to_index(case["size"], case["direction"], case["light"]) - case["expected"]) <= tol
    elif fn == "index_to_sun_direction":
        got = oracle.index_to_sun_direction(case["size"], case["direction"], case["sin_sun"], case["index"])
        assert np.linalg.norm(got - np.array(case["expected"], float)) <= tol
    elif fn == "forward":
        got = oracle.space_forward(_planet(oracle, case["planet"]), case["space"], case["shape"], case["point"],
                                   case.get("direction", (0, 0, 0)), case.get("light", (0, 0, 0)),
                                   case.get("above", False))
        assert np.linalg.norm(got - np.array(case["expected"], float)) <= tol
    elif fn == "backward":
        point, d, l, above = oracle.space_backward(_planet(oracle, case["planet"]), case["space"], case["shape"],
                                                   case["index"])
        if "expected_point" in case:
            assert np.linalg.norm(point - np.array(case["expected_point"], float)) <= tol
        if "expected_direction" in case:
            assert np.linalg.norm(d - np.array(case["expected_direction"], float)) <= tol
        if "expected_light" in case:
            assert np.linalg.norm(l - np.array(case["expected_light"], float)) <= tol
        if "expected_above" in case:
            assert above == case["expected_above"]
    elif fn == "height":
        assert abs(oracle.height(_planet(oracle, case["planet"]), case["point"]) - case["expected"]) <= tol
    elif fn == "ray_sphere_intersection":
        got = oracle.ray_sphere_intersection(case["centre"], case["radius"], case["origin"], case["direction"])
        assert np.linalg.norm(np.array(got) - np.array(case["expected"], float)) <= tol
    elif fn == "integral_sphere":
        got = oracle.integral_sphere_test(case["half"], case["steps"], case["normal"], case["kind"])
        exp = np.array(case["expected"], float)
        assert np.linalg.norm(got[:len(exp)] - exp) <= tol
    elif fn == "integral_ray":
        got = oracle.integral_ray_test(case["origin"], case["direction"], case["steps"], case["distance"], case["kind"])
        assert np.linalg.norm(got - np.array(case["expected"], float)) <= tol
    elif fn == "limit_quot":
        assert abs(oracle.limit_quot(case["a"], case["b"], case["limit"]) - case["expected"]) <= tol
    elif fn == "interpolate_linear":
        table = np.array(case["table"], float)
        ndim = len(case["minima"])
        if case.get("vector"):
            data, shape = table, list(table.shape[:ndim])
        else:
            shape = list(table.shape)
            data = np.repeat(table[..., None], 3, axis=-1)
        idx = _linear_index(case["x"], case["minima"], case["maxima"], shape)
        got = oracle.interpolate(data.reshape(-1, 3), shape, idx)
        exp = np.array(case["expected"], float)
        assert np.max(np.abs((got if exp.ndim else got[0]) - exp)) <= tol
    elif fn == "convert_4d_to_2d":
        table = np.array(case["table"], float)
        got = oracle.convert_4d_to_2d(np.repeat(table[..., None], 3, axis=-1))[..., 0]
        assert np.array_equal(got, np.array(case["expected"], float))
    elif fn in ("interpolate_4d_normalized", "interpolate_2d_normalized"):
        # GLSL index order is reversed relative to the Clojure table order and normalised to [0,1]
        table = np.array(case["table"], float)
        shape = list(table.shape)
        idx = [c * (n - 1) for c, n in zip(reversed(case["glsl_idx"]), shape)]
        rgb = np.repeat(table[..., None], 3, axis=-1)
        got = oracle.interpolate(rgb.reshape(-1, 3), shape, idx)
        assert abs(got[0] - case["expected"]) <= tol
        # the same value through the shader's texture path over the file layout (4D tables tiled)
        if len(shape) == 4:
            tex = oracle.convert_4d_to_2d(rgb)
            got = oracle.glsl_interpolate_4d(tex.reshape(-1, 3), *shape, case["glsl_idx"])
        else:
            got = oracle.glsl_interpolate_2d(rgb.reshape(-1, 3), shape[0], shape[1], case["glsl_idx"])
        assert abs(got[0] - case["expected"]) <= tol
    elif fn == "glsl_convert_2d_index":
        got = oracle.glsl_convert_2d_index(case["idx"], case["size_y"], case["size_x"])
        assert np.max(np.abs(got - np.array(case["expected"]))) <= tol
    elif fn == "glsl_make_2d_index_from_4d":
        got = oracle.glsl_make_2d_index_from_4d(case["idx"], *case["sizes_wzyx"])
        sel = got[:2] if case["selector"] == "st" else got[2:]
        assert np.max(np.abs(sel - np.array(case["expected"]))) <= tol
    elif fn.startswith("table_"):
        t = _size12_tables(oracle)
        earth = t["earth"]
        if fn == "table_transmittance_track":
            got = oracle.glsl_transmittance_track(earth, (12, 12), t["T"], case["p"], case["q"])
        elif fn == "table_transmittance_outer":
            got = oracle.glsl_transmittance_outer(earth, (12, 12), t["T"], case["p"], case["d"])
            # and through the texture path over the file layout
            tiled = oracle.glsl_tiled_transmittance_outer(earth, (12, 12), t["T"], case["p"], case["d"])
            assert abs(tiled[case["channel"]] - case["expected"]) <= tol
        elif fn == "table_ray_scatter_track":
            got = oracle.glsl_ray_scatter_track(earth, (12, 12, 12, 12), t["S"], t["S"], (12, 12), t["T"], 0.76,
                                                case["light"], case["p"], case["q"])
        else:
            got = oracle.glsl_surface_radiance(earth, (12, 12), t["E"], case["p"], case["light"])
            tiled = oracle.glsl_tiled_surface_radiance(earth, (12, 12), t["E"], case["p"], case["light"])
            assert abs(tiled[case["channel"]] - case["expected"]) <= tol
        assert abs(got[case["channel"]] - case["expected"]) <= tol
    else:
        pytest.fail(f"unknown golden function {fn}")


def test_scattering_components_relation(oracle):
    """t_atmosphere.clj:285-296 (unmocked): the component sum equals point-scatter-base, and the Mie
    strength times the Mie phase equals the Mie point-scatter-component."""
    earth = oracle.planet(6378000.0, 100000.0)
    mie, ray = oracle.mie(), oracle.rayleigh()
    comps = [mie, ray]
    x, v, l = (6379000.0, 0, 0), (0, 1, 0), (0.36, 0.48, 0.8)
    a = oracle.single_scatter_J(earth, comps, mie, 100, (1, 1, 1), 0, x, v, l)
    b = oracle.single_scatter_J(earth, comps, ray, 100, (1, 1, 1), 0, x, v, l)
    base = oracle.point_scatter_base(earth, comps, 100, (1, 1, 1), x, v, l)
    assert np.linalg.norm(a + b - base) <= 1e-12
    strength = oracle.single_scatter_J(earth, comps, mie, 100, (1, 1, 1), 1, x, v, l)
    assert np.linalg.norm(strength * oracle.phase(0.76, np.dot(v, l)) - a) <= 1e-12


def test_sphere_direction_counts(oracle):
    """Ring-quadrature sizes of sphere.clj:70-105 for the step counts the tables use."""
    assert oracle.sphere_direction_count(False, 14) == 66
    assert oracle.sphere_direction_count(False, 15) == 71
    assert oracle.sphere_direction_count(False, 28) == 256
    assert oracle.sphere_direction_count(True, 64) == 660
    assert oracle.sphere_direction_count(True, 100) == 1605


def test_quadrature_frame(oracle):
    """quaternion.clj:166-171 + matrix.clj:207-213: for n = x the frame is (x, z, -y)."""
    m = oracle.oriented_matrix((1.0, 0.0, 0.0))
    assert np.allclose(m, [[1, 0, 0], [0, 0, 1], [0, -1, 0]])
    m = oracle.oriented_matrix((0.0, 1.0, 0.0))
    assert np.allclose(m[1], [0, 0, -1])  # n x e_x = (0,1,0) x (1,0,0) = (0,0,-1)
