"""GPU parity of libsfsim_atmo against the FP64 oracle (and, through it, the reference).

* reference table probes: the size-12 transmittance and first-order scatter tables and the analytic
  surface-radiance table of the reference's tests, built on the GPU and read through the oracle's
  restatement of the renderer's shader functions, against the reference's own expected values;
* every stage at small sizes, fed identical inputs, against the oracle's stage (whole tables);
* the whole driver (3 orders) at small sizes against the oracle's driver;
* every stage at the full 5-order workload (T 64x256, E 16x64, S 32x64x32x8, 64 steps), checked on
  seeded random entry subsets with the GPU's own previous-stage tables as inputs;
* error paths that the host must turn into error codes.

Tolerances: transmittance 1e-4, everything else 1e-3 relative (floor: 1e-6 of the table's channel
maximum, see tests/parity.py).
"""
import json
import os

import numpy as np
import pytest

from tests.parity import TOL_SCATTER, TOL_TRANSMITTANCE, assert_close, fixture_errors, oracle_params, rel_errors

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def atm():
    from sfsim_amd import atmosphere
    return atmosphere


@pytest.fixture(scope="module")
def gpu(atm):
    ctx = atm.Atmosphere(0)
    yield ctx
    ctx.close()


def small_params(atm, iterations=3, height=100000.0, brightness=0.3):
    return atm.make_params(planet={"height": height, "brightness": (brightness,) * 3}, transmittance_shape=(10, 17),
                           surface_radiance_shape=(6, 11), scatter_shape=(6, 11, 7, 4), ray_steps=16, sphere_steps=10,
                           iterations=iterations)


# ------------------------------------------------------------------ reference table probes
def test_reference_size12_table_probes(atm, gpu, oracle):
    """t_atmosphere.clj:569-578, 629-635, 661-667, 769-775 with the tables built on the GPU."""
    golden = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))
    P = atm.make_params(planet={"height": 100000.0}, transmittance_shape=(12, 12), surface_radiance_shape=(12, 12),
                        scatter_shape=(12, 12, 12, 12), ray_steps=10, sphere_steps=8, iterations=0)
    T = gpu.transmittance_table(P).astype(np.float64).reshape(-1, 3)
    R1, _ = gpu.single_scatter_tables(P)
    S = R1.astype(np.float64).reshape(-1, 3)
    earth = oracle.planet(6378000.0, 100000.0)
    checked = 0
    for case in golden["cases"]:
        if case["fn"] == "table_transmittance_track":
            got = oracle.glsl_transmittance_track(earth, (12, 12), T, case["p"], case["q"])
        elif case["fn"] == "table_transmittance_outer":
            got = oracle.glsl_transmittance_outer(earth, (12, 12), T, case["p"], case["d"])
        elif case["fn"] == "table_ray_scatter_track":
            got = oracle.glsl_ray_scatter_track(earth, (12, 12, 12, 12), S, S, (12, 12), T, 0.76, case["light"],
                                                case["p"], case["q"])
        else:
            continue
        assert abs(got[case["channel"]] - case["expected"]) <= case["tol"], case["src"]
        checked += 1
    assert checked == 9


# ------------------------------------------------------------------ stages at small sizes
def test_stage_transmittance_small(atm, gpu, oracle):
    P = small_params(atm)
    assert_close("T", gpu.transmittance_table(P), oracle.stage_transmittance(oracle_params(oracle, P)),
                 TOL_TRANSMITTANCE)


def test_stage_surface_radiance_base_small(atm, gpu, oracle):
    P = small_params(atm)
    assert_close("E0", gpu.surface_radiance_base_table(P),
                 oracle.stage_surface_radiance_base(oracle_params(oracle, P)), TOL_TRANSMITTANCE)


def test_stage_single_scatter_small(atm, gpu, oracle):
    P = small_params(atm)
    OP = oracle_params(oracle, P)
    R1, M1 = gpu.single_scatter_tables(P)
    assert_close("R1", R1, oracle.stage_single_scatter(OP, 0), TOL_SCATTER)
    assert_close("M1", M1, oracle.stage_single_scatter(OP, 1), TOL_SCATTER)


@pytest.fixture(scope="module")
def small_inputs(atm, oracle):
    P = small_params(atm)
    OP = oracle_params(oracle, P)
    R1 = oracle.stage_single_scatter(OP, 0)
    M1 = oracle.stage_single_scatter(OP, 1)
    E0 = oracle.stage_surface_radiance_base(OP)
    J1 = oracle.stage_point_scatter(OP, 1, R1, M1, E0)
    return P, OP, R1, M1, E0, J1


def test_stage_point_scatter_small(gpu, oracle, small_inputs):
    P, OP, R1, M1, E0, J1 = small_inputs
    # first iteration: R1 + M1 * phase(mie)
    assert_close("dJ1", gpu.point_scatter_table(P, R1, E0, mie_strength=M1), J1, TOL_SCATTER)
    # later iterations: a plain dS table
    dS = oracle.stage_ray_scatter(OP, J1)
    assert_close("dJ2", gpu.point_scatter_table(P, dS, E0), oracle.stage_point_scatter(OP, 0, dS, None, E0),
                 TOL_SCATTER)


def test_stage_ray_scatter_small(gpu, oracle, small_inputs):
    P, OP, R1, M1, E0, J1 = small_inputs
    assert_close("dS", gpu.ray_scatter_table(P, J1), oracle.stage_ray_scatter(OP, J1), TOL_SCATTER)


def test_stage_calls_between_builds_use_their_own_maps(atm, gpu, oracle, small_inputs):
    """Every build and every single-stage call writes the per-build re-sampling and sun maps for its own
    grid before the kernels that read them: stage calls on the small grid between two builds of the full
    workload match the oracle, and the second build (a graph replay) equals the first bit for bit."""
    P, OP, R1, M1, E0, J1 = small_inputs
    F = atm.workload("full-5-order")
    gpu.build(F)
    first = [gpu.fetch(F, i) for i in range(4)]
    assert_close("dJ1", gpu.point_scatter_table(P, R1, E0, mie_strength=M1), J1, TOL_SCATTER)
    assert_close("dS", gpu.ray_scatter_table(P, J1), oracle.stage_ray_scatter(OP, J1), TOL_SCATTER)
    R1g, M1g = gpu.single_scatter_tables(P)
    assert_close("R1", R1g, R1, TOL_SCATTER)
    assert_close("M1", M1g, M1, TOL_SCATTER)
    gpu.build(F)
    for i in range(4):
        assert np.array_equal(gpu.fetch(F, i), first[i]), i


def test_stage_surface_radiance_small(gpu, oracle, small_inputs):
    P, OP, R1, M1, E0, J1 = small_inputs
    assert_close("dE1", gpu.surface_radiance_table(P, R1, mie_strength=M1),
                 oracle.stage_surface_radiance(OP, 1, R1, M1), TOL_SCATTER)
    dS = oracle.stage_ray_scatter(OP, J1)
    assert_close("dE2", gpu.surface_radiance_table(P, dS), oracle.stage_surface_radiance(OP, 0, dS, None),
                 TOL_SCATTER)


@pytest.mark.parametrize("space", [2, 3, 4])
def test_stage_resample_small(atm, gpu, oracle, space):
    P = small_params(atm)
    OP = oracle_params(oracle, P)
    shape = {2: P.transmittance_shape, 3: P.surface_radiance_shape, 4: P.scatter_shape}[space]
    rng = np.random.default_rng(20261202 + space)
    a = rng.uniform(0.1, 1.0, (int(np.prod(shape)), 3))
    b = rng.uniform(0.1, 1.0, (int(np.prod(shape)), 3))
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    assert_close(f"resample{space}", gpu.resample_table(P, space, a32, b32),
                 oracle.stage_resample(OP, space, a32.astype(np.float64), b32.astype(np.float64)), 1e-5)


# ------------------------------------------------------------------ whole driver, small
@pytest.mark.parametrize("iterations,brightness", [(0, 0.3), (1, 0.3), (3, 0.3), (2, 0.0)])
def test_generate_luts_small(atm, gpu, oracle, iterations, brightness):
    P = small_params(atm, iterations=iterations, brightness=brightness)
    OP = oracle_params(oracle, P)
    T, E, S, M = oracle.generate_atmosphere_luts(OP)
    got = gpu.generate_atmosphere_luts(P)
    assert_close("transmittance", got["transmittance"], T, TOL_TRANSMITTANCE)
    if iterations:
        assert_close("surface-radiance", got["surface-radiance"], E, TOL_SCATTER)
    else:
        assert not got["surface-radiance"].any()
    assert_close("ray-scatter", got["ray-scatter"], oracle.convert_4d_to_2d(S), TOL_SCATTER)
    assert_close("mie-strength", got["mie-strength"], oracle.convert_4d_to_2d(M), TOL_SCATTER)


@pytest.mark.parametrize("scatter_shape", [(4, 6, 24, 16), (3, 5, 32, 24)])
def test_generate_luts_wide_slabs(atm, gpu, oracle, scatter_shape):
    """Sun-elevation x heading slabs of 384 and 768 entries: the ray-scatter kernel's instantiations with two
    and four entries per thread (one- and two-sample chunks), against the oracle's whole driver."""
    P = atm.make_params(planet={"height": 100000.0, "brightness": (0.3,) * 3}, transmittance_shape=(10, 17),
                        surface_radiance_shape=(6, 11), scatter_shape=scatter_shape, ray_steps=16, sphere_steps=10,
                        iterations=2)
    OP = oracle_params(oracle, P)
    T, E, S, M = oracle.generate_atmosphere_luts(OP)
    got = gpu.generate_atmosphere_luts(P)
    assert_close("transmittance", got["transmittance"], T, TOL_TRANSMITTANCE)
    assert_close("surface-radiance", got["surface-radiance"], E, TOL_SCATTER)
    assert_close("ray-scatter", got["ray-scatter"], oracle.convert_4d_to_2d(S), TOL_SCATTER)
    assert_close("mie-strength", got["mie-strength"], oracle.convert_4d_to_2d(M), TOL_SCATTER)
    gpu.build(P)  # the graph-replayed build returns the same tables
    for which, name in enumerate(TABLES):
        assert np.array_equal(gpu.fetch(P, which), got[name]), name


@pytest.mark.parametrize("scatter_shape,ray_steps", [((4, 6, 32, 8), 13), ((4, 6, 7, 5), 11)])
def test_generate_luts_odd_ray_steps(atm, gpu, oracle, scatter_shape, ray_steps):
    """Odd numbers of view samples for the one-entry-per-lane ray-scatter kernel, whose fold wavefronts split
    the samples by parity: 13 samples in chunks of four (the last chunk one even sample) over a 256-entry
    slab, and 11 samples in chunks of eight (a second group of steps inside a chunk, and lanes past the
    35-entry slab), against the oracle's whole driver; the graph-replayed build returns the same tables."""
    P = atm.make_params(planet={"height": 100000.0, "brightness": (0.3,) * 3}, transmittance_shape=(10, 17),
                        surface_radiance_shape=(6, 11), scatter_shape=scatter_shape, ray_steps=ray_steps,
                        sphere_steps=10, iterations=2)
    OP = oracle_params(oracle, P)
    T, E, S, M = oracle.generate_atmosphere_luts(OP)
    got = gpu.generate_atmosphere_luts(P)
    assert_close("transmittance", got["transmittance"], T, TOL_TRANSMITTANCE)
    assert_close("surface-radiance", got["surface-radiance"], E, TOL_SCATTER)
    assert_close("ray-scatter", got["ray-scatter"], oracle.convert_4d_to_2d(S), TOL_SCATTER)
    assert_close("mie-strength", got["mie-strength"], oracle.convert_4d_to_2d(M), TOL_SCATTER)
    gpu.build(P)
    for which, name in enumerate(TABLES):
        assert np.array_equal(gpu.fetch(P, which), got[name]), name


TABLES = ["transmittance", "surface-radiance", "ray-scatter", "mie-strength"]


def test_tiled_tables_read_through_the_renderer_path(atm, gpu, oracle):
    """The GPU's final tables in file layout, read the way the renderer reads them (the oracle's restatement
    of transmittance-outer.glsl, ray-scatter-outer.glsl and surface-radiance.glsl over the tiled textures),
    against the oracle's own tables read the same way, at random points, views and sun directions."""
    P = small_params(atm, iterations=2)
    OP = oracle_params(oracle, P)
    T, E, S, M = oracle.generate_atmosphere_luts(OP)
    got = gpu.generate_atmosphere_luts(P)
    shape = tuple(P.scatter_shape)
    St, Mt = oracle.convert_4d_to_2d(S).reshape(-1, 3), oracle.convert_4d_to_2d(M).reshape(-1, 3)
    gS, gM = got["ray-scatter"].reshape(-1, 3), got["mie-strength"].reshape(-1, 3)
    rng = np.random.default_rng(31)
    R, H = P.planet.radius, P.planet.height
    res = {"T": [], "S": [], "E": []}
    for _ in range(300):
        point = np.array([R + rng.uniform(0, H), 0, 0])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        l = rng.normal(size=3)
        l /= np.linalg.norm(l)
        if oracle.is_above_horizon(OP.planet, point, d):
            res["T"].append((oracle.glsl_tiled_transmittance_outer(OP.planet, P.transmittance_shape,
                                                                   got["transmittance"].reshape(-1, 3), point, d),
                             oracle.glsl_tiled_transmittance_outer(OP.planet, P.transmittance_shape, T.reshape(-1, 3),
                                                                   point, d)))
            res["S"].append((oracle.glsl_tiled_ray_scatter_outer(OP.planet, shape, gS, gM, 0.76, l, point, d),
                             oracle.glsl_tiled_ray_scatter_outer(OP.planet, shape, St, Mt, 0.76, l, point, d)))
        res["E"].append((oracle.glsl_tiled_surface_radiance(OP.planet, P.surface_radiance_shape,
                                                            got["surface-radiance"].reshape(-1, 3), point, l),
                         oracle.glsl_tiled_surface_radiance(OP.planet, P.surface_radiance_shape, E.reshape(-1, 3),
                                                            point, l)))
    for key, tol in (("T", TOL_TRANSMITTANCE), ("S", TOL_SCATTER), ("E", TOL_SCATTER)):
        g = np.array([a for a, _ in res[key]])
        o = np.array([b for _, b in res[key]])
        assert len(g) > 50
        assert_close(key, g, o, tol)


def test_build_fetch_matches_generate(atm, gpu):
    P = small_params(atm, iterations=2)
    ref = gpu.generate_atmosphere_luts(P)
    gpu.build_timed(P)
    for which, name in enumerate(TABLES):
        assert np.array_equal(gpu.fetch(P, which), ref[name]), name
    timings = gpu.stage_timings()
    assert timings["single_scatter"][1] == 1 and timings["point_scatter"][1] == 2


@pytest.mark.parametrize("iterations", [0, 2])
def test_graph_replays_equal_eager_builds(atm, gpu, iterations):
    """sfsim_atmo_build replays a captured graph: every replay, with eager work and other parameters in
    between, equals the eager kernel-by-kernel build bit for bit (iterations 0 covers the zero-filled
    surface-radiance table)."""
    P = small_params(atm, iterations=iterations)
    Q = small_params(atm, iterations=1, brightness=0.1)
    eager = gpu.generate_atmosphere_luts(P)
    other = gpu.generate_atmosphere_luts(Q)
    for rep in range(4):
        gpu.build(P)
        for which, name in enumerate(TABLES):
            assert np.array_equal(gpu.fetch(P, which), eager[name]), (rep, name)
        gpu.transmittance_table(Q)  # eager work between replays
        if rep == 2:  # a different parameter set: re-capture, then back
            gpu.build(Q)
            for which, name in enumerate(TABLES):
                assert np.array_equal(gpu.fetch(Q, which), other[name]), name


def test_determinism(atm, gpu):
    P = small_params(atm, iterations=2)
    a = gpu.generate_atmosphere_luts(P)
    b = gpu.generate_atmosphere_luts(P)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ------------------------------------------------------------------ full-size stages on subsets
@pytest.fixture(scope="module")
def full(atm, gpu):
    P = atm.workload("full-5-order")
    T = gpu.transmittance_table(P)
    E0 = gpu.surface_radiance_base_table(P)
    R1, M1 = gpu.single_scatter_tables(P)
    J1 = gpu.point_scatter_table(P, R1, E0, mie_strength=M1)
    dE1 = gpu.surface_radiance_table(P, R1, mie_strength=M1)
    dS1 = gpu.ray_scatter_table(P, J1)
    J2 = gpu.point_scatter_table(P, dS1, dE1)
    return P, T, E0, R1, M1, J1, dE1, dS1, J2


def _subset(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))


def test_full_transmittance(oracle, full):
    P, T = full[0], full[1]
    assert_close("T", T, oracle.stage_transmittance(oracle_params(oracle, P)), TOL_TRANSMITTANCE)


def test_full_surface_radiance_base(oracle, full):
    P, E0 = full[0], full[2]
    assert_close("E0", E0, oracle.stage_surface_radiance_base(oracle_params(oracle, P)), TOL_TRANSMITTANCE)


def test_full_single_scatter_subset(oracle, full):
    P, R1, M1 = full[0], full[3], full[4]
    OP = oracle_params(oracle, P)
    idx = _subset(R1.size // 3, 384, 1)
    assert_close("R1", R1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 0, idx), TOL_SCATTER)
    assert_close("M1", M1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 1, idx), TOL_SCATTER)


def _grazing_entries(shape, rows):
    """Entries of the 4D grid whose view ray grazes the ground sphere: elevation index 0 (the ray
    tangent to the ground at the horizon point), the two indices next to the ground/sky split and the
    last index (the longest ray to the top, whose extension touches the ground sphere); height index 0
    holds the horizontal rays (mu = 0 at h = 0).  All sun elevations and headings of the given rows."""
    NH, NE, NL, NA = shape
    elev = sorted({0, 1, NE // 2 - 1, NE // 2, NE - 2, NE - 1})
    ih, ie, k = np.meshgrid(rows, elev, np.arange(NL * NA), indexing="ij")
    return np.sort(((ih * NE + ie) * NL * NA + k).reshape(-1))


def test_full_grazing_rays(oracle, full):
    """Transmittance and single scatter on grazing and tangent view rays (the FP64 segment path)."""
    P, T, R1, M1 = full[0], full[1], full[3], full[4]
    OP = oracle_params(oracle, P)
    NT = tuple(P.transmittance_shape)
    tidx = np.array([h * NT[1] + e for h in range(NT[0]) for e in sorted({0, 1, NT[1] // 2 - 1, NT[1] // 2,
                                                                            NT[1] - 2, NT[1] - 1})])
    assert_close("T grazing", T.reshape(-1, 3)[tidx], oracle.stage_transmittance(OP, tidx), TOL_TRANSMITTANCE)
    idx = _grazing_entries(tuple(P.scatter_shape), [0, 1, 7, 31])
    assert_close("R1 grazing", R1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 0, idx), TOL_SCATTER)
    assert_close("M1 grazing", M1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 1, idx), TOL_SCATTER)


def test_production_grazing_rays(atm, gpu, oracle):
    """The reference's production grid (35 km, T 64x255, S 32x127x32x8, 100 steps): odd elevation sizes
    put an index exactly on the ground/sky split (a vertical ray by the reference's limit-quot)."""
    P = atm.workload("production")
    OP = oracle_params(oracle, P)
    assert_close("T production", gpu.transmittance_table(P), oracle.stage_transmittance(OP), TOL_TRANSMITTANCE)
    R1, M1 = gpu.single_scatter_tables(P)
    idx = _grazing_entries(tuple(P.scatter_shape), [0, 3, 31])
    assert_close("R1 production", R1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 0, idx), TOL_SCATTER)
    assert_close("M1 production", M1.reshape(-1, 3)[idx], oracle.stage_single_scatter(OP, 1, idx), TOL_SCATTER)


def test_full_point_scatter_subset(oracle, full):
    P, T, E0, R1, M1, J1, dE1, dS1, J2 = full
    OP = oracle_params(oracle, P)
    idx = _subset(R1.size // 3, 384, 2)
    f64 = lambda a: a.astype(np.float64).reshape(-1, 3)
    ref1 = oracle.stage_point_scatter(OP, 1, f64(R1), f64(M1), f64(E0), idx)
    assert_close("dJ1", J1.reshape(-1, 3)[idx], ref1, TOL_SCATTER)
    ref2 = oracle.stage_point_scatter(OP, 0, f64(dS1), None, f64(dE1), idx)
    assert_close("dJ2", J2.reshape(-1, 3)[idx], ref2, TOL_SCATTER)


def test_full_ray_scatter_subset(oracle, full):
    P, J1, dS1 = full[0], full[5], full[7]
    OP = oracle_params(oracle, P)
    idx = _subset(J1.size // 3, 384, 3)
    assert_close("dS1", dS1.reshape(-1, 3)[idx], oracle.stage_ray_scatter(OP, J1.astype(np.float64).reshape(-1, 3), idx),
                 TOL_SCATTER)


def test_full_surface_radiance(oracle, full):
    P, R1, M1, dE1 = full[0], full[3], full[4], full[6]
    OP = oracle_params(oracle, P)
    ref = oracle.stage_surface_radiance(OP, 1, R1.astype(np.float64).reshape(-1, 3), M1.astype(np.float64).reshape(-1, 3))
    assert_close("dE1", dE1, ref, TOL_SCATTER)


def test_full_build_is_finite_and_positive(atm, gpu):
    P = atm.workload("full-5-order")
    luts = gpu.generate_atmosphere_luts(P)
    for name, t in luts.items():
        assert np.all(np.isfinite(t)), name
        assert np.all(t >= 0), name
    assert np.all(luts["transmittance"] <= 1.0)


# ------------------------------------------------------------------ error paths
def test_invalid_params_are_errors(atm, gpu):
    P = small_params(atm)
    P.scatter_shape[2] = 1
    with pytest.raises(atm.AtmoError) as e:
        gpu.transmittance_table(P)
    assert e.value.code == atm.ERR_ARGUMENT
    P = small_params(atm)
    P.ray_steps = 0
    with pytest.raises(atm.AtmoError) as e:
        gpu.build(P)
    assert e.value.code == atm.ERR_ARGUMENT
    P = small_params(atm)
    P.mie.scatter_scale = float("nan")
    with pytest.raises(atm.AtmoError) as e:
        gpu.build(P)
    assert e.value.code == atm.ERR_NONFINITE


def test_nonfinite_input_table_is_an_error(atm, gpu):
    P = small_params(atm)
    n = int(np.prod(P.scatter_shape))
    dJ = np.ones((n, 3), np.float32)
    dJ[5, 1] = np.nan
    with pytest.raises(atm.AtmoError) as e:
        gpu.ray_scatter_table(P, dJ)
    assert e.value.code == atm.ERR_NONFINITE


def test_fetch_before_build_is_an_error(atm):
    ctx = atm.Atmosphere(0)
    P = small_params(atm)
    with pytest.raises(atm.AtmoError):
        ctx.fetch(P, 0)
    ctx.close()


# ------------------------------------------------------------------ full-size end to end
TOLS = {"transmittance": TOL_TRANSMITTANCE, "surface-radiance": TOL_SCATTER, "ray-scatter": TOL_SCATTER,
        "mie-strength": TOL_SCATTER}


def _fixture(name):
    path = os.path.join(HERE, "golden", f"oracle_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"oracle fixture {path} not generated")
    return np.load(path)


@pytest.mark.parametrize("workload", ["full-5-order", "production"])
def test_full_build_against_oracle_fixture(atm, gpu, workload):
    """The whole 5-order build at full size against the FP64 oracle's build of the same workload
    (tests/golden/make_oracle_subsets.py): every transmittance and surface-radiance entry, and of the 4D
    tables every view ray at five (sun elevation, heading) pairs, every horizon-row entry, a sample of the
    entries whose sun direction the reference clamps, and a uniform sample.  The benchmark's path (the
    captured graph replayed by sfsim_atmo_build) equals the eager kernel-by-kernel build bit for bit."""
    ref = _fixture(workload)
    P = atm.workload(workload)
    eager = gpu.generate_atmosphere_luts(P)
    gpu.build(P)
    gpu.build(P)  # the replay
    for which, name in enumerate(TABLES):
        assert np.array_equal(gpu.fetch(P, which), eager[name]), name
    errs = fixture_errors(ref, eager, tuple(P.scatter_shape))
    for name, r in errs.items():
        print(workload, name, r)
        assert r["max_rel"] <= TOLS[name], (name, r)
    assert errs["ray-scatter"]["entries"] >= 4 * P.scatter_shape[0] * P.scatter_shape[1]


def test_production_scatter_files(atm, gpu, tmp_path):
    """atmosphere_lut.clj:102-105 at the production shapes: the four files have the sizes the loader's
    constants imply (atmosphere.clj:572-579; little-endian FP32, no header) and hold, at the oracle
    fixture's entries, the oracle's values in the 4D->2D tiling the loader uploads (image.clj:299-312,
    texture.clj:380-398)."""
    ref = _fixture("production")
    P = atm.workload("production")
    luts = gpu.generate_atmosphere_luts(P)
    atm.write_scatter_files(str(tmp_path), luts)
    sizes = {"transmittance": 64 * 255, "surface-radiance": 16 * 63, "ray-scatter": 32 * 127 * 32 * 8,
             "mie-strength": 32 * 127 * 32 * 8}
    files = {}
    for name, n in sizes.items():
        raw = (tmp_path / f"{name}.scatter").read_bytes()
        assert len(raw) == n * 3 * 4, name
        files[name] = np.frombuffer(raw, dtype="<f4")
        assert np.array_equal(files[name], luts[name].reshape(-1)), name
    # the ray-scatter texture: width heading x elevation = 1016, height sun elevation x height = 1024
    assert luts["ray-scatter"].shape == (32 * 32, 127 * 8, 3)
    for name, r in fixture_errors(ref, files, (32, 127, 32, 8)).items():
        assert r["max_rel"] <= TOLS[name], (name, r)
