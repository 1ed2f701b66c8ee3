"""The single-scatter stage's two ways of getting each sample's sun optical depth: the tabulated node
sums (default, 640 x 768, read with bilinear interpolation in log2) and the node-by-node sums of
every sun segment (``sun_table = (0, 0)``, the reference's own procedure, atmosphere.clj:154-160).

Tolerances: the table path against the FP64 oracle at the scatter tolerance 1e-3 (measured 5.8e-5 max
on the full 5-order workload); the per-segment path at 1e-4 (measured 7e-6); the two paths against each
other at 2e-4; a 64 x 64 table must be measurably worse than the default (the error comes from the
table and shrinks with it)."""
import numpy as np
import pytest

from tests.parity import TOL_SCATTER, assert_close, oracle_params, rel_errors

DEFAULT = (640, 768)  # the library's default sun table (rows, cols)

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


def _grazing(shape):
    NH, NE, NL, NA = shape
    elev = sorted({0, 1, NE // 2 - 1, NE // 2, NE - 2, NE - 1})
    ih, ie, k = np.meshgrid([0, 1, NH // 4, NH - 1], elev, np.arange(NL * NA), indexing="ij")
    return np.sort(((ih * NE + ie) * NL * NA + k).reshape(-1))


@pytest.fixture(scope="module")
def full_tables(atm, gpu):
    """R1, M1 of the full 5-order workload with the default table, the per-segment sums and a 64 x 64
    table; the context is left at its default afterwards."""
    P = atm.workload("full-5-order")
    out = {}
    for shape in (DEFAULT, (0, 0), (64, 64)):
        gpu.sun_table = shape
        out[shape] = gpu.single_scatter_tables(P)
    gpu.sun_table = DEFAULT
    return P, out


def test_default_and_settings(atm, gpu):
    assert gpu.sun_table == DEFAULT
    gpu.sun_table = (0, 0)
    assert gpu.sun_table == (0, 0)
    for bad in ((8, 1024), (1024, 9000), (0, 64), (64, 0), (-1, -1)):
        with pytest.raises(atm.AtmoError) as e:
            gpu.sun_table = bad
        assert e.value.code == atm.ERR_ARGUMENT
        assert gpu.sun_table == (0, 0)  # a refused setting leaves the last one
    gpu.sun_table = DEFAULT


def test_both_paths_against_oracle(oracle, full_tables):
    P, tabs = full_tables
    OP = oracle_params(oracle, P)
    rng = np.random.default_rng(7)
    idx = np.sort(np.concatenate([rng.choice(int(np.prod(P.scatter_shape)), 384, replace=False),
                                  _grazing(tuple(P.scatter_shape))[::7]]))
    refR, refM = oracle.stage_single_scatter(OP, 0, idx), oracle.stage_single_scatter(OP, 1, idx)
    for shape, tol in ((DEFAULT, TOL_SCATTER), ((0, 0), 1e-4)):
        R1, M1 = tabs[shape]
        r = assert_close(f"R1 {shape}", R1.reshape(-1, 3)[idx], refR, tol)
        m = assert_close(f"M1 {shape}", M1.reshape(-1, 3)[idx], refM, tol)
        print(shape, "R1 max %.2e, M1 max %.2e" % (r["max_rel"], m["max_rel"]))


def test_paths_agree_on_every_entry(full_tables):
    P, tabs = full_tables
    worst = {}
    for shape in (DEFAULT, (64, 64)):
        worst[shape] = max(float(rel_errors(tabs[shape][c], tabs[(0, 0)][c]).max()) for c in (0, 1))
    print("table vs per-segment sums, max rel over all entries:", worst)
    assert worst[DEFAULT] <= 2e-4, worst
    assert worst[(64, 64)] > 4 * worst[DEFAULT], worst


def test_table_path_in_full_build(atm, gpu):
    """The whole 5-order build with the table and with per-segment sums: the final tables agree to the
    scatter tolerance (the later orders inherit the first order's difference, not amplify it)."""
    P = atm.workload("full-5-order")
    builds = {}
    for shape in (DEFAULT, (0, 0)):
        gpu.sun_table = shape
        builds[shape] = gpu.generate_atmosphere_luts(P)
    gpu.sun_table = DEFAULT
    for name in ("surface-radiance", "ray-scatter", "mie-strength"):
        e = float(rel_errors(builds[DEFAULT][name], builds[(0, 0)][name]).max())
        print(name, "%.2e" % e)
        assert e <= 2e-4, (name, e)
