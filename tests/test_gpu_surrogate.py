"""GPU checks of the learned stand-in for scattering orders >= 2 (no counterpart in the reference).

* forward pass (bf16 matrix-core products, FP32 accumulation) against the FP64 reference network
  oracle/mlp.py, both exact and with the GPU's bf16 roundings emulated;
* loss and gradient (every parameter block; all three operand layouts of the matrix-core product:
  A B, A^T B and A B^T) against oracle/mlp.py;
* the order sums P3, P4, P5 that the network learns, against the oracle's own iteration;
* training: the loss falls, and two runs from one seed give identical weights;
* the seeded 5-order build: better than the same step without the network, falls back to the plain
  iteration (bit-identical result) when the residual exceeds the tolerance, and refuses an untrained
  network.
"""
import numpy as np
import pytest

from oracle import mlp
from tests.parity import TOL_SCATTER, assert_close, oracle_params, rel_errors

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


def small5(atm, **kw):
    base = dict(planet={"height": 100000.0}, transmittance_shape=(10, 17), surface_radiance_shape=(6, 11),
                scatter_shape=(6, 11, 7, 4), ray_steps=16, sphere_steps=10, iterations=5)
    base.update(kw)
    return atm.make_params(**base)


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def xavier_theta(rng):
    parts = []
    for s in mlp.SHAPES:
        if len(s) == 2:
            a = np.sqrt(6.0 / (s[0] + s[1]))
            parts.append(rng.uniform(-a, a, s))
        else:
            parts.append(rng.normal(0.0, 0.1, s))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def emulated_forward(theta, x):
    """The GPU's forward pass with its bf16 roundings (features, weights, activations), FP64 sums."""
    W1, b1, W2, b2, W3, b3, W4, b4 = mlp.unflatten(theta)
    a = bf16(mlp.encode(x).astype(np.float32))
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        a = bf16(mlp.silu(a @ bf16(W) + b).astype(np.float32))
    return a @ bf16(W4) + b4


@pytest.fixture(scope="module")
def net(atm, gpu):
    rng = np.random.default_rng(20261202)
    theta = xavier_theta(rng)
    sg = atm.Surrogate(gpu, seed=1, batch=256)
    sg.set_weights(theta, trained=False)
    x = rng.uniform(0, 1, (200, 10)).astype(np.float32)
    return sg, theta.astype(np.float64), x


def test_parameter_count(atm):
    assert atm.surrogate_size() == mlp.n_params()


def test_forward_matches_fp64_network(net):
    sg, theta, x = net
    y, _, _ = sg.eval(x)
    ref = mlp.forward(theta, mlp.encode(x))[0]
    emu = emulated_forward(theta, x)
    scale = np.max(np.abs(ref))
    assert scale > 0.1
    # bf16 operands: about 3 significant digits; a transposed operand or a wrong accumulator layout
    # gives errors of the order of the outputs themselves
    assert np.max(np.abs(y - emu)) <= 1e-2 * scale
    assert np.max(np.abs(y - ref)) <= 5e-2 * scale
    assert np.median(np.abs(y - emu)) <= 1e-3 * scale


def test_forward_single_row_and_weight_roundtrip(atm, net):
    sg, theta, x = net
    y_all, _, _ = sg.eval(x)
    y_one, _, _ = sg.eval(x[17:18])
    assert np.array_equal(y_one[0], y_all[17])
    assert np.array_equal(sg.weights(), theta.astype(np.float32))


def test_loss_and_gradient_match_fp64_network(net):
    sg, theta, x = net
    rng = np.random.default_rng(7)
    T = np.zeros((x.shape[0], 16), np.float32)
    T[:, :15] = rng.normal(0.0, 1.0, (x.shape[0], 15))
    y, g, loss = sg.eval(x, T)
    mask = np.zeros((x.shape[0], 16))
    mask[:, :15] = 1.0
    ref_loss, ref_g = mlp.loss_and_grad(theta, mlp.encode(x), T.astype(np.float64), mask)
    assert loss == pytest.approx(ref_loss, rel=2e-2)
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in mlp.SHAPES])
    names = ["W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"]
    for b, name in enumerate(names):
        gg, rr = g[offsets[b]:offsets[b + 1]].astype(np.float64), ref_g[offsets[b]:offsets[b + 1]]
        rel = np.linalg.norm(gg - rr) / np.linalg.norm(rr)
        cos = gg @ rr / (np.linalg.norm(gg) * np.linalg.norm(rr))
        print(f"{name}: relative error {rel:.2e}, cosine {cos:.6f}")
        assert rel <= 3e-2 and cos >= 0.999, name
    # output 15 is masked: its weight and bias gradients are exactly zero
    W4g = g[offsets[6]:offsets[7]].reshape(128, 16)
    assert not W4g[:, 15:].any() and not g[offsets[7] + 15:offsets[8]].any()


def atmosphere_code(P):
    """The six conditioning inputs of an atmosphere in [0, 1] (as the library maps them)."""
    def unit(v, lo, hi):
        return min(max((v - lo) / (hi - lo), 0.0), 1.0)
    return np.array([unit(P.rayleigh.scatter_scale, 6000.0, 10000.0),
                     unit(P.rayleigh.scatter_base[0] / 5.8e-6, 0.5, 2.0), unit(P.mie.scatter_scale, 800.0, 2000.0),
                     unit(P.mie.scatter_base[0], 5e-6, 5e-5), unit(P.mie.scatter_g, 0.6, 0.85),
                     unit(P.planet.brightness[0], 0.1, 0.5)], np.float32)


def entry_inputs(P, entries):
    shape = tuple(P.scatter_shape)
    coords = np.stack(np.unravel_index(entries, shape), axis=1).astype(np.float32)
    coords /= np.array([n - 1 for n in shape], np.float32)
    code = np.broadcast_to(atmosphere_code(P), (len(entries), 6))
    return np.ascontiguousarray(np.concatenate([coords, code], axis=1), np.float32)


def layerwise_prediction(sg, P, entries, batch):
    """P_1..P_5 of the given entries through the training path's products (sfsim_atmo_surrogate_eval):
    array [5, len(entries), 3]."""
    x = entry_inputs(P, entries)
    y = np.concatenate([sg.eval(x[i:i + batch])[0] for i in range(0, len(x), batch)])
    v = np.maximum(np.exp(y[:, :15].astype(np.float64) - 10.0) - 1e-8, 0.0)
    return v.reshape(-1, 5, 3).transpose(1, 0, 2)


def assert_same_prediction(fused, layerwise):
    """Both paths round to bf16 at the same places; only FP32 summation order differs, which can flip the
    bf16 rounding of an activation.  The log outputs (log(P + 1e-8) + 10) reach ~10, where one bf16 unit
    (2^-8 relative) is ~4e-2: a flip moves an output by at most about that, so the bound on the largest
    difference is 5e-2; almost everywhere the two paths agree (median ~1e-7), bounded at 2e-3 (a
    transposed operand or a wrong tile would move every output)."""
    for k in range(5):
        d = np.abs(np.log(fused[k].astype(np.float64) + 1e-8) - np.log(layerwise[k] + 1e-8))
        print(f"P{k + 1}: max |d log| {d.max():.2e}, median {np.median(d):.2e}")
        assert d.max() <= 5e-2 and np.median(d) <= 2e-3, k + 1


def test_fused_prediction_matches_layerwise_full_grid(atm, net):
    """The fused kernel (layer-1 tables, layers 2-4 in registers) against the layer-by-layer products at
    the full 5-order grid (S 32x64x32x8: many tiles per wavefront)."""
    sg, theta, _ = net
    sg.set_weights(theta.astype(np.float32), trained=False)
    P = atm.workload("full-5-order")
    q = sg.predict(P)
    n = int(np.prod(P.scatter_shape))
    entries = np.sort(np.random.default_rng(9).choice(n, 2048, replace=False))
    entries[-1] = n - 1
    assert_same_prediction(q.reshape(5, n, 3)[:, entries], layerwise_prediction(sg, P, entries, 256))


def _oracle_order_sums(oracle, OP):
    R1 = oracle.stage_single_scatter(OP, 0)
    M1 = oracle.stage_single_scatter(OP, 1)
    dS, dE = R1, oracle.stage_surface_radiance_base(OP)
    sums, total = [], 0.0
    for k in range(1, 6):
        first = 1 if k == 1 else 0
        mie = M1 if first else None
        J = oracle.stage_point_scatter(OP, first, dS, mie, dE)
        dE = oracle.stage_surface_radiance(OP, first, dS, mie)
        dS = oracle.stage_ray_scatter(OP, J)
        total = total + dS
        sums.append(total)
    return sums


def test_order_sums_match_oracle(atm, gpu, oracle):
    P = small5(atm)
    got = atm.order_sums(gpu, P)
    ref = _oracle_order_sums(oracle, oracle_params(oracle, P))
    for k in range(5):
        assert_close(f"P{k + 1}", got[k], ref[k], TOL_SCATTER)
    assert (got[0] >= 0).all() and all((got[k + 1] >= got[k]).all() for k in range(4))


@pytest.fixture(scope="module")
def trained(atm, gpu):
    """A network trained on the test atmosphere and two neighbours (small grid)."""
    P = small5(atm)
    rng = np.random.default_rng(3)
    family = [P] + [atm.random_atmosphere(rng, P) for _ in range(2)]
    sg = atm.Surrogate(gpu, seed=11, batch=1024)
    losses = sg.train(family, steps=1500, lr=3e-3, seed=5)
    return sg, P, family, losses


def test_training_reduces_loss(trained):
    sg, P, family, losses = trained
    assert np.all(np.isfinite(losses))
    first, last = losses[:10].mean(), losses[-50:].mean()
    print(f"loss {first:.4g} -> {last:.4g}")
    assert last < 0.02 * first


def test_training_is_deterministic(atm, gpu, trained):
    sg, P, family, _ = trained
    w = sg.weights()
    other = atm.Atmosphere(0)
    try:
        sg2 = atm.Surrogate(other, seed=11, batch=1024)
        sg2.train(family, steps=1500, lr=3e-3, seed=5)
        assert np.array_equal(sg2.weights(), w)
    finally:
        other.close()


def test_fused_prediction_matches_layerwise_small_grid(atm, trained):
    """A grid whose entry count (1848) is not a multiple of the 32-entry tile."""
    sg, P, family, _ = trained
    n = int(np.prod(P.scatter_shape))
    assert_same_prediction(sg.predict(P).reshape(5, n, 3), layerwise_prediction(sg, P, np.arange(n), 1024))


def test_prediction_tracks_order_sums(atm, gpu, trained):
    sg, P, family, _ = trained
    sums = atm.order_sums(gpu, P)
    pred = sg.predict(P)
    for k in range(5):
        err = rel_errors(pred[k], sums[k])
        print(f"P{k + 1}: median {np.median(err):.3e}, max {err.max():.3e}")
        assert np.median(err) < 0.1


def _error_stats(got, ref):
    e = np.concatenate([rel_errors(g, r).reshape(-1) for g, r in zip(got, ref)])
    return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e))), "median": float(np.median(e))}


@pytest.mark.parametrize("applications", [1, 2])
def test_seeded_build_beats_the_steps_without_network(atm, gpu, trained, applications):
    """What the trained network gains: the seeded build against the same exact steps taken from a zero
    seed (applications + 1 scattering orders instead of five), both measured against the plain 5-order
    build (surface radiance and ray scatter tables)."""
    sg, P, family, _ = trained
    gpu.build(P)
    plain = [gpu.fetch(P, i) for i in range(4)]
    r, fell_back = atm.build_seeded(gpu, P, applications, tolerance=np.inf)
    assert not fell_back and np.isfinite(r) and r >= 0
    seeded = [gpu.fetch(P, i) for i in range(4)]
    assert np.array_equal(seeded[0], plain[0]) and np.array_equal(seeded[3], plain[3])
    w = sg.weights()
    zero = np.zeros_like(w)
    zero[-16:-1] = 10.0 + np.log(1e-8)  # decodes to P = 0
    sg.set_weights(zero, trained=True)
    try:
        atm.build_seeded(gpu, P, applications, tolerance=np.inf)
        trunc = [gpu.fetch(P, i) for i in range(4)]
    finally:
        sg.set_weights(w, trained=True)
    es = _error_stats(seeded[1:3], plain[1:3])
    et = _error_stats(trunc[1:3], plain[1:3])
    print(f"m={applications}: seeded {es} (residual {r:.3e}); without network {et}")
    assert es["rms"] < 0.5 * et["rms"] and es["median"] < 0.5 * et["median"]


@pytest.mark.parametrize("grid,applications", [("small", 1), ("small", 2), ("small", 4), ("full", 1), ("full", 2)])
def test_seeded_build_with_exact_sums_is_the_plain_build(atm, gpu, grid, applications):
    """Seeded with the plain iteration's own order sums, m applications reproduce the plain 5-order
    build up to FP32 rounding and the residual vanishes to the same level: the seeded algebra is exact,
    including the order-by-order re-sampling of the accumulated ray-scatter table, which is not the
    identity where the sun direction is clamped (the full grid has such entries)."""
    P = small5(atm) if grid == "small" else atm.workload("full-5-order")
    gpu.build(P)
    plain = [gpu.fetch(P, i) for i in range(4)]
    sums = atm.order_sums(gpu, P)
    r, fell_back = atm.build_from_sums(gpu, P, sums, applications, tolerance=1e-4)
    assert not fell_back and r <= 1e-4
    for i, name in enumerate(("transmittance", "surface-radiance", "ray-scatter", "mie-strength")):
        assert_close(name, gpu.fetch(P, i), plain[i], 1e-4)


def test_seeded_build_falls_back_to_plain(atm, gpu, trained):
    sg, P, family, _ = trained
    gpu.build(P)
    plain = [gpu.fetch(P, i) for i in range(4)]
    r, fell_back = atm.build_seeded(gpu, P, 2, tolerance=0.0)
    assert r > 0 and fell_back
    for i in range(4):
        assert np.array_equal(gpu.fetch(P, i), plain[i])


def test_seeded_build_refuses_untrained_network(atm):
    ctx = atm.Atmosphere(0)
    try:
        atm.Surrogate(ctx, seed=1, batch=256)
        with pytest.raises(atm.AtmoError):
            atm.build_seeded(ctx, small5(atm))
        with pytest.raises(atm.AtmoError):
            atm.build_seeded(ctx, small5(atm, iterations=3))
        with pytest.raises(atm.AtmoError):
            atm.build_seeded(ctx, small5(atm), applications=5)
    finally:
        ctx.close()


def test_surrogate_argument_errors(atm, gpu):
    ctx = atm.Atmosphere(0)
    try:
        with pytest.raises(atm.AtmoError):
            atm.Surrogate(ctx, seed=1, batch=100)  # not a multiple of 64
        sg = atm.Surrogate(ctx, seed=1, batch=64)
        with pytest.raises(atm.AtmoError):
            sg.eval(np.zeros((65, 10), np.float32))  # more rows than the batch capacity
        bad = np.zeros(atm.surrogate_size(), np.float32)
        bad[3] = np.nan
        with pytest.raises(atm.AtmoError):
            sg.set_weights(bad)
    finally:
        ctx.close()
