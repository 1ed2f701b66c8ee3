"""The learned part on the full 5-order grid (T 64x256, E 16x64, S 32x64x32x8, 64 ray steps): the network is
trained on the GPU on the plain iteration's order sums of four atmospheres of that grid (2.1 M entries; the
base atmosphere and three drawn from the training ranges by a seeded generator), then the base atmosphere
is built seeded, with 1 to 4 exact applications of the iteration, and compared with its plain build.

What is asserted: training converges; the network's order sums are within 5% (RMS) of the exact ones;
every extra exact application shrinks the seeded build's error (the iteration contracts the seed's error);
the seeded build beats the same number of exact steps without the network; and at the 1e-3 bound the build
falls back to the plain iteration, whose tables it then returns bit for bit: the network's error after four
applications is above 1e-3 at the entries where the sun direction is clamped, so a 5-order table within
1e-3 needs the plain iteration.  (Four training atmospheres do not cover the six-parameter range: an
atmosphere drawn anew from the whole range is outside what this network has seen.)
"""
import numpy as np
import pytest

from tests.parity import rel_errors

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


@pytest.fixture(scope="module")
def family_net(atm, gpu):
    P = atm.workload("full-5-order")
    rng = np.random.default_rng(20261202)
    family = [P] + [atm.random_atmosphere(rng, P) for _ in range(3)]  # 4 x 524,288 entries
    sg = atm.Surrogate(gpu, seed=1, batch=8192)
    print("training on 4 atmospheres of the full grid", flush=True)
    losses = sg.train(family, steps=30000, lr=3e-3, seed=2)
    return sg, P, losses


def _stats(got, ref):
    e = np.concatenate([rel_errors(g, r).reshape(-1) for g, r in zip(got, ref)])
    return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}


def test_family_training_converges(family_net):
    _, _, losses = family_net
    first, last = float(losses[:10].mean()), float(losses[-100:].mean())
    print(f"loss {first:.4g} -> {last:.4g}")
    assert np.all(np.isfinite(losses)) and last < 2e-3 * first


def test_network_order_sums(atm, gpu, family_net):
    sg, Q, _ = family_net
    sums, pred = atm.order_sums(gpu, Q), sg.predict(Q)
    for k in range(5):
        s = _stats([pred[k]], [sums[k]])
        print(f"P{k + 1}: {s}")
        assert s["rms"] < 0.05


def test_seeded_builds(atm, gpu, family_net):
    sg, Q, _ = family_net
    gpu.build(Q)
    plain = [gpu.fetch(Q, i) for i in range(4)]
    errs = []
    for m in (1, 2, 3, 4):
        r, fell_back = atm.build_seeded(gpu, Q, m, tolerance=np.inf)
        assert not fell_back
        seeded = [gpu.fetch(Q, i) for i in range(4)]
        assert np.array_equal(seeded[0], plain[0]) and np.array_equal(seeded[3], plain[3])  # T, M1: exact stages
        errs.append(_stats(seeded[1:3], plain[1:3]))
        print(f"m={m}: surface radiance + ray scatter vs plain {errs[-1]}, residual {r:.3e}")
    for a, b in zip(errs, errs[1:]):
        assert b["rms"] < a["rms"] and b["max"] < a["max"]
    # the same two exact steps from a zero seed (3 orders instead of 5)
    w = sg.weights()
    zero = np.zeros_like(w)
    zero[-16:-1] = 10.0 + np.log(1e-8)  # decodes to P = 0
    sg.set_weights(zero, trained=True)
    try:
        atm.build_seeded(gpu, Q, 2, tolerance=np.inf)
        trunc = _stats([gpu.fetch(Q, i) for i in (1, 2)], plain[1:3])
    finally:
        sg.set_weights(w, trained=True)
    print(f"m=2 without the network: {trunc}")
    assert errs[1]["rms"] < 0.2 * trunc["rms"]
    # within the 1e-3 bound only the plain iteration: the seeded build falls back and returns its tables
    r, fell_back = atm.build_seeded(gpu, Q, 4, tolerance=1e-3)
    assert fell_back and r > 1e-3
    for i in range(4):
        assert np.array_equal(gpu.fetch(Q, i), plain[i])
