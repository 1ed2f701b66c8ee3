"""Writes tests/golden/oracle_<workload>.npz: the final tables of a whole LUT build computed by the FP64
oracle (oracle/atmo_oracle.c), every entry of the transmittance and surface-radiance tables and a
structured subset of the two 4D tables (ray scatter, Mie strength):

  - every view ray: each (height, elevation) grid pair at four (sun-elevation, heading) pairs, the
    corners and the middle of that plane;
  - every entry on the horizon rows of the elevation axis (index (N - 1) / 2 for odd N, both rows
    next to it for even N; atmosphere.clj:239-270);
  - a seeded sample of the entries where index-to-sun-direction clamps the sun direction
    (atmosphere.clj:375-384), so that forward(backward(i)) != i and the accumulations re-sample
    between grid points (atmosphere_lut.clj:95-101);
  - a seeded uniform sample.

Keys: <table>/index (flat grid-order index: [h][e] for 2D tables, [h][e][s][a] for 4D tables, before
the 4D->2D tiling of the files), <table>/value (float32 RGB), <table>/channel_max (float64, over the
whole table), and for the 4D tables <table>/kind (bit mask of the categories above: 1 ray, 2 horizon,
4 clamped, 8 uniform).

The full oracle build takes minutes (production: ~10 min on 8 cores), too long for a test; the
fixture lets the GPU tests and bench.py check the device's end-to-end output at full size.

    python tests/golden/make_oracle_subsets.py <workload> [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import atmo as oracle  # noqa: E402
from sfsim_amd import atmosphere as atm  # noqa: E402
from tests.parity import oracle_params  # noqa: E402

SEED = 20261202
CLAMPED_SAMPLE = 8192
UNIFORM_SAMPLE = 4096
KIND_RAY, KIND_HORIZON, KIND_CLAMPED, KIND_UNIFORM = 1, 2, 4, 8


def clamped_entries(OP, shape):
    """Boolean mask over the 4D grid: forward(backward(i)) differs from i (the sun direction was clamped)."""
    mask = np.zeros(shape, bool)
    for i in np.ndindex(*shape):
        x, v, l, above = oracle.space_backward(OP.planet, 4, shape, i)
        f = oracle.space_forward(OP.planet, 4, shape, x, v, l, above)
        mask[i] = np.max(np.abs(f - np.asarray(i, np.float64))) > 1e-7
    return mask


def horizon_rows(n):
    return [(n - 1) // 2] if n % 2 else [n // 2 - 1, n // 2]


def scatter_subset(OP, shape, rng):
    NH, NE, NL, NA = shape
    kind = np.zeros(shape, np.uint8)
    for il, ia in ((0, 0), (0, NA - 1), (NL - 1, 0), (NL - 1, NA - 1), (NL // 2, NA // 2)):
        kind[:, :, il, ia] |= KIND_RAY
    kind[:, horizon_rows(NE), :, :] |= KIND_HORIZON
    clamped = np.flatnonzero(clamped_entries(OP, shape))
    pick = rng.choice(clamped, size=min(CLAMPED_SAMPLE, clamped.size), replace=False)
    kind.reshape(-1)[pick] |= KIND_CLAMPED
    uniform = rng.choice(kind.size, size=min(UNIFORM_SAMPLE, kind.size), replace=False)
    kind.reshape(-1)[uniform] |= KIND_UNIFORM
    idx = np.flatnonzero(kind.reshape(-1))
    return idx, kind.reshape(-1)[idx], int(clamped.size)


def main(name, threads=0):
    P = atm.workload(name)
    OP = oracle_params(oracle, P)
    t0 = time.time()
    T, E, S, M = oracle.generate_atmosphere_luts(OP, threads=threads)
    elapsed = time.time() - t0
    rng = np.random.default_rng(SEED)
    out = {"oracle_seconds": np.float64(elapsed), "oracle_threads": np.int64(threads or oracle.max_threads())}
    for key, tab in (("transmittance", T), ("surface-radiance", E)):
        flat = tab.reshape(-1, 3)
        out[key + "/index"] = np.arange(flat.shape[0], dtype=np.int64)
        out[key + "/value"] = flat.astype(np.float32)
        out[key + "/channel_max"] = np.max(np.abs(flat), axis=0)
    idx, kind, n_clamped = scatter_subset(OP, tuple(P.scatter_shape), rng)
    out["clamped_entries"] = np.int64(n_clamped)
    for key, tab in (("ray-scatter", S), ("mie-strength", M)):
        flat = tab.reshape(-1, 3)
        out[key + "/index"] = idx.astype(np.int64)
        out[key + "/value"] = flat[idx].astype(np.float32)
        out[key + "/channel_max"] = np.max(np.abs(flat), axis=0)
        out[key + "/kind"] = kind
    path = os.path.join(HERE, f"oracle_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: oracle build {elapsed:.1f} s, {idx.size} of {S.shape[0] * S.shape[1] * S.shape[2] * S.shape[3]} "
          f"4D entries stored ({n_clamped} clamped in the grid)")


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 0)
