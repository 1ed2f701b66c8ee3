"""Writes tests/golden/oracle_<workload>.npz: a seeded subset of every final table of a full LUT build
computed by the FP64 oracle (oracle/atmo_oracle.c), in .scatter file layout.

The full oracle build of these workloads takes minutes on a few cores, too long for a test; the
fixture lets the GPU tests and bench.py check the device's end-to-end output at full size.

    python tests/golden/make_oracle_subsets.py full-5-order [threads]
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

SUBSET = 4096
SEED = 20261202


def main(name, threads=0):
    P = atm.workload(name)
    OP = oracle_params(oracle, P)
    t0 = time.time()
    T, E, S, M = oracle.generate_atmosphere_luts(OP, threads=threads)
    elapsed = time.time() - t0
    tables = {"transmittance": T.reshape(-1, 3), "surface-radiance": E.reshape(-1, 3),
              "ray-scatter": oracle.convert_4d_to_2d(S).reshape(-1, 3),
              "mie-strength": oracle.convert_4d_to_2d(M).reshape(-1, 3)}
    rng = np.random.default_rng(SEED)
    out = {"oracle_seconds": np.float64(elapsed), "oracle_threads": np.int64(threads or oracle.max_threads())}
    for key, tab in tables.items():
        n = tab.shape[0]
        idx = np.sort(rng.choice(n, size=min(SUBSET, n), replace=False))
        out[key + "/index"] = idx.astype(np.int64)
        out[key + "/value"] = tab[idx]
        out[key + "/channel_max"] = np.max(np.abs(tab), axis=0)
    path = os.path.join(HERE, f"oracle_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{path}: oracle build {elapsed:.1f} s")


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 0)
