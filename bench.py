"""Benchmark of the atmosphere LUT precompute on one MI355X.

A step is one complete LUT build (transmittance, surface radiance, ray scatter and Mie strength with
5 scattering orders) of the full 5-order workload: T 64x256, E 16x64, S/M 32x64x32x8, 64 ray steps,
14 sphere steps (66 directions), 100 km atmosphere.  Its inputs are the parameters alone, so the
timed region starts with nothing but parameters on the host and ends with the four final tables in
device memory (sfsim_atmo_build, which replays the build's captured HIP graph; copying the tables
out is not part of a step).  Per-kernel durations come from the same number of kernel-by-kernel
builds with HIP events (sfsim_atmo_build_timed) after the timed steps.

Prints one JSON line with the throughput (LUT entries/s), the roofline of the dominant kernel
(ray scattering, five launches per build, measured with HIP events on the library's stream) and the
FP64 CPU oracle timed on a bounded sample of the same workload on this host.

    python bench.py [--gpus 1] [--steps 50] [--warmup 10] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LUT entries/s (5-order multi-scatter, 1 GPU) at ≤1e-3 rel err vs FP64 oracle"
WORKLOAD = "full-5-order"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
OPS_PER_LOOKUP4 = 136     # 4D multilinear RGB lookup: 15 lerps x 3 channels x 2 + 46 coordinate/index operations
OPS_PER_RAY_SAMPLE = OPS_PER_LOOKUP4 + 6  # ray scatter: one dJ lookup and T(x, p_i) * value accumulated (3 FMA)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-seeded", action="store_true", help="skip the surrogate-seeded leg")
    ap.add_argument("--train-steps", type=int, default=24000)
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: min(16, cores))")
    return ap.parse_args()


def check_against_fixture(atm, ctx, P, name):
    """Max and RMS relative error of this build against the committed FP64-oracle subset, per table."""
    path = os.path.join(ROOT, "tests", "golden", f"oracle_{name}.npz")
    if not os.path.exists(path):
        return None
    from tests.parity import report
    ref = np.load(path)
    out = {}
    for which, key in enumerate(["transmittance", "surface-radiance", "ray-scatter", "mie-strength"]):
        tab = ctx.fetch(P, which).reshape(-1, 3)
        idx = ref[key + "/index"]
        r = report(key, tab[idx], ref[key + "/value"], ref[key + "/channel_max"])
        out[key] = {"max": float(f"{r['max_rel']:.3e}"), "rms": float(f"{r['rms_rel']:.3e}")}
    return out


def cpu_baseline(P, threads):
    """Time the FP64 oracle per stage on seeded entry subsets and extrapolate to a whole build."""
    from oracle import atmo as oracle
    from tests.parity import oracle_params
    OP = oracle_params(oracle, P)
    rng = np.random.default_rng(7)
    nT = P.transmittance_shape[0] * P.transmittance_shape[1]
    nE = P.surface_radiance_shape[0] * P.surface_radiance_shape[1]
    nS = int(np.prod(P.scatter_shape))
    tabS = rng.uniform(1e-6, 1e-3, (nS, 3))
    tabE = rng.uniform(1e-3, 1e-1, (nE, 3))
    sub = lambda n, k: np.sort(rng.choice(n, size=min(n, k), replace=False))
    k4 = 32768
    stages = [
        ("transmittance", nT, lambda i: oracle.stage_transmittance(OP, i, threads)),
        ("surface_radiance_base", nE, lambda i: oracle.stage_surface_radiance_base(OP, i, threads)),
        ("single_scatter_rayleigh", nS, lambda i: oracle.stage_single_scatter(OP, 0, i, threads)),
        ("single_scatter_mie", nS, lambda i: oracle.stage_single_scatter(OP, 1, i, threads)),
    ]
    per_iter = [
        ("point_scatter_first", nS, lambda i: oracle.stage_point_scatter(OP, 1, tabS, tabS, tabE, i, threads), 1),
        ("point_scatter", nS, lambda i: oracle.stage_point_scatter(OP, 0, tabS, None, tabE, i, threads), -1),
        ("surface_radiance_first", nE, lambda i: oracle.stage_surface_radiance(OP, 1, tabS, tabS, i, threads), 1),
        ("surface_radiance", nE, lambda i: oracle.stage_surface_radiance(OP, 0, tabS, None, i, threads), -1),
        ("ray_scatter", nS, lambda i: oracle.stage_ray_scatter(OP, tabS, i, threads), 0),
        ("accumulate_E", nE, lambda i: oracle.stage_resample(OP, 3, tabE, tabE, i, threads), 0),
        ("accumulate_S", nS, lambda i: oracle.stage_resample(OP, 4, tabS, tabS, i, threads), 0),
    ]
    finals = [("final_T", nT, lambda i: oracle.stage_resample(OP, 2, tabE[:1].repeat(nT, 0), None, i, threads)),
              ("final_E", nE, lambda i: oracle.stage_resample(OP, 3, tabE, None, i, threads)),
              ("final_S", nS, lambda i: oracle.stage_resample(OP, 4, tabS, None, i, threads)),
              ("final_M", nS, lambda i: oracle.stage_resample(OP, 4, tabS, None, i, threads))]
    total, sampled_entries, sampled_seconds = 0.0, 0, 0.0
    iters = P.iterations

    def timed(n, fn):
        nonlocal sampled_entries, sampled_seconds
        idx = sub(n, k4 if n > 2 * k4 else n)
        t0 = time.perf_counter()
        fn(idx)
        dt = time.perf_counter() - t0
        sampled_entries += len(idx)
        sampled_seconds += dt
        return dt / len(idx) * n

    for _, n, fn in stages + finals:
        total += timed(n, fn)
    for _, n, fn, kind in per_iter:
        count = {1: min(iters, 1), -1: max(iters - 1, 0), 0: iters}[kind]
        if count:
            total += count * timed(n, fn)
    return total, sampled_entries, sampled_seconds


def seeded_leg(atm, ctx, P, workload, steps, train_steps):
    """The surrogate-seeded 5-order build next to the plain one: the network is trained on the GPU on
    the plain iteration's order sums of this atmosphere and 7 neighbours drawn from the training ranges
    (training time reported separately), then seeded builds with m = 1 and m = 2 exact applications are
    timed and checked against the same FP64-oracle fixture.  Tolerance infinity: the residual is
    reported, not acted on, so the timed work is the seeded path itself."""
    rng = np.random.default_rng(20261202)
    family = [P] + [atm.random_atmosphere(rng, P) for _ in range(7)]
    sg = atm.Surrogate(ctx, seed=1, batch=8192)
    t0 = time.perf_counter()
    losses = sg.train(family, steps=train_steps, lr=3e-3, seed=2)
    train_s = time.perf_counter() - t0
    out = {"train": {"atmospheres": len(family), "steps": train_steps, "batch": 8192, "seconds": round(train_s, 2),
                     "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-20:]))}}
    n_entries = atm.entries(P)
    for m in (1, 2):
        atm.build_seeded(ctx, P, m, tolerance=float("inf"))
        t0 = time.perf_counter()
        for _ in range(steps):
            r, _ = atm.build_seeded(ctx, P, m, tolerance=float("inf"))
        ms = (time.perf_counter() - t0) / steps * 1e3
        out[f"m{m}"] = {"ms_per_step": round(ms, 4), "value": round(n_entries / (ms / 1e3), 1), "residual": r,
                        "falls_back_at_1e-3": bool(r > 1e-3),
                        "rel_err_vs_oracle": check_against_fixture(atm, ctx, P, workload)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the LUT precompute runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported; across GPUs the work splits by atmosphere: independent "
              f"builds, one process per GPU, no exchange between stages).", file=sys.stderr)
        sys.exit(2)
    from sfsim_amd import atmosphere as atm
    P = atm.workload(args.workload)
    ctx = atm.Atmosphere(0)
    for _ in range(max(args.warmup, 1)):
        ctx.build(P)  # the first build captures the graph that the timed steps replay
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.build(P)  # synchronous: returns after the stream has drained
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    # per-kernel durations: the same build launched kernel by kernel with HIP events on its stream
    kernel_ms = {}
    for _ in range(args.steps):
        ctx.build_timed(P)
        for kind, (ms, n) in ctx.stage_timings().items():
            a = kernel_ms.setdefault(kind, [0.0, 0])
            a[0] += ms
            a[1] += n
    n_entries = atm.entries(P)
    value = n_entries / (ms_per_step / 1e3)
    errors = check_against_fixture(atm, ctx, P, args.workload)

    # roofline of the dominant kernel: ray scattering (algorithmic work / HIP-event duration per launch).
    # One launch evaluates every (entry, ray sample) pair once: scatter entries x ray_steps lookups.
    work = ctx.count_work(P)
    rs_ms, rs_n = kernel_ms["ray_scatter"]
    rs_avg_ms = rs_ms / max(rs_n, 1)
    rs_samples = int(np.prod(P.scatter_shape)) * P.ray_steps
    rs_flops = OPS_PER_RAY_SAMPLE * rs_samples
    achieved = rs_flops / (rs_avg_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_ray_scatter_full-5-order.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    roofline = {"bound": "valu", "kernel": "ray_scatter", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                "avg_launch_ms": round(rs_avg_ms, 4), "algorithmic_flops": rs_flops,
                "samples_per_launch": rs_samples, "work": work}

    seeded = None
    if not args.no_seeded and P.iterations == 5:
        seeded = seeded_leg(atm, ctx, P, args.workload, args.steps, args.train_steps)
        for m in ("m1", "m2"):
            seeded[m]["speedup_vs_plain"] = round(ms_per_step / seeded[m]["ms_per_step"], 3)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        est_s, k, s = cpu_baseline(P, threads)
        cpu = {"value": round(n_entries / est_s, 3), "unit": "entries/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle per stage on seeded entry subsets ({k} entries, {s:.1f} s wall), "
                         f"extrapolated to one whole build ({est_s:.0f} s)"}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 geometry)", "data": "synthetic (parameters only)",
        "config": {"workload": args.workload, "transmittance": list(P.transmittance_shape),
                   "surface_radiance": list(P.surface_radiance_shape), "scatter": list(P.scatter_shape),
                   "ray_steps": P.ray_steps, "sphere_steps": P.sphere_steps, "iterations": P.iterations,
                   "entries_per_step": n_entries},
        "rel_err_vs_oracle": errors,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kernel_ms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "surrogate_seeded": seeded,
    }
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
