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
HBM_PEAK_GBS = 8000.0     # MI355X published HBM3E bandwidth
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
    ap.add_argument("--batch", type=int, default=16,
                    help="atmospheres of the batched-build leg (sfsim_atmo_build_batch); 0 skips it")
    ap.add_argument("--train-steps", type=int, default=60000)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (default: the CPUs this process may use, capped by OMP_NUM_THREADS)")
    ap.add_argument("--cpu-stage-seconds", type=float, default=0.5,
                    help="CPU baseline: grow each stage's sample until it runs this long")
    return ap.parse_args()


TOLERANCE = {"transmittance": 1e-4, "surface-radiance": 1e-3, "ray-scatter": 1e-3, "mie-strength": 1e-3}


def check_against_fixture(atm, ctx, P, name):
    """Max and RMS relative error of the tables of the last build on ctx against the committed FP64-oracle
    fixture of the workload (every T and E entry; structured subsets of the 4D tables), per table, and
    whether every table is within its bound (1e-4 transmittance, 1e-3 the rest)."""
    path = os.path.join(ROOT, "tests", "golden", f"oracle_{name}.npz")
    if not os.path.exists(path):
        return None, None
    from tests.parity import fixture_errors
    ref = np.load(path)
    tables = {key: ctx.fetch(P, which) for which, key in enumerate(TOLERANCE)}
    errs = fixture_errors(ref, tables, tuple(P.scatter_shape))
    out = {k: {"max": float(f"{r['max_rel']:.3e}"), "rms": float(f"{r['rms_rel']:.3e}"), "entries": r["entries"]}
           for k, r in errs.items()}
    return out, all(errs[k]["max_rel"] <= TOLERANCE[k] for k in errs)


def host_info():
    """The host's CPU as this process sees it: logical CPUs, the ones this process may run on, the OpenMP
    thread limit of the environment and the model name."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": model}


def cpu_baseline(P, threads, structure, per_stage_s):
    """Time the FP64 oracle per stage of a whole build on seeded entry subsets, each subset grown until the
    stage ran for per_stage_s, and extrapolate to the whole build.  structure = 1: every stage parallel over
    the table's first dimension only (at most shape[0] threads), as the reference's pmap over dimension 0
    (interpolate.clj:72); separate Rayleigh and Mie single-scatter passes and the O(n^2) transmittance
    sums in every mode, as the reference computes them."""
    from oracle import atmo as oracle
    from tests.parity import oracle_params
    OP = oracle_params(oracle, P)
    rng = np.random.default_rng(7)
    nT = P.transmittance_shape[0] * P.transmittance_shape[1]
    nE = P.surface_radiance_shape[0] * P.surface_radiance_shape[1]
    nS = int(np.prod(P.scatter_shape))
    tabS = rng.uniform(1e-6, 1e-3, (nS, 3))
    tabE = rng.uniform(1e-3, 1e-1, (nE, 3))
    tabT = rng.uniform(0.1, 1.0, (nT, 3))
    kw = {"threads": threads, "structure": structure}
    stages = [
        (1, nT, lambda i: oracle.stage_transmittance(OP, i, **kw)),
        (1, nE, lambda i: oracle.stage_surface_radiance_base(OP, i, **kw)),
        (1, nS, lambda i: oracle.stage_single_scatter(OP, 0, i, **kw)),
        (1, nS, lambda i: oracle.stage_single_scatter(OP, 1, i, **kw)),
        (min(P.iterations, 1), nS, lambda i: oracle.stage_point_scatter(OP, 1, tabS, tabS, tabE, i, **kw)),
        (max(P.iterations - 1, 0), nS, lambda i: oracle.stage_point_scatter(OP, 0, tabS, None, tabE, i, **kw)),
        (min(P.iterations, 1), nE, lambda i: oracle.stage_surface_radiance(OP, 1, tabS, tabS, i, **kw)),
        (max(P.iterations - 1, 0), nE, lambda i: oracle.stage_surface_radiance(OP, 0, tabS, None, i, **kw)),
        (P.iterations, nS, lambda i: oracle.stage_ray_scatter(OP, tabS, i, **kw)),
        (P.iterations, nE, lambda i: oracle.stage_resample(OP, 3, tabE, tabE, i, **kw)),
        (P.iterations, nS, lambda i: oracle.stage_resample(OP, 4, tabS, tabS, i, **kw)),
        (1, nT, lambda i: oracle.stage_resample(OP, 2, tabT, None, i, **kw)),
        (1, nE, lambda i: oracle.stage_resample(OP, 3, tabE, None, i, **kw)),
        (2, nS, lambda i: oracle.stage_resample(OP, 4, tabS, None, i, **kw)),
    ]
    total, sampled_entries, sampled_seconds = 0.0, 0, 0.0
    for count, n, fn in stages:
        if not count:
            continue
        k = min(n, 64)
        while True:
            idx = np.sort(rng.choice(n, size=k, replace=False))
            t0 = time.perf_counter()
            fn(idx)
            dt = time.perf_counter() - t0
            sampled_entries += k
            sampled_seconds += dt
            if dt >= per_stage_s or k == n:
                break
            k = min(n, max(2 * k, int(k * per_stage_s / max(dt, 1e-6) * 1.2)))
        total += count * dt / k * n
    return total, sampled_entries, sampled_seconds


def seeded_leg(atm, ctx, P, workload, steps, train_steps):
    """The surrogate-seeded 5-order build next to the plain one: the network is trained on the GPU on
    the plain iteration's order sums of this atmosphere and 3 neighbours drawn from the training ranges
    (4 x 524,288 = 2.1 M entries; training time reported separately), then seeded builds with m = 1..4 exact
    applications are timed and checked against the same FP64-oracle fixture.  Tolerance infinity: the
    residual is reported, not acted on, so the timed work is the seeded path itself."""
    rng = np.random.default_rng(20261202)
    family = [P] + [atm.random_atmosphere(rng, P) for _ in range(3)]
    sg = atm.Surrogate(ctx, seed=1, batch=8192)
    t0 = time.perf_counter()
    losses = sg.train(family, steps=train_steps, lr=3e-3, seed=2)
    train_s = time.perf_counter() - t0
    out = {"train": {"atmospheres": len(family), "steps": train_steps, "batch": 8192, "seconds": round(train_s, 2),
                     "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-20:]))}}
    n_entries = atm.entries(P)
    for m in (1, 2, 3, 4):
        atm.build_seeded(ctx, P, m, tolerance=float("inf"))
        t0 = time.perf_counter()
        for _ in range(steps):
            r, _ = atm.build_seeded(ctx, P, m, tolerance=float("inf"))
        ms = (time.perf_counter() - t0) / steps * 1e3
        out[f"m{m}"] = {"ms_per_step": round(ms, 4), "value": round(n_entries / (ms / 1e3), 1), "residual": r,
                        "falls_back_at_1e-3": bool(r > 1e-3),
                        "rel_err_vs_oracle": check_against_fixture(atm, ctx, P, workload)[0]}
    return out


def batch_leg(atm, ctx, P, count, steps):
    """The batched build (sfsim_atmo_build_batch): `count` atmospheres drawn from the training ranges on the
    workload's grid in one launch sequence per build, after one untimed warm-up build; throughput per
    atmosphere and for the whole batch."""
    rng = np.random.default_rng(20261203)
    family = [P] + [atm.random_atmosphere(rng, P) for _ in range(count - 1)]
    ctx.build_batch(family)
    reps = max(1, min(steps, 10))
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.build_batch(family)
    ms = (time.perf_counter() - t0) / reps * 1e3
    n_entries = atm.entries(P) * count
    return {"atmospheres": count, "builds": reps, "ms_per_batch": round(ms, 3),
            "ms_per_atmosphere": round(ms / count, 4), "value": round(n_entries / (ms / 1e3), 1),
            "unit": "entries/s"}


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
    # the tables of the timed (graph-replayed) builds against the FP64 oracle's
    errors, within = check_against_fixture(atm, ctx, P, args.workload)
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

    # roofline of the dominant kernel: ray scattering (algorithmic work / HIP-event duration per launch).
    # One launch evaluates every (entry, ray sample) pair once: scatter entries x ray_steps lookups.
    work = ctx.count_work(P)
    rs_ms, rs_n = kernel_ms["ray_scatter"]
    rs_avg_ms = rs_ms / max(rs_n, 1)
    rs_samples = int(np.prod(P.scatter_shape)) * P.ray_steps
    rs_flops = OPS_PER_RAY_SAMPLE * rs_samples
    achieved = rs_flops / (rs_avg_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "kernel": "ray_scatter", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "avg_launch_ms": round(rs_avg_ms, 4), "launches": rs_n, "algorithmic_flops": rs_flops,
                "samples_per_launch": rs_samples, "work": work}
    # counters of the same kernel on the same workload (rocprofv3 --pmc runs committed under profiles/):
    # HBM bytes per launch and the fraction of cycles the SIMDs issued VALU instructions
    tpath = os.path.join(ROOT, "profiles", f"traffic_ray_scatter_{args.workload}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
        roofline["traffic"] = traffic
        roofline["hbm_gbs"] = round(traffic / (rs_avg_ms * 1e-3) / 1e9, 1)
        roofline["hbm_frac"] = round(roofline["hbm_gbs"] / HBM_PEAK_GBS, 4)
    vpath = os.path.join(ROOT, "profiles", f"valu_ray_scatter_{args.workload}.json")
    if os.path.exists(vpath):
        v = json.load(open(vpath))
        roofline["valu_busy"] = round(v["valu_busy"], 3)
        roofline["valu_issue"] = round(v["valu_issue"], 3)

    seeded = None
    if not args.no_seeded and P.iterations == 5:
        seeded = seeded_leg(atm, ctx, P, args.workload, args.steps, args.train_steps)
        for m in ("m1", "m2", "m3", "m4"):
            seeded[m]["speedup_vs_plain"] = round(ms_per_step / seeded[m]["ms_per_step"], 3)

    batch = batch_leg(atm, ctx, P, args.batch, args.steps) if args.batch > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        host = host_info()
        share = args.cpu_threads or min(host["affinity"], int(host["omp_num_threads"] or host["affinity"]))
        modes = {}
        for mode, threads, structure in (("all-entries", share, 0), ("reference-structure", share, 1),
                                         ("single-thread", 1, 0)):
            est_s, k, s = cpu_baseline(P, threads, structure, args.cpu_stage_seconds)
            modes[mode] = {"value": round(n_entries / est_s, 3), "threads": threads, "build_s": round(est_s, 2),
                           "sampled_entries": k, "sampled_s": round(s, 2)}
        best = modes["all-entries"]
        cpu = {"value": best["value"], "unit": "entries/s", "cores": share, "kind": "port",
               "sample": f"FP64 oracle (a literal port) per stage on seeded entry subsets grown to "
                         f"{args.cpu_stage_seconds} s per stage, extrapolated to one whole build; headline mode: "
                         f"every stage parallel over all entries on {share} threads",
               "modes": modes, "host": host}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 geometry)", "data": "synthetic (parameters only)",
        "config": {"workload": args.workload, "transmittance": list(P.transmittance_shape),
                   "surface_radiance": list(P.surface_radiance_shape), "scatter": list(P.scatter_shape),
                   "ray_steps": P.ray_steps, "sphere_steps": P.sphere_steps, "iterations": P.iterations,
                   "entries_per_step": n_entries},
        "rel_err_vs_oracle": errors,
        "rel_err_within_bounds": within,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kernel_ms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "surrogate_seeded": seeded,
        "batched": batch,
    }
    ctx.close()
    print(json.dumps(line), flush=True)
    if within is False:
        print("bench.py: the timed build's tables exceed the error bound against the FP64 oracle", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
