#!/usr/bin/env python3
"""Benchmark: fused SSP-RK3 lattice solver on one MI355X.

Headline workload: a single 512^3 periodic f32 box, 32 bubbles with seeded
centres (the reference's Pcg64Mcg stream, seed 1) and nucleation steps drawn
uniformly in [0, 200], eta drawn log-uniformly in [0.02, 1] from a seeded
sweep generator, default lambda / eps_ratio / alpha / cfl.  One "step" is one
full RK3 step (three fused stage launches, plus the nucleation stamp on the
steps where bubbles appear) of the whole box; the lattice is resident in HBM.

Prints ONE JSON line: site-updates/s (cells x steps / s), the roofline of the
fused stage kernel (192 algorithmic bytes per site-update against 8 TB/s),
and the CPU baseline (the oracle's f32 restatement of the reference scheme on
the host cores, timed on a bounded 128^3 sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "bubble-simulation_amd"), ROOT]

HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE_UPDATE = 192  # stage 0: 24 R + 24 W; stages 1, 2: 48 R + 24 W


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sweep_eta(bs, seed=7):
    u = float(bs.pcg64mcg_f32(seed, 1)[0])
    return math.exp(math.log(0.02) + u * (math.log(1.0) - math.log(0.02)))


def headline_spec(bs, n=512):
    model = bs.Model(eta=sweep_eta(bs))
    return bs.make_spec(grid=n, model=model, nucleation=bs.NUC_UNIFORM_STEPS, bubbles=32, seed=1,
                        uniform_max_step=200, time_seed=11, precision=bs.PRECISION_F32, math=bs.MATH_FAST,
                        kernel=bs.KERNEL_FUSED)


def cpu_baseline(bs, budget_s=12.0):
    """The oracle (f32, reference operation order) on the host cores."""
    from oracle import cpu as orc
    n = 128
    spec = headline_spec(bs, n)
    model = bs.Model(eta=sweep_eta(bs))
    params = bs.sim_params_words(model, (n, n, n))
    sched = bs.make_schedule(bs.NUC_UNIFORM_STEPS, 32, model.seed_radius(), n, 60.0, 1, 200, 11, model.dt())
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = orc.OracleSim(0, params, sched)
    o.advance(1)  # warm
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.advance(1)
        steps += 1
    dt = time.perf_counter() - t0
    o.close()
    del spec
    return {"value": n ** 3 * steps / dt, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n}^3 box of the same physics, {steps} RK3 steps in {dt:.1f} s, OpenMP {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import bubblesim as bs
    if bs.device_count() == 0:
        print("no HIP device", file=sys.stderr)
        return 1
    n = args.grid
    spec = headline_spec(bs, n)
    log(f"bench: {n}^3 f32 fused, eta={spec.model.eta:.4f}, device {bs.device_name()}")
    sim = bs.Simulation(spec)
    cells = sim.cell_count()
    sim.advance(args.warmup)
    sim.synchronize()
    log(f"bench: warmup {args.warmup} steps done")
    kernel_ms, wall_ms = sim.advance_timed(args.steps)
    # whole-job throughput without the per-launch event pairs
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance(args.steps)
    sim.synchronize()
    wall_s = time.perf_counter() - t0
    d = sim.diagnostics()
    log(f"bench: {args.steps} steps: wall {wall_s * 1e3:.1f} ms, stage kernels {kernel_ms:.1f} ms (timed pass); "
        f"broken {d.broken_fraction:.4f} max|v| {d.max_velocity:.4f} nonfinite {d.nonfinite}")
    value = cells * args.steps / wall_s
    launches = 3 * args.steps
    avg_launch_s = kernel_ms / 1e3 / launches
    bytes_per_launch = BYTES_PER_SITE_UPDATE / 3 * cells
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_fused_stage_512.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    cpu = None if args.no_cpu else cpu_baseline(bs)
    out = {
        "metric": "RK3 site-updates/s at 512^3 f32 (fused stage kernel, % of HBM peak)",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"single periodic {n}^3 box, 32 bubbles (seeded centres, steps U[0,200]), "
                               f"eta log-U[0.02,1] draw = {spec.model.eta:.4f}, SSP-RK3, fused stage kernel",
                   "grid": n, "bubbles": 32, "precision": "f32", "math": "fast"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "stage_kernel_ms_per_launch": avg_launch_s * 1e3,
    }
    print(json.dumps(out))
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
