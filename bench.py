#!/usr/bin/env python3
"""Benchmark: fused SSP-RK3 lattice solver on one MI355X.

Headline workload: a single 512^3 periodic f32 box, 32 bubbles with seeded
centres (the reference's Pcg64Mcg stream, seed 1) and nucleation steps drawn
uniformly in [0, 200], eta drawn log-uniformly in [0.02, 1] from a seeded
sweep generator, default lambda / eps_ratio / alpha / cfl.  One "step" is one
full RK3 step (three fused stage launches) of the whole box; the lattice is
resident in HBM.  Before the timed window, and outside timing, the box is
advanced past its last nucleation (step 197 of this schedule), then by the
warm-up steps: the stage kernel's instruction and byte counts do not depend
on the state, but its clock does (the power limit allows a higher clock
while the box is still nearly uniform), so a window at the start of the run
would overstate the rate.  The default window (800 steps after the warm-up)
runs through the rest of the transition to the broken phase; the first 30
steps of the run and 30 steps after step 1000 (the box fully broken, the
slowest state) are reported beside it for scale.

Prints ONE JSON line: site-updates/s (cells x steps / s), the roofline of the
fused stage kernel (192 algorithmic bytes per site-update against 8 TB/s),
the CPU baseline (the oracle's f32 restatement of the reference scheme on
the host cores, timed on a bounded 128^3 sample), and the accuracy of the
measured f32 path: its field error against the same box stepped in f64, and
the energy drift of both.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "bubble-simulation_amd"), ROOT]

HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE_UPDATE = 192  # stage 0: 24 R + 24 W; stages 1, 2: 48 R + 24 W


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sweep_eta(bs, seed=7):
    u = float(bs.pcg64mcg_f32(seed, 1)[0])
    return math.exp(math.log(0.02) + u * (math.log(1.0) - math.log(0.02)))


def settle_steps(bs, sim):
    """Steps until every scheduled bubble of `sim` has been stamped: an
    event due at step s is stamped before step s."""
    sched = sim.schedule()
    if len(sched) == 0:
        return 0
    return int(bs.event_steps(sched, float(sim.spec.model.cfl * sim.spec.model.dx)).max()) + 1


def headline_spec(bs, n=512, eos=0, eos_gamma=4.0 / 3.0, precision=None, math_mode=None):
    model = bs.Model(eta=sweep_eta(bs))
    prec = bs.PRECISION_F32 if precision is None else precision
    return bs.make_spec(grid=n, model=model, nucleation=bs.NUC_UNIFORM_STEPS, bubbles=32, seed=1,
                        uniform_max_step=200, time_seed=11, precision=prec,
                        math=bs.MATH_FAST if math_mode is None else math_mode,
                        kernel=bs.KERNEL_FUSED, eos=eos, eos_gamma=eos_gamma)


ACC_STEPS = 260  # every bubble of the headline schedule has nucleated by step 200
BLOCKS = (("field", ("phi", "pi")), ("fluid", ("E", "Zx", "Zy", "Zz")))


def block_errors(a32, a64, names):
    """Squared-difference and squared-reference sums, and the largest
    |difference| and |reference|, of each component of an f32 state against
    the f64 one, accumulated in float64 one z plane at a time."""
    k = a64.shape[-1]
    se, sr, me, mr = np.zeros(k), np.zeros(k), np.zeros(k), np.zeros(k)
    for z in range(a64.shape[0]):
        r = a64[z].reshape(-1, k)
        d = a32[z].reshape(-1, k).astype(np.float64) - r
        se += np.einsum("ij,ij->j", d, d)
        sr += np.einsum("ij,ij->j", r, r)
        me = np.maximum(me, np.abs(d).max(axis=0))
        mr = np.maximum(mr, np.abs(r).max(axis=0))
    comps = {nm: {"rel_l2": float(np.sqrt(se[i] / sr[i])) if sr[i] > 0 else None,
                  "rel_linf": float(me[i] / mr[i]) if mr[i] > 0 else None} for i, nm in enumerate(names)}
    block = {"rel_l2": float(np.sqrt(se.sum() / sr.sum())), "rel_linf": float(me.max() / mr.max())}
    return block, comps


def accuracy(bs, n, eos, eos_gamma, steps=ACC_STEPS, chunk=20):
    """Field error and energy drift of the measured path.  The headline box
    is stepped from the same schedule in f64 (the ExactMath fused kernel:
    the oracle's scheme and operation order, which it equals bit for bit
    wherever the oracle runs), in f32 FastMath (the measured path) and in f32
    ExactMath (the reference's f32 operation order, for scale).  Reported:
    relative L2 and L-infinity of each f32 state against f64 after `steps`
    steps, per block (phi/pi and the fluid) and per component, and each run's
    energy drift E/E0 - 1, E0 taken at the first report after the last
    nucleation (headless.rs:54-78)."""
    runs = (("f32_fast", bs.PRECISION_F32, bs.MATH_FAST), ("f32_exact", bs.PRECISION_F32, bs.MATH_EXACT),
            ("f64", bs.PRECISION_F64, bs.MATH_EXACT))
    states, drift = {}, {}
    base_step = None
    for name, prec, mm in runs:
        sim = bs.Simulation(headline_spec(bs, n, eos, eos_gamma, precision=prec, math_mode=mm))
        try:
            base, done = None, 0
            while done < steps:
                c = min(chunk, steps - done)
                sim.advance(c)
                done += c
                if base is None and sim.remaining_bubbles() == 0:
                    base, base_step = sim.diagnostics().mean_energy, done
            d = sim.diagnostics()
            if d.nonfinite:
                raise RuntimeError(f"{name} run went non-finite")
            drift[name] = d.mean_energy / base - 1.0 if base else None
            states[name] = sim.download()
        finally:
            sim.close()
    out = {"reference": "f64 ExactMath fused kernel (bit-identical to the oracle at oracle sizes)", "steps": steps,
           "drift_from_step": base_step, "energy_drift": drift, "blocks": {}, "components": {}}
    for name in ("f32_fast", "f32_exact"):
        out["blocks"][name], out["components"][name] = {}, {}
        for k, (bname, names) in enumerate(BLOCKS):
            blk, comps = block_errors(states[name][k], states["f64"][k], names)
            out["blocks"][name][bname] = blk
            out["components"][name].update(comps)
    fb, eb = out["blocks"]["f32_fast"], out["blocks"]["f32_exact"]
    log(f"bench: accuracy after {steps} steps against f64, rel L2 field / fluid: f32 fast "
        f"{fb['field']['rel_l2']:.2e} / {fb['fluid']['rel_l2']:.2e}, f32 exact {eb['field']['rel_l2']:.2e} / "
        f"{eb['fluid']['rel_l2']:.2e}; energy drift from step {base_step}: "
        + ", ".join(f"{k} {v:+.3e}" for k, v in drift.items() if v is not None))
    return out


def cpu_baseline(bs, state, budget_s=12.0, single_budget_s=5.0):
    """The oracle (f32, reference operation order) on the host cores, stepping
    the headline box itself: `state` is the GPU's (field, fluid) at the end of
    the timed window, loaded into an oracle lattice of the same parameters
    with no nucleation left (every bubble is in by then), and whole RK3 steps
    are timed with OpenMP over this process's CPU share (OMP_NUM_THREADS,
    else the visible cores up to 16).  One thread beside it, on a 128^3 box
    of the same physics (a 512^3 step takes over a minute on one core)."""
    from oracle import cpu as orc
    if state is None:
        raise RuntimeError("no state of the headline box to step")
    field, fluid = state
    nz, ny, nx = field.shape[:3]
    model = bs.Model(eta=sweep_eta(bs))
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(visible, 16)

    def timed(o, cells, nthreads, budget):
        orc.set_threads(nthreads)
        steps = 0
        t0 = time.perf_counter()
        while steps == 0 or time.perf_counter() - t0 < budget:
            o.advance(1)
            steps += 1
        dt = time.perf_counter() - t0
        o.close()
        return cells * steps / dt, steps, dt

    o = orc.OracleSim(0, bs.sim_params_words(model, (nx, ny, nz)), np.zeros((0, 5), dtype=np.float32))
    o.set_state(field, fluid)
    rate, steps, dt = timed(o, nx * ny * nz, threads, budget_s)
    n1 = 128
    sched = bs.make_schedule(bs.NUC_UNIFORM_STEPS, 32, model.seed_radius(), n1, 60.0, 1, 200, 11, model.dt())
    o1 = orc.OracleSim(0, bs.sim_params_words(model, (n1, n1, n1)), sched)
    o1.advance(1)  # warm
    rate1, steps1, dt1 = timed(o1, n1 ** 3, 1, single_budget_s)
    log(f"bench: CPU baseline {rate / 1e6:.2f} M/s on {threads} threads ({visible} cores visible), "
        f"{rate1 / 1e6:.2f} M/s on 1")
    return {"value": rate, "unit": "site-updates/s", "cores": threads, "kind": "port",
            "sample": f"the headline {nx}x{ny}x{nz} box from the GPU's state at the end of the timed window, "
                      f"{steps} RK3 steps in {dt:.1f} s, OpenMP {threads} threads",
            "host_cores_visible": visible, "single_thread_value": rate1,
            "single_thread_sample": f"{steps1} RK3 steps of a {n1}^3 box of the same physics in {dt1:.1f} s"}


def sweep_specs(bs, n_boxes=64, master_seed=2024):
    """Batched-sweep boxes: edge U{32,48,64,96,128}, eta log-U[0.02,1],
    seed_factor U[0.7,1.5], eps_ratio U[0.6,1.0), 1-4 bubbles at seeded
    positions (seed k+1), nucleation steps U[0,200]."""
    u = bs.pcg64mcg_f32(master_seed, 6 * n_boxes)
    specs = []
    for k in range(n_boxes):
        a, b, c, d, e, f = (float(v) for v in u[6 * k:6 * k + 6])
        edge = [32, 48, 64, 96, 128][min(int(a * 5), 4)]
        eta = math.exp(math.log(0.02) + b * (0.0 - math.log(0.02)))
        model = bs.Model(eta=eta, seed_factor=0.7 + 0.8 * c, eps_ratio=min(0.6 + 0.4 * d, 0.999))
        specs.append(bs.make_spec(grid=edge, model=model, nucleation=bs.NUC_UNIFORM_STEPS, bubbles=1 + min(int(e * 4), 3),
                                  seed=k + 1, uniform_max_step=200, time_seed=1000 + k))
    return specs


def field_errors(bs_states, ref_states):
    """Per block (phi/pi, fluid): relative L2 and L-infinity of f32 states
    against f64 ones of the same boxes, over all the boxes together."""
    out = {}
    for k, (bname, _) in enumerate(BLOCKS):
        se = sr = 0.0
        me = mr = 0.0
        for a32, a64 in zip(bs_states, ref_states):
            r = a64[k].astype(np.float64)
            d = a32[k].astype(np.float64) - r
            se += float(np.sum(d * d))
            sr += float(np.sum(r * r))
            me = max(me, float(np.abs(d).max()))
            mr = max(mr, float(np.abs(r).max()))
        out[bname] = {"rel_l2": float(np.sqrt(se / sr)) if sr > 0 else None, "rel_linf": me / mr if mr > 0 else None}
    return out


def f64_spec(bs, spec):
    """The same box in f64 ExactMath."""
    s64 = bs.CSpec.from_buffer_copy(spec)
    s64.precision = bs.PRECISION_F64
    s64.math = bs.MATH_EXACT
    return s64


def exact32_spec(bs, spec):
    """The same box in f32 ExactMath (the reference's operation order)."""
    s = bs.CSpec.from_buffer_copy(spec)
    s.precision = bs.PRECISION_F32
    s.math = bs.MATH_EXACT
    return s


def accuracy_sample(specs):
    """Indices of the first box of each edge 32, 64, 96, 128 (or fewer)."""
    idx, seen = [], set()
    for i, s in enumerate(specs):
        e = int(s.grid[0])
        if e in (32, 64, 96, 128) and e not in seen:
            seen.add(e)
            idx.append(i)
    return idx


def batched_sweep(bs, steps=1000, report_at=200, use_graph=True):
    """64 uneven boxes in one launch per RK stage (hipGraph replay) against
    the same boxes solved one after another on the plain path, both as whole
    jobs of `steps` steps from the initial state.  The batch reports
    diagnostics of a sample of boxes (one per edge 32, 64, 96, 128) at step
    `report_at` (every bubble has been stamped by then) and at the end, as
    the headless run reports, inside its timed region; the same sample is
    then run in f64 through the same engine for the field error and drift.
    use_graph=False steps eagerly (for a run under a profiler)."""
    specs = sweep_specs(bs)
    sample = accuracy_sample(specs)
    b = bs.Batch(specs)
    cells = b.total_cells()
    b.synchronize()
    t0 = time.perf_counter()
    b.advance(report_at, use_graph=use_graph)
    e0 = {i: b.diagnostics(i).mean_energy for i in sample}
    b.advance(steps - report_at, use_graph=use_graph)
    b.synchronize()
    dt = time.perf_counter() - t0
    e1 = {i: b.diagnostics(i) for i in sample}
    states = [b.download(i) for i in sample]
    plan = b.plan()
    graph_steps = b.graph_steps()
    b.close()
    batched = cells * steps / dt
    log(f"bench: batched {len(specs)} boxes, {cells / 1e6:.1f} M cells, {steps} steps in {dt:.2f} s "
        f"= {batched / 1e9:.2f} G/s (plan {plan}, {graph_steps} steps by graph replay)")
    if any(e1[i].nonfinite for i in sample):
        raise RuntimeError("a batched box went non-finite")
    # f64 of the sample through the same engine
    def sample_run(make):
        bb = bs.Batch([make(bs, specs[i]) for i in sample])
        bb.advance(report_at, use_graph=use_graph)
        a0 = [bb.diagnostics(k).mean_energy for k in range(len(sample))]
        bb.advance(steps - report_at, use_graph=use_graph)
        a1 = [bb.diagnostics(k).mean_energy for k in range(len(sample))]
        st = [bb.download(k) for k in range(len(sample))]
        bb.close()
        return a0, a1, st

    f0, f1, ref = sample_run(f64_spec)          # the accuracy reference
    x0, x1, exact32 = sample_run(exact32_spec)  # the reference's f32 operation order, for scale
    acc = {"reference": "the same boxes in f64 ExactMath through the batched engine", "steps": steps,
           "boxes": [{"index": i, "edge": int(specs[i].grid[0]),
                      "energy_drift_f32": e1[i].mean_energy / e0[i] - 1.0,
                      "energy_drift_f32_exact": x1[k] / x0[k] - 1.0,
                      "energy_drift_f64": f1[k] / f0[k] - 1.0} for k, i in enumerate(sample)],
           "drift_from_step": report_at, "blocks": field_errors(states, ref),
           "blocks_f32_exact": field_errors(exact32, ref)}
    log("bench: batched accuracy (sample " + ", ".join(str(int(specs[i].grid[0])) for i in sample) + "): rel L2 "
        + ", ".join(f"{k} {v['rel_l2']:.2e}" for k, v in acc["blocks"].items()) + "; f32 exact "
        + ", ".join(f"{k} {v['rel_l2']:.2e}" for k, v in acc["blocks_f32_exact"].items()))
    seq_t, seq_work = 0.0, 0
    for s in specs:
        sim = bs.Simulation(s)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.advance(steps)
        sim.synchronize()
        seq_t += time.perf_counter() - t1
        seq_work += sim.cell_count() * steps
        sim.close()
    sequential = seq_work / seq_t
    log(f"bench: sequential plain solves {sequential / 1e9:.2f} G/s")
    return {"boxes": len(specs), "cells": cells, "steps": steps, "site_updates_per_s": batched,
            "sequential_site_updates_per_s": sequential, "speedup_vs_sequential": batched / sequential,
            "sequential_note": SEQUENTIAL_NOTE,
            "plan": plan, "graph_replayed_steps": graph_steps, "accuracy": acc}


SEQUENTIAL_NOTE = ("each box or job alone on the plain Simulation path, one after another, whole; that path "
                   "gives boxes too small to fill the device short z chunks (2 planes at 32^3-64^3, 4 at 96^3, "
                   "8 at 128^3) and 16x16 tiles where the x edge is an odd multiple of 16, so small boxes run "
                   "at 1.4-19 G site-updates/s alone rather than latency-bound")


def served_jobs(bs, n_jobs, rate_hz, seed):
    """Sweep jobs (edge in {48..192} step 16, 500-2000 steps, parameters as
    the batched sweep) with Poisson arrival times at rate_hz."""
    rng_u = bs.pcg64mcg_f32(seed, 8 * n_jobs)
    jobs = []
    t = 0.0
    for k in range(n_jobs):
        u = [float(v) for v in rng_u[8 * k:8 * k + 8]]
        edge = 48 + 16 * min(int(u[0] * 10), 9)
        steps = 500 + int(u[1] * 1501)
        eta = math.exp(math.log(0.02) + u[2] * (0.0 - math.log(0.02)))
        model = bs.Model(eta=eta, seed_factor=0.7 + 0.8 * u[3], eps_ratio=min(0.6 + 0.4 * u[4], 0.999))
        spec = bs.make_spec(grid=edge, model=model, nucleation=bs.NUC_UNIFORM_STEPS, bubbles=1 + min(int(u[5] * 4), 3),
                            seed=k + 1, uniform_max_step=200, time_seed=5000 + k)
        t += -math.log(max(1.0 - u[6], 1e-12)) / rate_hz
        jobs.append((t, spec, steps))
    return jobs


def serve(bs, jobs, burst, keep=(), flags=0):
    """Submit the jobs to the continuous-batching server under a 288 GB -
    16 GB budget, at their arrival times or (burst) all at once; returns
    (wall seconds, per-job results, server stats, kept states by job index)."""
    srv = bs.Server(budget_bytes=0, reserve_bytes=16 << 30, chunk_steps=20, flags=flags)
    try:
        t0 = time.perf_counter()
        ids = []
        for k, (ta, spec, steps) in enumerate(jobs):
            now = time.perf_counter() - t0
            if not burst and ta > now:
                time.sleep(ta - now)
            ids.append(srv.submit(spec, steps, keep_state=k in keep))
        res = [srv.wait(j, 600) for j in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
        kept = {k: srv.fetch_state(ids[k], jobs[k][1]) for k in keep if res[k].status == 0}
    finally:
        srv.close()
    return wall, res, st, kept


def served_accuracy(bs, jobs, keep, kept, res, flags=0):
    """Field error and drift of sampled served jobs against the same jobs run
    in f64 ExactMath through a server; the drift of each is taken from the
    job's baseline diagnostics (first tick after its last nucleation) to its
    final ones."""
    def sample_run(precision, make):
        srv = bs.Server(budget_bytes=0, reserve_bytes=16 << 30, chunk_steps=20, precision=precision,
                        math=bs.MATH_EXACT, flags=flags)
        try:
            ids = [srv.submit(make(bs, jobs[k][1]), jobs[k][2], keep_state=True) for k in keep]
            rr = [srv.wait(j, 600) for j in ids]
            st = [srv.fetch_state(j, jobs[k][1]) for j, k in zip(ids, keep)]
        finally:
            srv.close()
        return rr, st

    r64, ref = sample_run(bs.PRECISION_F64, f64_spec)
    r32, exact32 = sample_run(bs.PRECISION_F32, exact32_spec)

    def drift(r):
        return r.final_diag.mean_energy / r.baseline_diag.mean_energy - 1.0 if r.baseline_step else None

    return {"reference": "the same jobs in f64 ExactMath through the server",
            "jobs": [{"index": k, "edge": int(jobs[k][1].grid[0]), "steps": jobs[k][2],
                      "drift_from_step": int(res[k].baseline_step), "energy_drift_f32": drift(res[k]),
                      "energy_drift_f32_exact": drift(x), "energy_drift_f64": drift(r)}
                     for k, r, x in zip(keep, r64, r32)],
            "blocks": field_errors([kept[k] for k in keep], ref), "blocks_f32_exact": field_errors(exact32, ref)}


def served_stream(bs, n_jobs=48, rate_hz=6.0, seed=99, flags=0):
    """The continuous-batching server on a stream of sweep jobs: Poisson
    arrivals at rate_hz (p50/p99 latency and aggregate throughput, which the
    arrival rate bounds), then the same jobs submitted at once (the server's
    capacity), both against sequential plain solves of the same jobs (whole
    jobs from step 0, as the server runs them).  flags: server flags
    (SERVER_NO_GRAPHS steps eagerly, for a run under a profiler)."""
    jobs = served_jobs(bs, n_jobs, rate_hz, seed)
    work = sum(spec.grid[0] * spec.grid[1] * spec.grid[2] * steps for (_, spec, steps) in jobs)
    # a sample for the accuracy check: the first job of each of four edges
    keep, seen = [], set()
    for k, (_, spec, _) in enumerate(jobs):
        e = int(spec.grid[0])
        if e not in seen and len(keep) < 4:
            seen.add(e)
            keep.append(k)
    wall, res, st, kept = serve(bs, jobs, burst=False, keep=keep, flags=flags)
    lat = np.array([r.latency_s for r in res])
    served = work / wall
    bwall, bres, _, _ = serve(bs, jobs, burst=True, flags=flags)
    blat = np.array([r.latency_s for r in bres])
    burst = work / bwall
    ok = all(r.status == 0 for r in res) and all(r.status == 0 for r in bres)
    seq_t, seq_work, service = 0.0, 0, []
    for (_, spec, steps) in jobs:
        sim = bs.Simulation(spec)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.advance(steps)
        sim.synchronize()
        service.append(time.perf_counter() - t1)
        seq_t += service[-1]
        seq_work += sim.cell_count() * steps
        sim.close()
    sequential = seq_work / seq_t
    # the same arrivals served one job at a time, in arrival order, each for
    # its measured plain-solve time: the latencies a sequential server gives
    # at the same offered load
    free, seq_lat = 0.0, []
    for (arrival, _, _), t in zip(jobs, service):
        start = max(arrival, free)
        free = start + t
        seq_lat.append(free - arrival)
    seq_lat = np.array(seq_lat)
    log(f"bench: served {n_jobs} jobs at {rate_hz} Hz: wall {wall:.1f} s, {served / 1e9:.2f} G/s, "
        f"p50 {np.percentile(lat, 50):.2f} s p99 {np.percentile(lat, 99):.2f} s; all at once: wall {bwall:.1f} s, "
        f"{burst / 1e9:.2f} G/s; sequential {sequential / 1e9:.2f} G/s, at the same arrivals one job at a time "
        f"p50 {np.percentile(seq_lat, 50):.2f} s p99 {np.percentile(seq_lat, 99):.2f} s, last done at {free:.1f} s")
    acc = served_accuracy(bs, jobs, keep, kept, res, flags) if len(kept) == len(keep) else None
    if acc:
        log("bench: served accuracy (sample edges " + ", ".join(str(j["edge"]) for j in acc["jobs"]) + "): rel L2 "
            + ", ".join(f"{k} {v['rel_l2']:.2e}" for k, v in acc["blocks"].items()) + "; f32 exact "
            + ", ".join(f"{k} {v['rel_l2']:.2e}" for k, v in acc["blocks_f32_exact"].items()))
    return {"jobs": n_jobs, "arrival_rate_hz": rate_hz, "budget_bytes": int(st.budget_bytes),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "site_updates_per_s": served, "sequential_site_updates_per_s": sequential,
            "sequential_note": SEQUENTIAL_NOTE,
            "speedup_vs_sequential": burst / sequential,
            "speedup_note": "capacity: every job submitted at once (burst) against the same jobs solved one after "
                            "another; at the Poisson arrivals site_updates_per_s is bound by the offered load, and "
                            "the comparison there is latency against sequential_fifo",
            "latency_vs_sequential_fifo": {"p50": float(np.percentile(lat, 50) / np.percentile(seq_lat, 50)),
                                           "p99": float(np.percentile(lat, 99) / np.percentile(seq_lat, 99))},
            "wall_s": wall, "all_ok": ok,
            "graphs": not (flags & bs.SERVER_NO_GRAPHS),
            "sequential_fifo": {"p50_latency_s": float(np.percentile(seq_lat, 50)),
                                "p99_latency_s": float(np.percentile(seq_lat, 99)), "wall_s": free,
                                "note": "the same arrival times, one job at a time in arrival order, each for its "
                                        "measured plain-solve time"},
            "burst": {"wall_s": bwall, "site_updates_per_s": burst, "speedup_vs_sequential": burst / sequential,
                      "p50_latency_s": float(np.percentile(blat, 50)),
                      "p99_latency_s": float(np.percentile(blat, 99))},
            "accuracy": acc}


def copy_calibration(nbytes=1 << 30, repeats=3):
    """Three plain device-to-device copies of a known size (hipMemcpy).

    Run under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE): the counter's
    reading for the copy kernel against the nbytes it moves gives the factor
    that converts the counter into HBM bytes for the stage kernel."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    size = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(a), size) or hip.hipMalloc(ctypes.byref(b), size):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, size)
    for _ in range(repeats):
        if hip.hipMemcpy(b, a, size, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {repeats} device copies of {nbytes} bytes")
    return 0


def copy_bandwidth(nbytes=1 << 30, repeats=5):
    """Achieved HBM bandwidth of the runtime's own device-to-device copy
    (read + write bytes per second), timed with HIP events: the practical
    ceiling the stage kernel's traffic is compared against."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    size = ctypes.c_size_t(nbytes)
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        if hip.hipMalloc(ctypes.byref(a), size) or hip.hipMalloc(ctypes.byref(b), size):
            return None
        hip.hipMemset(a, 1, size)
        hip.hipMemcpy(b, a, size, 3)  # warm
        hip.hipEventCreate(ctypes.byref(ev0))
        hip.hipEventCreate(ctypes.byref(ev1))
        hip.hipEventRecord(ev0, None)
        for _ in range(repeats):
            hip.hipMemcpyAsync(b, a, size, 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(ev1, None)
        hip.hipEventSynchronize(ev1)
        ms = ctypes.c_float()
        if hip.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1) or ms.value <= 0:
            return None
        return 2.0 * nbytes * repeats / (ms.value / 1e3) / 1e9
    finally:
        if ev0:
            hip.hipEventDestroy(ev0)
        if ev1:
            hip.hipEventDestroy(ev1)
        if a:
            hip.hipFree(a)
        if b:
            hip.hipFree(b)


def side_measurement(name, fn, *args):
    """Run one of the bench's side measurements; on failure log it and
    report {"error": ...} in its place."""
    try:
        return fn(*args)
    except Exception as e:  # noqa: BLE001 -- any failure is reported, not raised
        log(f"bench: {name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}


ALG_READ_SHARE = 120.0 / 192.0  # reads among the algorithmic bytes of one step


def hbm_context(bs, achieved, traffic, avg_launch_s):
    """The practical HBM ceilings beside the published peak: streaming read,
    write and copy kernels (bs_hbm_probe) and the runtime's device copy; the
    ceiling those imply for the stage kernel's algorithmic mix of reads and
    writes (62.5 % reads over a step), against which its algorithmic rate --
    a lower bound on its HBM traffic -- is set.  The counter traffic
    (FETCH_SIZE + WRITE_SIZE, scaled) counts every L2 miss, including those
    the last-level cache in front of HBM serves, so it can exceed what HBM
    alone could deliver in the time."""
    probe = bs.hbm_probe()
    out = dict(probe)
    out["device_copy_GBps"] = copy_bandwidth()
    mix = None
    if probe["read_GBps"] > 0 and probe["write_GBps"] > 0:
        mix = 1.0 / (ALG_READ_SHARE / probe["read_GBps"] + (1.0 - ALG_READ_SHARE) / probe["write_GBps"])
    out["algorithmic_read_share"] = ALG_READ_SHARE
    out["algorithmic_mix_ceiling_GBps"] = mix
    out["algorithmic_GBps"] = achieved
    out["algorithmic_frac_of_mix_ceiling"] = achieved / mix if mix else None
    out["l2_miss_traffic_GBps"] = traffic / avg_launch_s / 1e9 if traffic else None
    log(f"bench: HBM streaming read {probe['read_GBps']:.0f}, write {probe['write_GBps']:.0f}, "
        f"copy {probe['copy_GBps']:.0f} GB/s; device copy {out['device_copy_GBps'] or 0:.0f} GB/s"
        + (f"; algorithmic {achieved:.0f} GB/s = {achieved / mix:.2f} of the {mix:.0f} GB/s its read/write mix "
           f"allows" if mix else "")
        + (f"; L2-miss traffic {out['l2_miss_traffic_GBps']:.0f} GB/s" if traffic else ""))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=800,
                    help="timed steps, after the box has been advanced past its last nucleation and the warm-up")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream")
    ap.add_argument("--no-graphs", action="store_true",
                    help="step the batched sweep and the served stream eagerly, without graph capture (for a run "
                         "under a profiler)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the f32-against-f64 field error and drift")
    ap.add_argument("--eos", choices=["bag", "gamma-law"], default="bag",
                    help="equation of state of the stepped box (the headline is the bag model)")
    ap.add_argument("--eos-gamma", type=float, default=1.45, help="gamma-law adiabatic index")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 1 GiB device copies (the reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import bubblesim as bs
    if bs.device_count() == 0:
        print("no HIP device", file=sys.stderr)
        return 1
    n = args.grid
    eos = 1 if args.eos == "gamma-law" else 0
    spec = headline_spec(bs, n, eos, args.eos_gamma if eos else 4.0 / 3.0)
    log(f"bench: {n}^3 f32 fused, eta={spec.model.eta:.4f}, device {bs.device_name()}")
    sim = bs.Simulation(spec)
    cells = sim.cell_count()
    settle = settle_steps(bs, sim)
    start = settle + args.warmup  # first timed step
    sim.advance(start)
    sim.synchronize()
    log(f"bench: advanced past the last nucleation ({settle} steps) and {args.warmup} warm-up steps")
    # whole-job throughput, no per-launch events
    t0 = time.perf_counter()
    sim.advance(args.steps)
    sim.synchronize()
    wall_s = time.perf_counter() - t0
    d = sim.diagnostics()
    # the same window again from the same initial state, with HIP events
    # around every stage launch on the solver's stream (the kernel's cost
    # depends on the state through the clock the power limit allows)
    sim.reset()
    sim.advance(start)
    kernel_ms, _ = sim.advance_timed(args.steps)
    # the CPU baseline steps this box from here (the state at the window's end)
    cpu_state = None
    if not args.no_cpu:
        try:
            cpu_state = sim.download()
        except Exception as e:  # noqa: BLE001 -- the baseline then reports the failure
            log(f"bench: the CPU baseline's state could not be read back: {type(e).__name__}: {e}")
    log(f"bench: steps {start}..{start + args.steps}: wall {wall_s * 1e3:.1f} ms, stage kernels "
        f"{kernel_ms:.1f} ms (second pass); broken {d.broken_fraction:.4f} max|v| {d.max_velocity:.4f} "
        f"nonfinite {d.nonfinite}")

    def window_rate(first, count):
        sim.reset()
        sim.advance(first)
        sim.synchronize()
        t1 = time.perf_counter()
        sim.advance(count)
        sim.synchronize()
        return cells * count / (time.perf_counter() - t1)

    # for scale: the first steps of the run, while the box is still nearly
    # uniform, and steps after the transition, the box fully broken
    n_scale = 30
    early = window_rate(args.warmup, n_scale)
    late_first = max(1000, start + args.steps)
    late = window_rate(late_first, n_scale)
    # and the whole transition, from the first steps through nucleation to
    # the fully broken box (the power limit lowers the clock as it fills)
    transition = window_rate(args.warmup, 1000)
    log(f"bench: steps {args.warmup}..{args.warmup + n_scale}: {early / 1e9:.2f} G/s; "
        f"steps {late_first}..{late_first + n_scale}: {late / 1e9:.2f} G/s; "
        f"steps {args.warmup}..{args.warmup + 1000}: {transition / 1e9:.2f} G/s")
    value = cells * args.steps / wall_s
    launches = 3 * args.steps
    avg_launch_s = kernel_ms / 1e3 / launches
    bytes_per_launch = BYTES_PER_SITE_UPDATE / 3 * cells
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_fused_stage_512.json")
    if os.path.exists(tfile) and n == 512 and not eos:  # measured on the headline workload only
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    sim.close()
    # the side measurements never take the headline line down with them
    hbm = side_measurement("HBM context", hbm_context, bs, achieved, traffic, avg_launch_s)
    acc = None if args.no_accuracy else side_measurement(
        "accuracy", accuracy, bs, n, eos, args.eos_gamma if eos else 4.0 / 3.0)
    batch = None if args.no_batch else side_measurement("batched sweep", batched_sweep, bs, 1000, 200,
                                                        not args.no_graphs)
    served = None if args.no_serve else side_measurement(
        "served stream", served_stream, bs, 48, 6.0, 99, bs.SERVER_NO_GRAPHS if args.no_graphs else 0)
    cpu = None if args.no_cpu else side_measurement("CPU baseline", cpu_baseline, bs, cpu_state)
    cpu_state = None
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
                               f"eta log-U[0.02,1] draw = {spec.model.eta:.4f}, SSP-RK3, fused stage kernel; "
                               f"advanced untimed past its last nucleation ({settle} steps) and {args.warmup} "
                               f"warm-up steps, then steps {start}..{start + args.steps} timed",
                   "grid": n, "bubbles": 32, "precision": "f32", "math": "fast",
                   "timed_steps": [start, start + args.steps],
                   "eos": args.eos if not eos else f"gamma-law {args.eos_gamma}"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "stage_kernel_ms_per_launch": avg_launch_s * 1e3,
        "first_steps": {"steps": [args.warmup, args.warmup + n_scale], "site_updates_per_s": early},
        "late_steps": {"steps": [late_first, late_first + n_scale], "site_updates_per_s": late},
        "transition_steps": {"steps": [args.warmup, args.warmup + 1000], "site_updates_per_s": transition,
                             "frac": transition * BYTES_PER_SITE_UPDATE / 1e9 / HBM_PEAK_GBS},
        "hbm_context": hbm,
        "batched_sweep": batch,
        "served": served,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
