"""The reference's published headless runs, re-run on the GPU solver.

The only numbers the reference publishes for its solver (docs/NUMERICS.md,
docs/PHYSICS.md, transcribed in tests/golden/published_runs.json) come from
its f32 wgpu solver on an RTX 3090 with naive f32 reductions.  The runs use
the reference's seeded schedule (seed 1), which this build reproduces bit
for bit, so the same bubble is placed at the same cell.  Field values are
not published, so the comparison is on the diagnostics table at tolerances
that cover the reference's own f32 reduction error (~1e-4 relative on sums)
and the implementation-defined rounding of its driver (FMA contraction,
2.5-ulp division and square root).
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "published_runs.json")
PUB = json.load(open(GOLDEN))


def r_eq(broken_fraction, cells):
    # headless.rs:432-434
    return (np.float32(broken_fraction) * np.float32(cells) * np.float32(3.0) / np.float32(4.0 * math.pi)) ** (1 / 3)


def headless_table(bs, spec, steps, every):
    """The rows headless::run prints (headless.rs:345-377)."""
    sim = bs.Simulation(spec)
    rows, base, done = [], None, 0
    while done < steps:
        chunk = min(every, steps - done)  # headless.rs:346
        sim.advance(chunk)
        done += chunk
        d = sim.diagnostics()
        assert d.nonfinite == 0
        if base is None and sim.remaining_bubbles() == 0:
            base = d.mean_energy
        rows.append(dict(step=done, time=sim.time, broken_pct=100 * d.broken_fraction,
                         r_eq=float(r_eq(d.broken_fraction, sim.cell_count())), max_v=d.max_velocity,
                         k=d.kinetic_fraction, e_ratio=(d.mean_energy / base) if base else None))
    sim.close()
    return rows


def test_single_bubble_128_trajectory(bs, gpu_available):
    # docs/NUMERICS.md:179-189, reproduced to the printed digits with seed
    # factor 1.6 (the documented run predates the 1.15 default).  E/E0 in its
    # sixth decimal depends on how an f32 implementation rounds the RK stage
    # combination a0 U^n + a1 (...): fused into a multiply-add, as the
    # reference's driver evidently compiles it (a build of this solver that
    # fused it met every published E/E0 within 2e-6), the energy takes a
    # bias that grows over the run, 8e-6 by step 1000 and 1.3e-5 by 1500;
    # this build keeps the unfused combination, the reference's source
    # order, whose f32 state stays within 1e-5 of f64's over 1500 steps
    # where the fused one's does not.  Hence 1.5e-5 on the published E/E0;
    # the f64 scheme's own E/E0 is logged beside it.
    ref = PUB["single_bubble_128"]
    tables = {}
    for name, prec, math in (("fast", 0, bs.MATH_FAST), ("exact", 0, bs.MATH_EXACT), ("f64", 1, bs.MATH_EXACT)):
        spec = bs.make_spec(grid=128, model=bs.Model(eta=0.3, seed_factor=ref["seed_factor"]),
                            nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=1, precision=prec, math=math)
        tables[name] = {r["step"]: r for r in headless_table(bs, spec, 1500, 100)}
    bad = []
    for want in ref["rows"]:
        got = tables["fast"][want["step"]]
        ex = tables["exact"][want["step"]]
        print(f"step {want['step']}: broken {got['broken_pct']:.2f}% ({want['broken_pct']}), "
              f"R_eq {got['r_eq']:.1f} ({want['r_eq']}), max|v| {got['max_v']:.4f} ({want['max_v']}), "
              f"K {got['k']:.3e} ({want['k']}), E/E0 {got['e_ratio']:.6f} (reference order {ex['e_ratio']:.6f}; "
              f"f64 {tables['f64'][want['step']]['e_ratio']:.6f}; published {want['e_ratio']})", flush=True)
        assert f"{got['time']:.1f}" == f"{want['time']:.1f}"
        assert f"{got['broken_pct']:.2f}" == f"{want['broken_pct']:.2f}"
        assert f"{got['r_eq']:.1f}" == f"{want['r_eq']:.1f}"
        assert f"{got['max_v']:.4f}" == f"{want['max_v']:.4f}"
        assert f"{got['k']:.3e}" == f"{want['k']:.3e}"
        if abs(got["e_ratio"] - want["e_ratio"]) > 1.5e-5:
            bad.append((want["step"], got["e_ratio"], want["e_ratio"]))
    assert not bad, bad


def test_runaway_wall_128(bs, gpu_available):
    # docs/NUMERICS.md:204-211: "Same setup, eta = 0.02: R_eq goes 29.9 ->
    # 45.7 -> 63.3 over equal intervals, a wall speed near 0.85 rather than
    # 0.4.  Energy drift grows to -4.5e-3."  The same setup is the single
    # bubble run above with its seed factor 1.6; R_eq is compared at the
    # printed digit, the wall speed at the printed figure, and the drift
    # (printed to one figure, without a step) where the box has just broken
    # and at the end of the run.
    ref = PUB["runaway_wall_128"]
    spec = bs.make_spec(grid=128, model=bs.Model(eta=ref["eta"], seed_factor=ref["seed_factor"]),
                        nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=1)
    rows = headless_table(bs, spec, ref["steps"], ref["report_every"])
    by_step = {r["step"]: r for r in rows}
    for want in ref["r_eq_rows"]:
        got = by_step[want["step"]]
        print(f"step {want['step']}: R_eq {got['r_eq']:.1f} ({want['r_eq']}), max|v| {got['max_v']:.4f}", flush=True)
        assert f"{got['r_eq']:.1f}" == f"{want['r_eq']:.1f}"
    speeds = [(by_step[b]["r_eq"] - by_step[a]["r_eq"]) / (by_step[b]["time"] - by_step[a]["time"])
              for a, b in ((100, 200), (200, 300))]
    print(f"wall speeds {speeds[0]:.3f}, {speeds[1]:.3f} (published near {ref['wall_speed']})", flush=True)
    assert all(0.75 <= v <= 0.95 for v in speeds)
    broken = next(r for r in rows if r["broken_pct"] >= 99.5)
    drift_broken = broken["e_ratio"] - 1
    drift_end = rows[-1]["e_ratio"] - 1
    print(f"drift {drift_broken:.3e} at step {broken['step']} (box broken), {drift_end:.3e} at step {rows[-1]['step']} "
          f"(published {ref['drift']:.1e})", flush=True)
    assert f"{drift_broken:.1e}" == f"{ref['drift']:.1e}"
    assert -5.0e-3 < drift_end < drift_broken


def test_courant_study_96(bs, gpu_available):
    # docs/NUMERICS.md:213-222: 96^3, one bubble, eta = 0.02, to t = 450;
    # drifts -8.2e-4 (cfl 0.2, 2250 steps) and -1.06e-2 (cfl 0.5, 900 steps),
    # "0.5 is 13x less accurate".  The published drifts take E0 at the first
    # of five reports (t = 90, just after the box has broken, where the drift
    # falls steeply with the E0 time; the fixture's notes give the curves):
    # reproduced to 13 % and 6 %, the ratio to 12x.
    ref = PUB["courant_96"]
    drift = {}
    for row in ref["rows"]:
        spec = bs.make_spec(grid=96, model=bs.Model(eta=0.02, cfl=row["cfl"]), nucleation=bs.NUC_SIMULTANEOUS,
                            bubbles=1, seed=1)
        rows = headless_table(bs, spec, row["steps"], row["steps"] // ref["reports"])
        drift[row["cfl"]] = rows[-1]["e_ratio"] - 1
        print(f"cfl {row['cfl']}: E0 at t = {rows[0]['time']:.1f}, t = {rows[-1]['time']:.1f}, "
              f"drift {drift[row['cfl']]:.3e} (published {row['drift']:.2e})", flush=True)
        assert rows[-1]["time"] == pytest.approx(450.0, abs=0.5)
        assert rows[0]["time"] == pytest.approx(90.0, abs=0.5)
        assert abs(drift[row["cfl"]] / row["drift"] - 1.0) < 0.2
    ratio = drift[0.5] / drift[0.2]
    print(f"cfl 0.5 / cfl 0.2 drift: {ratio:.1f}x (published 13x)", flush=True)
    assert 10.0 < ratio < 16.0


@pytest.mark.parametrize("row", PUB["seed_threshold_128"]["rows"], ids=lambda r: f"eps{r['eps_ratio']}")
def test_seed_threshold_table(bs, gpu_available, row):
    factors = PUB["seed_threshold_128"]["factors"]
    for factor, want in zip(factors, row["broken_pct"]):
        spec = bs.make_spec(grid=128, model=bs.Model(eps_ratio=row["eps_ratio"], seed_factor=factor),
                            nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=1)
        sim = bs.Simulation(spec)
        sim.advance(700)
        got = 100 * sim.diagnostics().broken_fraction
        sim.close()
        print(f"eps_ratio {row['eps_ratio']} factor {factor}: broken {got:.2f}% (published {want})", flush=True)
        if want is None:
            assert got < 0.05  # "collapsed"
        else:
            assert f"{got:.1f}" == f"{want:.1f}"  # the published precision
