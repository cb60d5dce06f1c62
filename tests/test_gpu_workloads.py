"""The batched sweep and the served stream at their benchmark workloads
(bench.py's sweep_specs and served_jobs: the same boxes and jobs the bench
line measures).

* Batched sweep, 64 boxes of edge 32..128: the ExactMath batch, advanced
  through nucleation (every bubble is due by step 200) by graph replay,
  equals one plain Simulation per box bit for bit; the FastMath batch stays
  within 1e-5 relative L2 per block of the ExactMath batch, box by box.
* Boxes narrower than the tile (x half slopes shared between lanes across
  the periodic wrap): FastMath within the bar of f64 for every fast tile.
* Served stream, 48 jobs of edge 48..192 and 500..2000 steps: every job's
  final state checksum is the same run alone through the server and inside
  the mix at Poisson arrival times; sampled jobs also equal a plain
  Simulation of the same spec (bit for bit under ExactMath, within the 1e-5
  bar under FastMath).
"""
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def _rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-300)))


def _with_math(bs, spec, math):
    s = bs.CSpec.from_buffer_copy(spec)
    s.math = math
    return s


SWEEP_STEPS = 225  # the last bubble of the sweep is stamped at step 200 at the latest


def test_sweep_batch_exact_equals_solo_boxes(bs, gpu_available):
    specs = [_with_math(bs, s, bs.MATH_EXACT) for s in bench.sweep_specs(bs)]
    b = bs.Batch(specs)
    try:
        b.advance(SWEEP_STEPS, use_graph=True)
        assert b.graph_steps() == SWEEP_STEPS
        _log(f"batch plan {b.plan()}")
        for i, s in enumerate(specs):
            _log(f"box {i} grid {tuple(s.grid)} eta {s.model.eta:.4f}")
            assert b.steps(i) == SWEEP_STEPS
            sim = bs.Simulation(s)
            sim.advance(SWEEP_STEPS)
            assert sim.remaining_bubbles() == 0
            sf, su = sim.download()
            sim.close()
            bf, bu = b.download(i)
            assert np.array_equal(sf, bf) and np.array_equal(su, bu), f"box {i} differs from its solo run"
    finally:
        b.close()


def test_sweep_batch_fast_within_bar_of_exact(bs, gpu_available):
    base = bench.sweep_specs(bs)
    fast = bs.Batch(base)
    exact = bs.Batch([_with_math(bs, s, bs.MATH_EXACT) for s in base])
    try:
        fast.advance(SWEEP_STEPS, use_graph=True)
        exact.advance(SWEEP_STEPS, use_graph=True)
        _log(f"fast plan {fast.plan()}")
        worst = 0.0
        for i in range(len(base)):
            ff, fu = fast.download(i)
            ef, eu = exact.download(i)
            e = max(_rel_l2(ff, ef), _rel_l2(fu, eu))
            worst = max(worst, e)
            assert e < 1e-5, f"box {i} grid {tuple(base[i].grid)}: fast vs exact rel L2 {e:.3e}"
            assert fast.diagnostics(i).nonfinite == 0
        _log(f"worst per-block rel L2 of fast against exact over {len(base)} boxes: {worst:.3e}")
    finally:
        fast.close()
        exact.close()


FAST_TILES = ["32x16", "64x8", "32x8", "16x16", "16x8", "16x15"]


@pytest.mark.parametrize("tile", FAST_TILES)
def test_fast_boxes_narrower_than_the_tile(bs, gpu_available, monkeypatch, tile):
    # nx below the tile width: lanes past the box wrap onto its cells, and the
    # lane on a cell's right takes its x half slope across the periodic seam
    monkeypatch.setenv("BS_TILE", tile)
    tx = int(tile.split("x")[0])
    for nx in sorted({8, 12, max(8, tx // 2 + 3), max(8, tx - 5)}):
        if nx >= tx:
            continue
        grid = (nx, 20, 18)
        model = bs.Model(seed_cells=3.0)
        r = model.seed_radius()
        sched = np.array([(0.0, nx / 2, 10.0, 9.0, r), (0.4, 1.0, 3.0, 15.0, r)], dtype=np.float32)
        out = {}
        for name, prec, math in (("fast", 0, bs.MATH_FAST), ("f64", 1, bs.MATH_EXACT)):
            sim = bs.Simulation(bs.make_spec(grid=grid, model=model, precision=prec, math=math,
                                             nucleation=bs.NUC_SIMULTANEOUS, bubbles=1))
            sim.reset(sched)
            sim.advance(20)
            out[name] = sim.download()
            sim.close()
        for k in range(2):
            e = _rel_l2(out["fast"][k], out["f64"][k])
            _log(f"tile {tile} grid {grid} block {k}: fast vs f64 rel L2 {e:.3e}")
            assert e < 1e-5


def test_served_stream_alone_equals_mixed(bs, gpu_available):
    jobs = bench.served_jobs(bs, 48, 6.0, 99)
    srv = bs.Server(budget_bytes=0, reserve_bytes=16 << 30, chunk_steps=20)
    try:
        alone = []
        for k, (_, spec, steps) in enumerate(jobs):
            r = srv.wait(srv.submit(spec, steps, checksum=True), 600)
            assert r.status == bs.BS_OK and r.steps == steps
            alone.append(r.state_checksum)
        _log(f"{len(jobs)} jobs run alone")
        t0 = time.perf_counter()
        ids = []
        for (ta, spec, steps) in jobs:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            ids.append(srv.submit(spec, steps, checksum=True))
        spans = []
        for k, j in enumerate(ids):
            r = srv.wait(j, 600)
            spans.append((r.admit_s, r.finish_s))
            assert r.status == bs.BS_OK and r.steps == jobs[k][2]
            assert r.state_checksum == alone[k], f"job {k} (edge {jobs[k][1].grid[0]}) differs in the mix"
        # continuous batching: jobs admitted while others were on the device
        overlaps = sum(1 for k in range(1, len(spans)) if spans[k][0] < max(f for _, f in spans[:k]))
        _log(f"{overlaps} of {len(spans) - 1} jobs admitted while an earlier one was still running")
        assert overlaps > 0
        st = srv.stats()
        _log(f"mix done: {st.completed} completed, {st.ticks} ticks")
        assert st.completed == 2 * len(jobs) and st.failed == 0
        # a few of the jobs against a plain Simulation of the same spec:
        # within the 1e-5 bar under FastMath (the batched and single-box
        # kernels are separate instantiations, whose multiply-adds the
        # compiler may contract differently, and over 500-2000 steps the
        # rounding differences grow to the size of FastMath's own distance
        # to f64: 1.9e-6 for a 128^3 job over 1343 steps), bit for bit
        # under ExactMath
        for k in (0, 1, 2):
            spec, steps = jobs[k][1], jobs[k][2]
            j = srv.submit(spec, steps, keep_state=True)
            assert srv.wait(j, 600).status == bs.BS_OK
            f, u = srv.fetch_state(j, spec)
            sim = bs.Simulation(spec)
            sim.advance(steps)
            sf, su = sim.download()
            sim.close()
            e = max(_rel_l2(f, sf), _rel_l2(u, su))
            _log(f"job {k} (edge {spec.grid[0]}, {steps} steps): served vs plain Simulation rel L2 {e:.3e}")
            assert e < 1e-5
    finally:
        srv.close()
    exact = bs.Server(budget_bytes=0, reserve_bytes=16 << 30, chunk_steps=20, math=bs.MATH_EXACT)
    try:
        for k in (0, 1, 2):
            spec, steps = _with_math(bs, jobs[k][1], bs.MATH_EXACT), jobs[k][2]
            j = exact.submit(spec, steps, keep_state=True)
            assert exact.wait(j, 600).status == bs.BS_OK
            f, u = exact.fetch_state(j, spec)
            sim = bs.Simulation(spec)
            sim.advance(steps)
            sf, su = sim.download()
            sim.close()
            assert np.array_equal(f, sf) and np.array_equal(u, su), f"exact job {k} differs from its plain Simulation"
    finally:
        exact.close()
