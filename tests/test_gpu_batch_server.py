"""Batched sweep and served jobs: every box gives the bits it gives alone.

* Batch (one fused launch per RK stage across uneven boxes, device-driven
  nucleation, hipGraph replay) against one Simulation per box: bitwise under
  ExactMath (the per-box arithmetic is the same sequence), and graph replay
  bitwise equal to eager stepping under FastMath.
* Server (continuous batching under an HBM budget): each job's final state
  checksum is identical when it runs alone and inside a mix of concurrently
  admitted jobs; ExactMath jobs equal a plain Simulation bitwise.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(bs, math, n=6, seed=3):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        edge = int(rng.choice([16, 24, 32, 40]))
        grid = (edge, int(rng.choice([16, 24, 32])), int(rng.choice([16, 20, 32])))
        eta = float(np.exp(np.log(0.02) + rng.uniform() * (0 - np.log(0.02))))
        model = bs.Model(eta=eta, seed_factor=float(rng.uniform(0.7, 1.5)), eps_ratio=float(rng.uniform(0.6, 0.99)))
        specs.append(bs.make_spec(grid=grid, model=model, nucleation=bs.NUC_UNIFORM_STEPS,
                                  bubbles=int(rng.integers(1, 5)), seed=k + 1, uniform_max_step=8, time_seed=100 + k,
                                  math=math))
    return specs


def test_batch_equals_single_boxes_exact(bs, gpu_available):
    # 6 eager steps (every event due by step 8 is stamped in them or in the
    # graph), then 30 steps: one 25-step graph replayed, 5 eager
    specs = _specs(bs, bs.MATH_EXACT)
    b = bs.Batch(specs)
    b.advance(6, use_graph=False)
    assert b.graph_steps() == 0
    b.advance(30, use_graph=True)
    assert b.graph_steps() == 25
    for i, s in enumerate(specs):
        print(f"box {i} grid {tuple(s.grid)}", flush=True)
        sim = bs.Simulation(s)
        sim.advance(36)
        f, u = sim.download()
        bf, bu = b.download(i)
        assert b.steps(i) == 36
        assert np.array_equal(f, bf) and np.array_equal(u, bu)
        d1, d2 = sim.diagnostics(), b.diagnostics(i)
        assert d1.mean_energy == d2.mean_energy and d1.max_velocity == d2.max_velocity
        sim.close()
    b.close()


def test_batch_graph_equals_eager_fast(bs, gpu_available):
    specs = _specs(bs, bs.MATH_FAST, n=5, seed=9)
    a = bs.Batch(specs)
    a.advance(60, use_graph=True)
    assert a.graph_steps() == 50
    c = bs.Batch(specs)
    c.advance(60, use_graph=False)
    assert c.graph_steps() == 0
    for i in range(len(specs)):
        fa, ua = a.download(i)
        fc, uc = c.download(i)
        assert np.array_equal(fa, fc) and np.array_equal(ua, uc)
        assert np.isfinite(ua).all()
    a.close()
    c.close()


def test_batch_skips_the_stamp_only_when_nothing_is_due(bs, gpu_available):
    # events spread over steps 0..90: of the four 25-step graphs some carry the
    # stamp and some do not (each variant is captured once and replayed), the
    # eager remainder stamps step by step; every box still equals its solo run
    rng = np.random.default_rng(21)
    specs = []
    for k in range(4):
        grid = (int(rng.choice([16, 24, 32])), 24, int(rng.choice([16, 20])))
        specs.append(bs.make_spec(grid=grid, model=bs.Model(eta=0.1, seed_factor=1.2), nucleation=bs.NUC_UNIFORM_STEPS,
                                  bubbles=2, seed=k + 7, uniform_max_step=90, time_seed=40 + k, math=bs.MATH_EXACT))
    b = bs.Batch(specs)
    b.advance(110, use_graph=True)
    assert b.graph_steps() == 100
    print(f"stamps skipped: {b.stamps_skipped()} of 110 steps", flush=True)
    for i, s in enumerate(specs):
        sim = bs.Simulation(s)
        sim.advance(110)
        f, u = sim.download()
        bf, bu = b.download(i)
        assert b.steps(i) == 110
        assert np.array_equal(f, bf) and np.array_equal(u, bu), f"box {i}"
        sim.close()
    b.close()
    # simultaneous nucleation: every event is due at step 0, so the first
    # graph stamps and the second graph and the 10 eager steps do not
    specs = [bs.make_spec(grid=(24, 24, 16), nucleation=bs.NUC_SIMULTANEOUS, bubbles=3, seed=k + 1, math=bs.MATH_EXACT)
             for k in range(3)]
    b = bs.Batch(specs)
    b.advance(60, use_graph=True)
    assert b.graph_steps() == 50
    assert b.stamps_skipped() == 35
    b.advance(5, use_graph=False)
    assert b.stamps_skipped() == 40
    for i, s in enumerate(specs):
        sim = bs.Simulation(s)
        sim.advance(65)
        ref = sim.download()
        got = b.download(i)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), f"box {i}"
        sim.close()
    b.close()


def test_zero_and_short_runs(bs, gpu_available):
    # zero steps leave every path at the initialised lattice (the first
    # step's nucleation stamp included: nothing is stamped before a step
    # runs), runs shorter than a graph chunk or a server tick step eagerly
    # to exactly their count, and a batch needs at least one box
    specs = [bs.make_spec(grid=24, nucleation=bs.NUC_SIMULTANEOUS, bubbles=2, seed=3, math=bs.MATH_EXACT),
             bs.make_spec(grid=(16, 24, 20), nucleation=bs.NUC_UNIFORM_STEPS, bubbles=2, seed=4, uniform_max_step=2,
                          time_seed=9, math=bs.MATH_EXACT)]
    ref0, ref3 = [], []
    for s in specs:
        sim = bs.Simulation(s)
        init = sim.download()
        sim.advance(0)
        assert sim.steps == 0
        got = sim.download()
        assert np.array_equal(init[0], got[0]) and np.array_equal(init[1], got[1])
        ref0.append(init)
        sim.advance(3)
        ref3.append(sim.download())
        sim.close()
    b = bs.Batch(specs)
    b.advance(0, use_graph=True)
    for i in range(len(specs)):
        assert b.steps(i) == 0
        f, u = b.download(i)
        assert np.array_equal(f, ref0[i][0]) and np.array_equal(u, ref0[i][1]), f"box {i} after 0 steps"
    b.advance(3, use_graph=True)
    assert b.graph_steps() == 0
    for i in range(len(specs)):
        assert b.steps(i) == 3
        f, u = b.download(i)
        assert np.array_equal(f, ref3[i][0]) and np.array_equal(u, ref3[i][1]), f"box {i} after 3 steps"
    b.close()
    srv = bs.Server(budget_bytes=1 << 30, chunk_steps=5, math=bs.MATH_EXACT)
    ids = [(srv.submit(s, n, keep_state=True), s, n) for s in specs for n in (0, 3)]
    for j, s, n in ids:
        r = srv.wait(j, 120)
        assert r.status == 0 and r.steps == n
        f, u = srv.fetch_state(j, s)
        ref = (ref0 if n == 0 else ref3)[specs.index(s)]
        assert np.array_equal(f, ref[0]) and np.array_equal(u, ref[1]), f"job of {n} steps"
    srv.close()
    with pytest.raises(bs.BubbleSimError) as e:
        bs.Batch([])
    assert e.value.code == bs.BS_ERR_INVALID


def _wait_all(srv, ids, timeout=300):
    return {j: srv.wait(j, timeout) for j in ids}


def test_server_alone_equals_mixed(bs, gpu_available):
    specs = _specs(bs, bs.MATH_FAST, n=6, seed=5)
    steps = [30, 45, 20, 60, 35, 50]
    srv = bs.Server(budget_bytes=2 << 30, chunk_steps=5)
    alone = {}
    for k, (s, n) in enumerate(zip(specs, steps)):
        r = srv.wait(srv.submit(s, n, checksum=True), 300)
        assert r.status == 0 and r.steps == n
        alone[k] = r.state_checksum
    ids = [srv.submit(s, n, checksum=True) for s, n in zip(specs, steps)]
    res = _wait_all(srv, ids)
    for k, j in enumerate(ids):
        assert res[j].status == 0 and res[j].steps == steps[k]
        assert res[j].state_checksum == alone[k], f"job {k} differs in the mix"
    st = srv.stats()
    assert st.completed == 12 and st.rejected == 0
    srv.close()


def test_server_size_classes(bs, gpu_available):
    # jobs below 128^3 cells step in a class of their own, more steps per
    # tick while large jobs run: with equal step counts the small jobs of a
    # mix finish before the large one (in one lockstep set all four would
    # finish in the same tick), and every job still gives its solo bits
    small = [bs.make_spec(grid=32, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=k + 1) for k in range(3)]
    large = bs.make_spec(grid=128, nucleation=bs.NUC_SIMULTANEOUS, bubbles=2, seed=9)
    steps = 400
    srv = bs.Server(budget_bytes=4 << 30, chunk_steps=10)
    alone = [srv.wait(srv.submit(s, steps, checksum=True), 300).state_checksum for s in small + [large]]
    ids = [srv.submit(s, steps, checksum=True) for s in [large] + small]
    res = _wait_all(srv, ids)
    assert all(res[j].status == 0 and res[j].steps == steps for j in ids)
    assert [res[j].state_checksum for j in ids] == [alone[3]] + alone[:3]
    t_large = res[ids[0]].finish_s
    t_small = max(res[j].finish_s for j in ids[1:])
    print(f"small jobs done at {t_small - res[ids[0]].submit_s:.3f} s, large at {t_large - res[ids[0]].submit_s:.3f} s",
          flush=True)
    assert t_small < t_large
    srv.close()


def test_server_without_graphs_equals_graphs(bs, gpu_available):
    # SERVER_NO_GRAPHS (eager stepping, for runs under a profiler) launches the
    # same kernels as the captured graphs: the same bits per job
    specs = _specs(bs, bs.MATH_FAST, n=6, seed=7)
    steps = [30, 45, 20, 60, 35, 50]
    sums = []
    for flags in (0, bs.SERVER_NO_GRAPHS):
        srv = bs.Server(budget_bytes=2 << 30, chunk_steps=5, flags=flags)
        ids = [srv.submit(s, n, checksum=True) for s, n in zip(specs, steps)]
        res = _wait_all(srv, ids)
        assert all(res[j].status == 0 for j in ids)
        sums.append([res[j].state_checksum for j in ids])
        srv.close()
    assert sums[0] == sums[1]


def test_server_exact_equals_simulation(bs, gpu_available):
    specs = _specs(bs, bs.MATH_EXACT, n=3, seed=11)
    srv = bs.Server(budget_bytes=1 << 30, chunk_steps=4, math=bs.MATH_EXACT)
    ids = [srv.submit(s, 13, keep_state=True) for s in specs]
    res = _wait_all(srv, ids)
    for s, j in zip(specs, ids):
        assert res[j].status == 0
        f, u = srv.fetch_state(j, s)
        sim = bs.Simulation(s)
        sim.advance(13)
        sf, su = sim.download()
        assert np.array_equal(f, sf) and np.array_equal(u, su)
        sim.close()
    srv.close()


def test_server_admission_and_rejection(bs, gpu_available):
    s32 = bs.make_spec(grid=32, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
    one = bs.lattice_bytes((32, 32, 32))
    # room for exactly one 32^3 job at a time: jobs run one after another
    srv = bs.Server(budget_bytes=one + one // 2, chunk_steps=5)
    ids = [srv.submit(s32, 20) for _ in range(3)]
    res = _wait_all(srv, ids)
    fin = sorted((res[j].admit_s, res[j].finish_s) for j in ids)
    for (a0, f0), (a1, f1) in zip(fin, fin[1:]):
        assert a1 >= f0 - 1e-3  # admitted only after the previous one freed its memory
    with pytest.raises(bs.BubbleSimError) as e:
        srv.submit(bs.make_spec(grid=64), 10)
    assert e.value.code == bs.BS_ERR_OOM
    with pytest.raises(bs.BubbleSimError) as e:
        srv.submit(bs.make_spec(grid=32, model=bs.Model(cfl=0.9)), 10)
    assert e.value.code == bs.BS_ERR_INVALID
    assert srv.stats().rejected == 2
    srv.close()


def test_server_overlaps_jobs(bs, gpu_available):
    # continuous batching: a job submitted while another runs is admitted and
    # advanced in the same launches, not queued behind it
    srv = bs.Server(budget_bytes=4 << 30, chunk_steps=5)
    # a job long enough (192^3, 300 steps: about 0.1 s) to be running when
    # the second one arrives
    big = srv.submit(bs.make_spec(grid=192, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1), 300)
    time.sleep(0.02)
    small = srv.submit(bs.make_spec(grid=32, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1), 40)
    rs = srv.wait(small, 300)
    rb = srv.wait(big, 300)
    assert rs.finish_s < rb.finish_s
    assert rs.admit_s < rb.finish_s
    srv.close()
