"""Failure detection and containment, with the failure forced from the host
(a non-finite value written into a state through the upload entry points;
NaN arithmetic on the GPU traps nothing, so every kernel exits normally):

* the diagnostics reduction flags a non-finite state (`nonfinite`), which is
  what the headless driver aborts on (headless.rs:63-67), and the stepping
  itself keeps going without error;
* in a batch, a box driven non-finite leaves every other box bit for bit on
  the trajectory it follows alone (tasks never span boxes), graph replay
  included, while its own diagnostics report the failure;
* in the server, a job whose state goes non-finite is evicted at the tick
  that finds it, with BS_ERR_NONFINITE, and every other job of the mix ends
  with the state checksum it has alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def test_nonfinite_state_is_flagged(bs, gpu_available):
    spec = bs.make_spec(grid=32, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
    sim = bs.Simulation(spec)
    try:
        sim.advance(2)
        assert sim.diagnostics().nonfinite == 0
        f, u = sim.download()
        u[5, 6, 7, 1] = np.nan
        sim.upload(f, u)
        assert sim.diagnostics().nonfinite == 1
        sim.advance(3)  # the non-finite value spreads; nothing traps
        d = sim.diagnostics()
        assert d.nonfinite == 1
        _, u2 = sim.download()
        assert np.count_nonzero(~np.isfinite(u2)) > 1
    finally:
        sim.close()


def test_batch_contains_a_nonfinite_box(bs, gpu_available):
    specs = [bs.make_spec(grid=g, math=bs.MATH_EXACT, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=k + 1)
             for k, g in enumerate([(32, 24, 20), (40, 32, 24), (24, 24, 32)])]
    b = bs.Batch(specs)
    try:
        b.advance(2, use_graph=False)
        f, u = b.download(1)
        u[3, 4, 5, 0] = np.inf
        f[10, 11, 12, 0] = np.nan
        b.upload(1, f, u)
        b.advance(28, use_graph=True)  # one 25-step graph replayed, 3 eager steps
        assert b.graph_steps() == 25
        assert b.diagnostics(1).nonfinite == 1
        for i in (0, 2):
            _log(f"box {i} grid {tuple(specs[i].grid)} against its solo run")
            sim = bs.Simulation(specs[i])
            sim.advance(30)
            sf, su = sim.download()
            sim.close()
            bf, bu = b.download(i)
            assert b.steps(i) == 30
            assert np.array_equal(sf, bf) and np.array_equal(su, bu)
            assert b.diagnostics(i).nonfinite == 0
    finally:
        b.close()


def test_server_evicts_a_nonfinite_job(bs, gpu_available):
    # the 128^3 job steps in the large size class, the others and the
    # poisoned one in the small class: the eviction happens with both stepping
    specs = [bs.make_spec(grid=g, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=k + 1)
             for k, g in enumerate([(32, 32, 32), (48, 32, 40), (40, 40, 24), (128, 128, 128)])]
    steps = [60, 45, 70, 40]
    srv = bs.Server(budget_bytes=1 << 30, chunk_steps=5)
    try:
        alone = []
        for s, n in zip(specs, steps):
            r = srv.wait(srv.submit(s, n, checksum=True), 300)
            assert r.status == bs.BS_OK
            alone.append(r.state_checksum)
        # the poisoned job starts from the initial state with one NaN in it
        poison = bs.make_spec(grid=(32, 24, 24), nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, seed=9)
        e_ref = bs.Model().derived().e_ref
        field = np.zeros((24, 24, 32, 2), np.float32)
        fluid = np.zeros((24, 24, 32, 4), np.float32)
        fluid[..., 0] = e_ref
        fluid[7, 8, 9, 1] = np.nan
        ids = [srv.submit(s, n, checksum=True) for s, n in zip(specs, steps)]
        bad = srv.submit(poison, 200, checksum=True, init_state=(field, fluid))
        rb = srv.wait(bad, 300)
        _log(f"poisoned job: status {rb.status} after {rb.steps} steps, nonfinite {rb.final_diag.nonfinite}")
        assert rb.status == bs.BS_ERR_NONFINITE
        assert rb.steps < 200 and rb.final_diag.nonfinite == 1
        for k, j in enumerate(ids):
            r = srv.wait(j, 300)
            assert r.status == bs.BS_OK and r.steps == steps[k]
            assert r.state_checksum == alone[k], f"job {k} differs beside the non-finite one"
        st = srv.stats()
        assert st.nonfinite == 1 and st.failed == 1 and st.completed == 8
    finally:
        srv.close()
