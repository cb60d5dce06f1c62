"""Failure detection and containment, with the failure forced from the host
(a non-finite value written into a state through the upload entry points;
NaN arithmetic on the GPU traps nothing, so every kernel exits normally):

* the diagnostics reduction flags a non-finite state (`nonfinite`), which is
  what the headless driver aborts on (headless.rs:63-67), and the stepping
  itself keeps going without error;
* in a batch, a box driven non-finite leaves every other box bit for bit on
  the trajectory it follows alone (tasks never span boxes), while its own
  diagnostics report the failure.
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
        b.advance(4, use_graph=True)
        assert b.diagnostics(1).nonfinite == 1
        for i in (0, 2):
            _log(f"box {i} grid {tuple(specs[i].grid)} against its solo run")
            sim = bs.Simulation(specs[i])
            sim.advance(6)
            sf, su = sim.download()
            sim.close()
            bf, bu = b.download(i)
            assert b.steps(i) == 6
            assert np.array_equal(sf, bf) and np.array_equal(su, bu)
            assert b.diagnostics(i).nonfinite == 0
    finally:
        b.close()
