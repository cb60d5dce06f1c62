"""GPU checks at the headline size (512^3, f32) through properties that do not
depend on running the oracle there (it would take hours on the host):

* ExactMath fused and two-pass kernels both reproduce the oracle bit for bit
  at small sizes, so at 512^3 they must equal each other bit for bit;
* translation invariance: a state shifted by a vector that is not a multiple
  of the tile or z chunk evolves into the shifted state, bit for bit under
  ExactMath (every face and cell sees the same operands, whichever thread,
  tile-edge round or z-chunk prologue evaluates it), to rounding under
  FastMath;
* FastMath stays within its error bar of ExactMath (relative L2 per state
  block, 1e-5, the bar of the small-size parity tests);
* a uniform state is a fixed point of the FastMath kernel.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 512
SHIFT = (5, 3, 7)  # cells along x, y, z; not a multiple of any tile edge or z chunk


def _log(*a):
    print(*a, flush=True)


def _sched(bs, centres, shift=(0, 0, 0)):
    r = bs.Model().seed_radius()
    rows = [(0.0, (x + shift[0]) % N, (y + shift[1]) % N, (z + shift[2]) % N, r) for (x, y, z) in centres]
    return np.array(rows, dtype=np.float32).reshape(-1, 5)


# bubbles that straddle x/y tile edges, z-chunk boundaries and the periodic wrap
CENTRES = [(256.0, 256.0, 256.0), (31.0, 7.0, 32.0), (500.0, 130.0, 3.0), (64.5, 480.0, 450.0)]


def _run(bs, math, kernel, sched, steps):
    spec = bs.make_spec(grid=N, math=math, kernel=kernel, nucleation=bs.NUC_SIMULTANEOUS, bubbles=len(CENTRES))
    sim = bs.Simulation(spec)
    try:
        sim.reset(sched)
        sim.advance(steps)
        return sim.download()
    finally:
        sim.close()


def _rel_l2(a, b):
    d = 0.0
    s = 0.0
    for k in range(a.shape[0]):  # plane by plane, in f64
        x = a[k].astype(np.float64)
        y = b[k].astype(np.float64)
        d += float(np.sum((x - y) ** 2))
        s += float(np.sum(y ** 2))
    return (d / max(s, 1e-300)) ** 0.5


def test_headline_box_exact_fused_equals_two_pass(bs, gpu_available):
    _log("case 512^3 exact fused vs two-pass, 3 steps")
    sched = _sched(bs, CENTRES)
    f1, u1 = _run(bs, bs.MATH_EXACT, bs.KERNEL_FUSED, sched, 3)
    f2, u2 = _run(bs, bs.MATH_EXACT, bs.KERNEL_TWO_PASS, sched, 3)
    assert np.array_equal(f1, f2), "field differs"
    assert np.array_equal(u1, u2), "fluid differs"
    assert np.count_nonzero(f1[..., 0]) > 0  # the bubbles are there


def test_headline_box_translation_invariance_exact(bs, gpu_available):
    _log(f"case 512^3 exact fused, shift {SHIFT}, 3 steps")
    f1, u1 = _run(bs, bs.MATH_EXACT, bs.KERNEL_FUSED, _sched(bs, CENTRES), 3)
    f2, u2 = _run(bs, bs.MATH_EXACT, bs.KERNEL_FUSED, _sched(bs, CENTRES, SHIFT), 3)
    # arrays are (z, y, x, c): roll the unshifted result onto the shifted one
    roll = (SHIFT[2], SHIFT[1], SHIFT[0])
    assert np.array_equal(np.roll(f1, roll, axis=(0, 1, 2)), f2), "field is not translation invariant"
    assert np.array_equal(np.roll(u1, roll, axis=(0, 1, 2)), u2), "fluid is not translation invariant"


def test_headline_box_translation_invariance_fast(bs, gpu_available):
    # the production kernel: tile-edge faces go through the far-edge round and
    # z-chunk starts through the prologue, interior faces through the DPP and
    # LDS rounds.  The far-edge round evaluates y faces as x faces of the
    # vx <-> vy swapped states; with FMA contraction a sum of squares is then
    # rounded in the other order, so FastMath is translation invariant to
    # rounding (ExactMath, without contraction, bit for bit: above)
    _log(f"case 512^3 fast fused, shift {SHIFT}, 3 steps")
    f1, u1 = _run(bs, bs.MATH_FAST, bs.KERNEL_FUSED, _sched(bs, CENTRES), 3)
    f2, u2 = _run(bs, bs.MATH_FAST, bs.KERNEL_FUSED, _sched(bs, CENTRES, SHIFT), 3)
    roll = (SHIFT[2], SHIFT[1], SHIFT[0])
    rf = np.roll(f1, roll, axis=(0, 1, 2))
    ru = np.roll(u1, roll, axis=(0, 1, 2))
    ef = _rel_l2(rf, f2)
    eu = _rel_l2(ru, u2)
    _log(f"  values differing: field {np.count_nonzero(rf != f2)}, fluid {np.count_nonzero(ru != u2)}; "
         f"rel L2 field {ef:.3e} fluid {eu:.3e}")
    assert ef < 1e-6 and eu < 1e-6


def test_headline_box_fast_close_to_exact(bs, gpu_available):
    _log("case 512^3 fast vs exact fused, 6 steps")
    sched = _sched(bs, CENTRES)
    ff, uf = _run(bs, bs.MATH_FAST, bs.KERNEL_FUSED, sched, 6)
    fe, ue = _run(bs, bs.MATH_EXACT, bs.KERNEL_FUSED, sched, 6)
    ef = _rel_l2(ff, fe)
    eu = _rel_l2(uf, ue)
    _log(f"  rel L2 field {ef:.3e} fluid {eu:.3e}")
    assert ef < 1e-5 and eu < 1e-5


def test_headline_box_uniform_state_is_a_fixed_point_fast(bs, gpu_available):
    _log("case 512^3 fast, uniform state, 3 steps")
    spec = bs.make_spec(grid=N, math=bs.MATH_FAST, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
    sim = bs.Simulation(spec)
    try:
        sim.reset(np.zeros((0, 5), dtype=np.float32))
        f0, u0 = sim.download()
        sim.advance(3)
        f1, u1 = sim.download()
    finally:
        sim.close()
    assert np.array_equal(f0, f1) and np.array_equal(u0, u1)
