"""GPU parity of the HIP lattice solver against the oracle (CPU restatement).

Bars:
* ExactMath kernels (f32 and f64, fused and two-pass) reproduce the oracle's
  fields bit for bit (numpy equality; +0 == -0), nucleation included;
* FastMath f32 and exact f32 agree with the f64 solver to a relative L2
  error of 1e-5 per state block ((phi, pi) and (E, Z)) over 20 steps; per
  component the fast path stays within 2x the exact f32 path's own error;
* diagnostics equal the oracle's long-double sums of the same per-cell terms
  to 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def _setup(bs, orc, grid, precision, math, kernel, schedule, model=None):
    model = model or bs.Model()
    spec = bs.make_spec(grid=grid, model=model, precision=precision, math=math, kernel=kernel,
                        nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
    sim = bs.Simulation(spec)
    sim.reset(schedule)
    params = bs.sim_params_words(model, grid)
    o = orc.OracleSim(1 if precision == bs.PRECISION_F64 else 0, params, schedule)
    return sim, o


def _sched(bs, grid, events):
    r = bs.Model().seed_radius()
    return np.array([(t, x, y, z, r) for (t, x, y, z) in events], dtype=np.float32).reshape(-1, 5)


GRIDS = [(40, 24, 20), (32, 32, 32)]


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kernel", [0, 1])
def test_exact_matches_oracle_bitwise(bs, orc, gpu_available, grid, precision, kernel):
    _log(f"case grid={grid} precision={precision} kernel={kernel}")
    sched = _sched(bs, grid, [(0.0, grid[0] / 2, grid[1] / 2, grid[2] / 2), (0.6, 3.0, 5.0, 7.5)])
    sim, o = _setup(bs, orc, grid, precision, bs.MATH_EXACT, kernel, sched)
    for chunk in (1, 2, 3):
        sim.advance(chunk)
        o.advance(chunk)
        gf, gu = sim.download()
        of, ou = o.state()
        assert np.array_equal(gf, of), f"field differs after {sim.steps} steps: max {np.abs(gf - of).max()}"
        assert np.array_equal(gu, ou), f"fluid differs after {sim.steps} steps: max {np.abs(gu - ou).max()}"
    assert sim.remaining_bubbles() == o.remaining() == 0
    assert sim.time == o.time


@pytest.mark.parametrize("precision", [0, 1])
def test_diagnostics_match_oracle(bs, orc, gpu_available, precision):
    grid = (40, 24, 20)
    sched = _sched(bs, grid, [(0.0, 20.0, 12.0, 10.0)])
    sim, o = _setup(bs, orc, grid, precision, bs.MATH_EXACT, bs.KERNEL_FUSED, sched)
    sim.advance(4)
    o.advance(4)
    d = sim.diagnostics()
    ref = o.diagnostics_sums()
    n = np.prod(grid)
    got = np.array([d.mean_energy * n, d.broken_fraction * n, d.mean_kinetic * n, d.mean_fluid_energy * n])
    np.testing.assert_allclose(got, ref[:4], rtol=1e-12, atol=0)
    assert d.max_velocity == ref[4]
    assert d.nonfinite == 0


def test_quiescent_state_is_a_fixed_point(bs, gpu_available):
    # docs/NUMERICS.md:163-175: a uniform state is a discrete fixed point
    for math in (bs.MATH_FAST, bs.MATH_EXACT):
        spec = bs.make_spec(grid=32, math=math, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
        sim = bs.Simulation(spec)
        sim.reset(np.zeros((0, 5), dtype=np.float32))
        f0, u0 = sim.download()
        sim.advance(10)
        f1, u1 = sim.download()
        assert np.array_equal(f0, f1) and np.array_equal(u0, u1)
        d = sim.diagnostics()
        assert d.max_velocity == 0.0 and d.kinetic_fraction == 0.0 and d.broken_fraction == 0.0


def _rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-300)))


@pytest.mark.parametrize("kernel", [0, 1])
def test_fast_close_to_exact_and_f64(bs, gpu_available, kernel):
    # Error norm: relative L2 over the box of each state block -- (phi, pi)
    # and (E, Zx, Zy, Zz) -- against the f64 solver, bar 1e-5.  Per component,
    # the fast path is within 2x the reference's own f32 arithmetic
    # (the exact f32 path): a nearly-zero component such as Zx early on is
    # dominated by f32 roundoff relative to its own small norm.
    grid = (48, 48, 48)
    sched = _sched(bs, grid, [(0.0, 24.0, 24.0, 24.0), (1.0, 5.0, 9.0, 30.0)])
    out = {}
    for name, prec, math in (("fast", 0, bs.MATH_FAST), ("exact", 0, bs.MATH_EXACT), ("f64", 1, bs.MATH_EXACT)):
        spec = bs.make_spec(grid=grid, precision=prec, math=math, kernel=kernel)
        sim = bs.Simulation(spec)
        sim.reset(sched)
        sim.advance(20)
        out[name] = sim.download()
    for k in range(2):
        ref = out["f64"][k]
        e_fast = _rel_l2(out["fast"][k], ref)
        e_exact = _rel_l2(out["exact"][k], ref)
        _log(f"block {k}: fast-vs-f64 {e_fast:.3e}  exact-f32-vs-f64 {e_exact:.3e}")
        assert e_fast < 1e-5 and e_exact < 1e-5
        for c in range(ref.shape[-1]):
            if np.sum(ref[..., c].astype(np.float64) ** 2) == 0:
                assert np.all(out["fast"][k][..., c] == 0)
                continue
            ef = _rel_l2(out["fast"][k][..., c], ref[..., c])
            ee = _rel_l2(out["exact"][k][..., c], ref[..., c])
            _log(f"  comp {c}: fast-vs-f64 {ef:.3e}  exact-f32-vs-f64 {ee:.3e}")
            assert ef <= 2.0 * ee + 1e-7


def test_fused_equals_two_pass_fast_small_tiles(bs, gpu_available):
    # the fused and two-pass fast kernels evaluate the same faces; contraction
    # may differ between the two code paths, so this is a tolerance check
    grid = (64, 48, 40)
    sched = _sched(bs, grid, [(0.0, 30.0, 20.0, 17.0)])
    res = []
    for kernel in (bs.KERNEL_FUSED, bs.KERNEL_TWO_PASS):
        sim = bs.Simulation(bs.make_spec(grid=grid, kernel=kernel))
        sim.reset(sched)
        sim.advance(15)
        res.append(sim.download())
    for k in range(2):
        assert _rel_l2(res[0][k], res[1][k]) < 1e-6


TILES = ["32x16", "64x8", "32x8", "16x16", "64x4", "16x8", "32x4", "16x15"]


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("zc", ["16", "7", "4", "2"])
def test_every_tile_shape_matches_oracle(bs, orc, gpu_available, monkeypatch, tile, zc):
    # the fused kernel's tiling (x/y tile, z chunk) changes which thread and
    # which launch evaluates a face, never its operands: bitwise under ExactMath
    monkeypatch.setenv("BS_TILE", tile)
    monkeypatch.setenv("BS_ZC", zc)
    grid = (36, 20, 18)
    _log(f"case tile={tile} zc={zc}")
    sched = _sched(bs, grid, [(0.0, 18.0, 10.0, 9.0), (0.4, 2.0, 17.0, 1.0)])
    sim, o = _setup(bs, orc, grid, 0, bs.MATH_EXACT, bs.KERNEL_FUSED, sched)
    sim.advance(4)
    o.advance(4)
    gf, gu = sim.download()
    of, ou = o.state()
    assert np.array_equal(gf, of) and np.array_equal(gu, ou)


@pytest.mark.parametrize("tile,grid", [("16x16", (64, 32, 20)), ("16x8", (64, 32, 20)), ("16x8", (256, 16, 12)),
                                       ("32x8", (256, 32, 8)), ("32x8", (512, 8, 8))])
def test_xcd_tile_maps_match_oracle(bs, orc, gpu_available, monkeypatch, tile, grid):
    # tile counts per z chunk that divide by 8 switch the order in which
    # tiles are dealt to the XCDs: the first RK stage on tiles of 16 columns
    # takes bands of tile rows (8, 16 and 32 tiles per chunk here), stages 2
    # and 3 a 4 x 2 grid of rectangles where the tile columns divide by 4 and
    # the rows by 2 (1 x 2, 4 x 1 and 2 x 2 tiles per rectangle here; the
    # single tile row of the 512-wide plane keeps the plain order).  Every
    # tile must still be evaluated exactly once, so the state stays the
    # oracle's bit for bit
    monkeypatch.setenv("BS_TILE", tile)
    monkeypatch.setenv("BS_ZC", "8")
    _log(f"case tile={tile} grid={grid}")
    sched = _sched(bs, grid, [(0.0, grid[0] / 2 - 2, grid[1] / 2, grid[2] / 2), (0.4, 3.0, grid[1] - 2.0, 1.0)])
    sim, o = _setup(bs, orc, grid, 0, bs.MATH_EXACT, bs.KERNEL_FUSED, sched)
    sim.advance(4)
    o.advance(4)
    gf, gu = sim.download()
    of, ou = o.state()
    assert np.array_equal(gf, of) and np.array_equal(gu, ou)


@pytest.mark.parametrize("grid", [(8, 8, 8), (8, 9, 10), (17, 8, 33)])
def test_smallest_and_odd_boxes_match_oracle(bs, orc, gpu_available, grid):
    # boxes smaller than a tile wrap the tile onto itself (min edge 8, sim.rs:47)
    _log(f"case grid={grid}")
    sched = _sched(bs, grid, [(0.0, grid[0] / 2, grid[1] / 2, grid[2] / 2)])
    model = bs.Model(seed_cells=3.0)
    sched[:, 4] = model.seed_radius()
    for precision in (0, 1):
        sim, o = _setup(bs, orc, grid, precision, bs.MATH_EXACT, bs.KERNEL_FUSED, sched, model=model)
        sim.advance(5)
        o.advance(5)
        gf, gu = sim.download()
        of, ou = o.state()
        assert np.array_equal(gf, of) and np.array_equal(gu, ou)


def test_centre_bubble_64_against_f64_over_300_steps(bs, gpu_available):
    # 64^3, eta 0.3, eps_ratio 0.3, seed factor 1.15, one bubble at the
    # centre cell (the --bubble-at-centre extension), 300 steps: the f32
    # states (FastMath and the reference's f32 operation order) against the
    # f64 solver at steps 100, 200 and 300, per state block, with the energy
    # drift of each run from the first report after nucleation.
    model = bs.Model(eta=0.3, eps_ratio=0.3, seed_factor=1.15)
    runs = {}
    for name, prec, math in (("fast", 0, bs.MATH_FAST), ("exact", 0, bs.MATH_EXACT), ("f64", 1, bs.MATH_EXACT)):
        spec = bs.make_spec(grid=64, model=model, precision=prec, math=math, nucleation=bs.NUC_SIMULTANEOUS,
                            bubbles=1, bubble_at_centre=True)
        sim = bs.Simulation(spec)
        states, energy = [], []
        for _ in range(3):
            sim.advance(100)
            states.append(sim.download())
            energy.append(sim.diagnostics().mean_energy)
        sim.close()
        runs[name] = (states, energy)
    for r in range(3):
        for k, block in enumerate(("phi/pi", "fluid")):
            ref = runs["f64"][0][r][k]
            e_fast = _rel_l2(runs["fast"][0][r][k], ref)
            e_exact = _rel_l2(runs["exact"][0][r][k], ref)
            _log(f"step {100 * (r + 1)} {block}: fast-vs-f64 {e_fast:.3e}  exact-f32-vs-f64 {e_exact:.3e}")
            assert e_fast < 1e-5
            assert e_fast <= 2.0 * e_exact + 1e-7
    drift = {n: runs[n][1][2] / runs[n][1][0] - 1.0 for n in runs}
    _log("energy drift, steps 100 to 300: " + ", ".join(f"{n} {d:+.3e}" for n, d in drift.items()))
    # f32 drift is f64's plus f32 rounding; FastMath's no larger than twice
    # the reference's f32 order's
    assert abs(drift["fast"] - drift["f64"]) <= 2.0 * abs(drift["exact"] - drift["f64"]) + 1e-6
