"""Visualisation pack (vis.wgsl:32-55, sim.rs:920-942) on the GPU.

The texels must equal the oracle's channels of the same state, narrowed to
half with round-to-nearest-even, bit for bit, for f32 and f64 lattices: the
pack is evaluated with the exact arithmetic policy whatever the stepping
policy, so its operation sequence is the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(bs, grid, precision, math, steps):
    model = bs.Model()
    spec = bs.make_spec(grid=grid, model=model, precision=precision, math=math,
                        nucleation=bs.NUC_SIMULTANEOUS, bubbles=1)
    sim = bs.Simulation(spec)
    r = model.seed_radius()
    sim.reset(np.array([(0.0, grid[0] / 2, grid[1] / 2, grid[2] / 2, r), (0.4, 3.0, 5.0, 2.5, r)],
                       dtype=np.float32))
    sim.advance(steps)
    return sim, model


@pytest.mark.parametrize("precision,math", [(0, 1), (0, 0), (1, 1)])
def test_vis_texels_equal_oracle_channels(bs, orc, gpu_available, precision, math):
    grid = (40, 24, 20)
    print(f"case precision={precision} math={math}", flush=True)
    sim, model = _run(bs, grid, precision, math, 6)
    base = sim.device_memory_bytes()
    with pytest.raises(bs.BubbleSimError) as e:
        sim.download_visualisation()
    assert e.value.code == bs.BS_ERR_NOT_READY
    sim.record_visualisation()
    tex = sim.download_visualisation()
    assert tex.shape == (grid[2], grid[1], grid[0], 4)
    assert sim.device_memory_bytes() == base + 8 * sim.cell_count()
    ptr, nbytes = sim.visualisation_buffer()
    assert ptr != 0 and nbytes == 8 * sim.cell_count()

    field, fluid = sim.download()
    o = orc.OracleSim(precision, bs.sim_params_words(model, grid), np.zeros((0, 5), np.float32))
    o.set_state(field, fluid)
    want = o.vis_channels().astype(np.float16).reshape(tex.shape)
    got_bits, want_bits = tex.view(np.uint16), want.view(np.uint16)
    bad = np.argwhere(got_bits != want_bits)
    assert bad.size == 0, f"{len(bad)} texel channels differ, first {bad[:3].tolist()}"
    # channel sanity: the bubble is broken phase, the rest symmetric and
    # nearly at rest
    assert float(tex[..., 0].max()) > 0.9 and float(tex[..., 1].max()) < 1.0
    o.close()
    sim.close()


def test_vis_follows_the_state(bs, gpu_available):
    # a second record after more steps reflects the new state; a resize drops
    # the texel buffer with the lattice
    grid = (32, 32, 32)
    sim, _ = _run(bs, grid, 0, 1, 2)
    sim.record_visualisation()
    a = sim.download_visualisation().copy()
    sim.advance(10)
    sim.record_visualisation()
    b = sim.download_visualisation()
    assert not np.array_equal(a.view(np.uint16), b.view(np.uint16))
    sim.resize((24, 24, 24))
    with pytest.raises(bs.BubbleSimError):
        sim.download_visualisation()
    sim.record_visualisation()
    assert sim.download_visualisation().shape == (24, 24, 24, 4)
    sim.close()
