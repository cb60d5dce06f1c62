"""Memory cap and resize (sim.rs:705-782, vram.rs:52-69), after the
reference's GPU tests in src/gpu/gpu_tests.rs:80-253 (the renderer rebind test
is out of scope: there is no renderer here).

Each resized lattice must step and report sane diagnostics; a resize must
give back the old lattice's memory; the predicted lattice size must match the
allocation; a rejected resize must leave the simulation as it was.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


def spec(bs, grid):
    # gpu_tests.rs:33-43: default model, one simultaneous bubble, seed 1
    return bs.make_spec(grid=grid, nucleation=bs.NUC_SIMULTANEOUS, bubbles=1, duration=10.0, seed=1)


def step_and_check(sim, what):
    # gpu_tests.rs:45-70
    sim.advance(8)
    d = sim.diagnostics()
    assert math.isfinite(d.mean_energy), f"{what}: energy is not finite"
    assert math.isfinite(d.max_velocity), f"{what}: max |v| is not finite"
    assert d.mean_energy > 0.0, f"{what}: energy collapsed to {d.mean_energy}"
    assert 0.0 <= d.broken_fraction <= 1.0, f"{what}: broken fraction {d.broken_fraction} is outside [0, 1]"
    assert d.max_velocity < 1.0, f"{what}: max |v| = {d.max_velocity} is superluminal"


def test_resize_rebuilds_a_working_simulation(bs, gpu_available):
    # gpu_tests.rs:72-99
    sim = bs.Simulation(spec(bs, 64))
    assert sim.cell_count() == 64 ** 3
    step_and_check(sim, "initial 64^3")
    sim.resize((96, 96, 96))
    assert sim.grid == (96, 96, 96)
    assert sim.cell_count() == 96 ** 3
    assert sim.time == 0.0, "a resize restarts the run"
    step_and_check(sim, "grown to 96^3")
    sim.resize((32, 32, 32))
    assert sim.cell_count() == 32 ** 3
    step_and_check(sim, "shrunk to 32^3")
    sim.resize((48, 32, 72))
    assert sim.grid == (48, 32, 72)
    assert sim.cell_count() == 48 * 32 * 72
    step_and_check(sim, "non-cubic 48x32x72")
    sim.close()


def test_resize_releases_the_old_lattice(bs, gpu_available):
    # gpu_tests.rs:101-145.  The reference reads wgpu's allocator report; the
    # HIP runtime may keep freed blocks cached for reuse, so the release is
    # shown by what it allows: after shrinking a lattice that holds 45 % of
    # the free device memory, a second lattice of 60 % must still fit.
    free0, _ = bs.device_memory()
    side_big, _ = bs.max_cubic_size(int(0.45 * free0))
    side_two, _ = bs.max_cubic_size(int(0.60 * free0))
    big, small = (side_big,) * 3, (64, 64, 64)
    big_bytes = bs.lattice_bytes(big)
    print(f"free {free0 / 1e9:.1f} GB: lattice {side_big}^3 ({big_bytes / 1e9:.1f} GB), then {side_two}^3",
          flush=True)
    sim = bs.Simulation(spec(bs, big))
    step_and_check(sim, "big")
    free_peak, _ = bs.device_memory()
    assert free0 - free_peak >= big_bytes // 2, (
        f"expected the {big_bytes / 1e9:.2f} GB lattice to show up in the device's free memory")
    sim.resize(small)
    second = bs.Simulation(spec(bs, (side_two,) * 3))   # out of memory if the big lattice were still held
    step_and_check(second, "second lattice after the release")
    second.close()
    step_and_check(sim, "after release")
    sim.close()


@pytest.mark.parametrize("grid", [(64, 64, 64), (128, 64, 96), (192, 192, 192)])
def test_predicted_lattice_size_matches_the_allocation(bs, gpu_available, grid):
    # gpu_tests.rs:147-166
    sim = bs.Simulation(spec(bs, grid))
    predicted = bs.lattice_bytes(grid)
    actual = sim.device_memory_bytes()
    err = abs(actual - predicted) / predicted
    assert err < 0.01, f"{grid}: predicted {predicted} B but allocated {actual} B ({100 * err:.2f}% off)"
    sim.close()


def test_a_rejected_resize_leaves_the_simulation_intact(bs, gpu_available):
    # gpu_tests.rs:168-188
    sim = bs.Simulation(spec(bs, 64))
    with pytest.raises(bs.BubbleSimError):
        sim.resize((4, 4, 4))
    assert sim.grid == (64, 64, 64), "a rejected resize must not change the grid"
    free, total = bs.device_memory()
    side, _ = bs.max_cubic_size(total)
    with pytest.raises(bs.BubbleSimError) as e:
        sim.resize((side + 64,) * 3)
    assert e.value.code in (bs.BS_ERR_OOM, bs.BS_ERR_INVALID)
    assert sim.grid == (64, 64, 64)
    step_and_check(sim, "after rejected resizes")
    sim.close()
