"""The offscreen volume renderer on the GPU (volume.wgsl fs_main as a HIP
kernel) against the oracle's restatement, on the texels of real states:

* every field mode, a cutaway, a ragged box: the frames agree pixel for
  pixel, up to one step of 8 bits in a few pixels (powf / expf differ in
  their last bit between the device library and the host's libm; every other
  operation, the texel filter included, is the same IEEE sequence);
* bs_sim_screenshot writes the frame bs_sim_render gives for the same state,
  as a PNG;
* the CLI's --screenshot after --headless N writes that PNG at the requested
  size (headless.rs:102-113, capture.rs:114).
"""
import os
import struct
import subprocess
import time
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 320, 240


def _log(*a):
    print(*a, flush=True)


def _state(bs, grid, steps, bubbles=2):
    spec = bs.make_spec(grid=grid, nucleation=bs.NUC_SIMULTANEOUS, bubbles=bubbles, seed=3,
                        model=bs.Model(eta=0.1, seed_factor=1.3))
    sim = bs.Simulation(spec)
    sim.advance(steps)
    sim.record_visualisation()
    return sim


def _compare(gpu, ref, what):
    d = np.abs(gpu.astype(int) - ref.astype(int))
    frac = float((d > 0).any(axis=-1).mean())
    _log(f"{what}: max diff {d.max()}, pixels differing {frac:.5f}")
    assert d.max() <= 1, what
    assert frac <= 0.01, what


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        chunks[data[pos + 4:pos + 8]] = chunks.get(data[pos + 4:pos + 8], b"") + data[pos + 8:pos + 8 + n]
        pos += 12 + n
    w, h = struct.unpack(">II", chunks[b"IHDR"][:8])
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), dtype=np.uint8).reshape(h, w * 4 + 1)
    return raw[:, 1:].reshape(h, w, 4)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_render_equals_oracle_every_field(bs, orc, gpu_available, mode):
    grid = (48, 48, 48)
    sim = _state(bs, grid, 80)
    tex = sim.download_visualisation()
    st = bs.render_settings(mode)
    view = bs.render_view(grid, st, W, H, 1)
    img = sim.render(view, W, H)
    ref = orc.render(np.frombuffer(bytes(view), dtype=np.float32), tex, W, H)
    # the wall and the field show up: the frame is not background alone
    assert len(np.unique(img.reshape(-1, 4), axis=0)) > 200
    _compare(img, ref, f"field {mode}")
    sim.close()


@pytest.mark.parametrize("grid,clip", [((40, 24, 20), -1), ((48, 48, 48), 0), ((48, 48, 48), 2)])
def test_render_equals_oracle_ragged_and_cutaway(bs, orc, gpu_available, grid, clip):
    sim = _state(bs, grid, 60, bubbles=3)
    tex = sim.download_visualisation()
    st = bs.render_settings(bs.FIELD_PHI)
    st.clip_axis = clip
    st.clip_pos = 0.2
    for frame in (1, 7):
        view = bs.render_view(grid, st, 200, 150, frame)
        img = sim.render(view, 200, 150)
        ref = orc.render(np.frombuffer(bytes(view), dtype=np.float32), tex, 200, 150)
        _compare(img, ref, f"grid {grid} clip {clip} frame {frame}")
    sim.close()


def test_screenshot_is_the_rendered_frame(bs, gpu_available, tmp_path):
    grid = (48, 40, 32)
    sim = _state(bs, grid, 50)
    st = bs.render_settings(bs.FIELD_KINETIC)
    path = str(tmp_path / "shot.png")
    sim.screenshot(path, st, (160, 120))
    img = sim.render(bs.render_view(grid, st, 160, 120, 1), 160, 120)
    assert np.array_equal(_read_png(path), img)
    sim.close()


def test_cli_screenshot(bs, gpu_available, tmp_path):
    path = str(tmp_path / "cli.png")
    r = subprocess.run([bs.CLI_PATH, "--grid", "32", "--bubbles", "2", "--nucleation", "simultaneous", "--headless",
                        "20", "--field", "phi", "--screenshot", path, "--screenshot-size", "128x96"],
                       capture_output=True, text=True, timeout=300)
    _log(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    assert f"wrote {path} (128x96)" in r.stdout
    img = _read_png(path)
    assert img.shape == (96, 128, 4) and (img[..., 3] == 255).all()
    assert len(np.unique(img.reshape(-1, 4), axis=0)) > 50


def test_render_timing_at_the_default_size(bs, gpu_available):
    # for the record: one 1280x960 frame of the default 192^3 lattice, and of 512^3
    for n, steps in ((192, 200), (512, 40)):
        sim = _state(bs, (n, n, n), steps, bubbles=8)
        st = bs.render_settings()
        view = bs.render_view((n, n, n), st, 1280, 960, 1)
        sim.render(view, 1280, 960)  # warm
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            img = sim.render(view, 1280, 960)
        dt = (time.perf_counter() - t0) / reps
        _log(f"{n}^3: {dt * 1e3:.2f} ms per 1280x960 frame (including the 4.9 MB copy to the host)")
        assert img.shape == (960, 1280, 4)
        sim.close()


def test_screenshot_size_limit(bs, gpu_available, tmp_path):
    # the frame limit is checked before any buffer is sized
    sim = _state(bs, (32, 32, 32), 5)
    st = bs.render_settings(bs.FIELD_PHI)
    with pytest.raises(bs.BubbleSimError) as e:
        sim.screenshot(str(tmp_path / "big.png"), st, (100000, 100000))
    assert e.value.code == bs.BS_ERR_INVALID and "16384" in str(e.value)
    sim.close()
    r = subprocess.run([bs.CLI_PATH, "--grid", "32", "--headless", "2", "--screenshot", str(tmp_path / "c.png"),
                        "--screenshot-size", "100000x100000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "16384" in r.stderr
