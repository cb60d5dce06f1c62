"""Offscreen volume rendering on the CPU: the render settings and camera
against the reference's own tests and definitions, the PNG writer against a
decoder, and the oracle's ray march (volume.wgsl fs_main) against closed forms:

* a ray that misses the box returns the background gradient untouched;
* a ray through an empty box returns that background tone-mapped;
* a ray through a uniform broken phase composites the emission of every
  sample front to back, 1 - (1 - a)^k after k samples, until the opacity
  passes 0.995;
* a ray that crosses a flat wall (phi a step along x) picks up the wall's
  Blinn-Phong shading with the normal +x at the wall's opacity.

Rays are formed here with the textbook pinhole camera (eye, forward, right,
up, tan(fov/2)), not through the inverse view-projection matrix the renderer
and the oracle use; expected colours are computed in float64 and compared
within one step of 8 bits.
"""
import math
import struct
import zlib

import numpy as np
import pytest

W, H = 96, 72


def _camera(grid, width, height):
    # camera.rs:24-34 (OrbitCamera::default) and frame_box (camera.rs:98-101);
    # box_dimensions (render.rs:363-371)
    n = max(grid)
    half = np.array([0.5 * g / n for g in grid])
    dist = np.linalg.norm(half) * 2.4
    yaw, pitch = 0.6, 0.42
    eye = np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)]) * dist
    f = -eye / np.linalg.norm(eye)
    s = np.cross(f, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return half, eye, f, s, u, math.tan(math.radians(50.0) / 2), width / height


def _ray(cam, px, py, width, height):
    half, eye, f, s, u, t, aspect = cam
    ndc_x = (px + 0.5) / width * 2 - 1
    ndc_y = 1 - (py + 0.5) / height * 2
    d = f + s * ndc_x * t * aspect + u * ndc_y * t
    return eye, d / np.linalg.norm(d)


def _box_hit(half, ro, rd):
    with np.errstate(divide="ignore"):
        inv = 1.0 / rd
    t0, t1 = (-half - ro) * inv, (half - ro) * inv
    return max(np.minimum(t0, t1).max(), 0.0), np.maximum(t0, t1).min()


def _background(rd):
    h = min(max(rd[1] * 0.5 + 0.5, 0.0), 1.0)
    return np.array([0.015, 0.017, 0.028]) * (1 - h) + np.array([0.045, 0.055, 0.085]) * h


def _tonemap(c):
    c = c / (1 + c)
    return np.maximum(c, 0) ** (1 / 2.2)


def _u8(c):
    return np.floor(np.clip(c, 0, 1) * 255 + 0.5)


def _view_words(bs, grid, settings, width=W, height=H, frame=1):
    v = bs.render_view(grid, settings, width, height, frame)
    return np.frombuffer(bytes(v), dtype=np.float32).copy(), v


def _centre_pixels(cam, width, height, margin=0.15):
    # pixels whose rays cross the box well away from its edges
    half = cam[0]
    out = []
    for py in range(height // 4, 3 * height // 4, 7):
        for px in range(width // 4, 3 * width // 4, 7):
            ro, rd = _ray(cam, px, py, width, height)
            tn, tf = _box_hit(half, ro, rd)
            if tf - tn < 0.3:
                continue
            ok = True
            for p in (ro + rd * tn, ro + rd * tf):
                d = np.sort(half - np.abs(p))
                if d[1] < margin:  # the median distance to an edge
                    ok = False
            if ok:
                out.append((px, py))
    return out


def test_settings_follow_render_settings_with_field(bs):
    # RenderSettings::default and with_field (render.rs:96-126), FieldMode's
    # gains and colour maps (render.rs:47-71), every_field_mode_has_a_valid_default_colormap
    gains = {0: 0.5, 1: 4.0, 2: 15.0, 3: 20.0}
    maps = {0: 0, 1: 1, 2: 2, 3: 1}
    for m in range(4):
        st = bs.render_settings(m)
        assert st.field_mode == m and st.field_gain == gains[m] and st.colormap == maps[m]
        assert (st.iso_level, st.wall_opacity, st.samples_per_cell, st.exposure, st.absorption) == \
            (0.5, pytest.approx(0.55), 1.5, 1.5, 12.0)
        assert (st.show_iso, st.show_volume, st.show_box, st.background, st.clip_axis, st.clip_pos) == \
            (1, 1, 1, 1.0, -1, 0.0)
    assert bs.render_settings().field_mode == bs.FIELD_KINETIC  # FieldMode::default
    with pytest.raises(bs.BubbleSimError):
        bs.render_settings(4)


def test_view_uniform_matches_the_wgsl_struct(bs):
    # view_uniform_matches_the_wgsl_struct (render.rs:409-413): mat4x4 + 7 vec4
    assert bs.lib().bs_struct_size(15) == 176
    import ctypes as C
    assert C.sizeof(bs.CRenderView) == 176 and bs.lib().bs_struct_size(14) == C.sizeof(bs.CRenderSettings)


@pytest.mark.parametrize("grid", [(32, 32, 32), (64, 40, 24)])
def test_camera_frames_the_box(bs, grid):
    words, v = _view_words(bs, grid, bs.render_settings())
    cam = _camera(grid, W, H)
    half, eye = cam[0], cam[1]
    # box_dimensions: the longest axis spans 1, w = one cell
    assert np.allclose(list(v.box_half), list(half) + [1 / max(grid)], rtol=1e-6)
    assert list(v.grid)[:3] == list(map(float, grid))
    # eye_sits_at_the_requested_distance (camera.rs:110-114), framed at 2.4 |half|
    assert np.allclose(list(v.cam_pos)[:3], eye, rtol=1e-5)
    # projection_maps_the_target_to_the_screen_centre (camera.rs:133-138):
    # the view-projection is the inverse of the uniform's matrix
    m = np.array(v.inv_view_proj, dtype=np.float64).reshape(4, 4).T  # column-major
    vp = np.linalg.inv(m)
    clip = vp @ np.array([0.0, 0.0, 0.0, 1.0])
    assert abs(clip[0] / clip[3]) < 1e-4 and abs(clip[1] / clip[3]) < 1e-4
    # every pixel's ray through the matrix equals the pinhole camera's ray
    for px, py in [(0, 0), (W - 1, H - 1), (W // 3, H // 2), (W // 2, 5)]:
        nx, ny = (px + 0.5) / W * 2 - 1, 1 - (py + 0.5) / H * 2
        far = m @ np.array([nx, ny, 1.0, 1.0])
        d = far[:3] / far[3] - eye
        d /= np.linalg.norm(d)
        # (the f32 inverse over the 0.002-100 depth range resolves directions to ~3e-5)
        assert np.allclose(d, _ray(cam, px, py, W, H)[1], atol=1e-4)
    # the dither offset of frame k (render.rs:323-324)
    for k in (0, 1, 5, 64, 65):
        assert _view_words(bs, grid, bs.render_settings(), frame=k)[1].render3[0] == pytest.approx((k % 64) * 0.618034)


def test_png_writer_round_trip(bs, tmp_path):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, size=(37, 53, 4), dtype=np.uint8)
    path = tmp_path / "x.png"
    bs.png_write(str(path), img)
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        chunks[kind] = chunks.get(kind, b"") + body
        pos += 12 + n
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert (w, h, depth, ctype, comp, filt, inter) == (53, 37, 8, 6, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), dtype=np.uint8).reshape(37, 53 * 4 + 1)
    assert (raw[:, 0] == 0).all()
    assert np.array_equal(raw[:, 1:].reshape(37, 53, 4), img)
    assert b"IEND" in chunks


def test_oracle_background_and_empty_box(bs, orc):
    grid = (32, 32, 32)
    st = bs.render_settings(bs.FIELD_PHI)
    words, _ = _view_words(bs, grid, st)
    img = orc.render(words, np.zeros(grid + (4,), np.float16), W, H)
    cam = _camera(grid, W, H)
    misses = hits = 0
    centre = set(_centre_pixels(cam, W, H))
    for py in list(range(0, H, 5)) + [p for _, p in centre]:
        for px in list(range(0, W, 5)) + [p for p, _ in centre]:
            ro, rd = _ray(cam, px, py, W, H)
            tn, tf = _box_hit(cam[0], ro, rd)
            if tn > tf + 1e-3:
                want = _u8(_background(rd))  # a miss returns the gradient, no tone map
                misses += 1
            elif (px, py) in centre:
                want = _u8(_tonemap(_background(rd)))
                hits += 1
            else:
                continue
            assert np.abs(img[py, px, :3].astype(int) - want).max() <= 1, (px, py)
            assert img[py, px, 3] == 255
    assert misses > 20 and hits > 3


def test_oracle_uniform_broken_phase_composites_front_to_back(bs, orc):
    grid = (32, 32, 32)
    st = bs.render_settings(bs.FIELD_PHI)
    st.field_gain = 2.0
    st.show_box = 0
    words, _ = _view_words(bs, grid, st)
    tex = np.zeros(grid + (4,), np.float16)
    tex[..., 0] = 1.0  # phi = phi_b everywhere: no wall to cross
    img = orc.render(words, tex, W, H)
    cam = _camera(grid, W, H)
    ds = (1 / 32) / 1.5
    a = 1 - math.exp(-12.0 * (2.0 * 2.0) * ds)  # intensity 2, weight 4
    k = next(k for k in range(1, 100) if 1 - (1 - a) ** k > 0.995)
    alpha = 1 - (1 - a) ** k
    emit = np.array([0.993, 0.906, 0.144]) * 1.5  # viridis at 1, times the exposure
    n = 0
    for px, py in _centre_pixels(cam, W, H):
        ro, rd = _ray(cam, px, py, W, H)
        want = _u8(_tonemap(_background(rd) * (1 - alpha) + emit * alpha))
        assert np.abs(img[py, px, :3].astype(int) - want).max() <= 1, (px, py)
        n += 1
    assert n > 3


def test_oracle_flat_wall_shading(bs, orc):
    grid = (32, 32, 32)
    st = bs.render_settings(bs.FIELD_PHI)
    st.show_volume = 0
    st.show_box = 0
    words, _ = _view_words(bs, grid, st)
    tex = np.zeros(grid + (4,), np.float16)
    tex[:, :, :16, 0] = 1.0  # broken for x < 0 (world), symmetric beyond: a wall in the plane x = 0
    img = orc.render(words, tex, W, H)
    cam = _camera(grid, W, H)
    half = cam[0]
    n = np.array([1.0, 0.0, 0.0])  # down the gradient of phi
    l1 = np.array([0.55, 0.72, 0.42]) / np.linalg.norm([0.55, 0.72, 0.42])
    l2 = np.array([-0.5, -0.2, -0.7]) / np.linalg.norm([-0.5, -0.2, -0.7])
    base = np.array([0.16, 0.62, 0.72])  # no fluid speed at the wall: the cool end of the tint
    checked = 0
    for px, py in _centre_pixels(cam, W, H, margin=0.1):
        ro, rd = _ray(cam, px, py, W, H)
        tn, tf = _box_hit(half, ro, rd)
        pin, pout = ro + rd * tn, ro + rd * tf
        # the ray must cross x = 0 inside the box, away from the faces
        if not (pin[0] > 0.1 and pout[0] < -0.1):
            continue
        view = -rd
        diff = max(n @ l1, 0) + 0.35 * max(n @ l2, 0)
        h = l1 + view + np.array([0, 1e-5, 0])
        h /= np.linalg.norm(h)
        spec = max(n @ h, 0) ** 48
        fres = (1 - abs(n @ view)) ** 3
        shaded = base * (0.22 + 0.85 * diff) + spec * 0.35 + base * fres * 0.9
        want = _u8(_tonemap(_background(rd) * (1 - 0.55) + shaded * 0.55))
        assert np.abs(img[py, px, :3].astype(int) - want).max() <= 1, (px, py)
        checked += 1
    assert checked > 3
