// render.cpp -- host side of the offscreen renderer: the render settings
// (RenderSettings, render.rs:73-126), the orbit camera framed on the box
// (camera.rs:24-101) and the view uniform it yields (render.rs:288-334), and
// the PNG writer of the screenshot (capture.rs:99-114).
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace bs;

namespace {

static_assert(sizeof(bs_render_view) == 176, "the view uniform is mat4x4 + 7 vec4 (render.rs:409-413)");

// FieldMode::default_gain / default_colormap (render.rs:47-71)
float default_gain(int mode) { return mode == 0 ? 0.5f : (mode == 1 ? 4.0f : (mode == 2 ? 15.0f : 20.0f)); }
int default_colormap(int mode) { return mode == 0 ? 0 : (mode == 2 ? 2 : 1); }

struct Vec3 {
    float x, y, z;
};
Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec3 normalize(Vec3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }

// 4x4 matrices, column-major: m[c * 4 + r]
struct Mat4 {
    float m[16];
};
Mat4 mul(const Mat4 &a, const Mat4 &b) {
    Mat4 r{};
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + i] * b.m[c * 4 + k];
            r.m[c * 4 + i] = s;
        }
    return r;
}
// right-handed perspective onto clip depth [0, 1]: the camera looks down -z,
// z = -near maps to depth 0 and z = -far to depth 1
Mat4 perspective_rh(float fov_y, float aspect, float znear, float zfar) {
    const float h = 1.0f / std::tan(0.5f * fov_y);
    const float w = h / aspect;
    const float r = zfar / (znear - zfar);
    return Mat4{{w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1, 0, 0, r * znear, 0}};
}
// right-handed view matrix: the eye at the origin looking down -z, +y up
Mat4 look_at_rh(Vec3 eye, Vec3 center, Vec3 up) {
    const Vec3 f = normalize(center - eye);
    const Vec3 s = normalize(cross(f, up));
    const Vec3 u = cross(s, f);
    return Mat4{{s.x, u.x, -f.x, 0, s.y, u.y, -f.y, 0, s.z, u.z, -f.z, 0, -dot(eye, s), -dot(eye, u), dot(eye, f), 1}};
}
// inverse by Gauss-Jordan elimination with partial pivoting, in double
bool inverse(const Mat4 &a, Mat4 &out) {
    double m[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            m[r][c] = a.m[c * 4 + r];
            m[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
        if (m[p][c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < 8; ++k) std::swap(m[p][k], m[c][k]);
        const double d = m[c][c];
        for (int k = 0; k < 8; ++k) m[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
            for (int k = 0; k < 8; ++k) m[r][k] -= f * m[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out.m[c * 4 + r] = (float)m[r][c + 4];
    return true;
}

void put_be32(std::vector<unsigned char> &v, uint32_t x) {
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}
void put_chunk(std::vector<unsigned char> &png, const char type[4], const unsigned char *data, size_t n) {
    put_be32(png, (uint32_t)n);
    const size_t at = png.size();
    png.insert(png.end(), type, type + 4);
    if (n) png.insert(png.end(), data, data + n);
    const uLong crc = crc32(0L, png.data() + at, (uInt)(n + 4));
    put_be32(png, (uint32_t)crc);
}

}  // namespace

extern "C" {

int bs_render_settings_default(int32_t field_mode, bs_render_settings *out) {
    if (!out) return set_error(BS_ERR_INVALID, "null argument");
    if (field_mode < 0 || field_mode > 3)
        return set_error(BS_ERR_INVALID, "field mode " + std::to_string(field_mode) +
                                             " is not one of phi (0), speed (1), temperature (2), kinetic (3)");
    out->field_mode = field_mode;
    out->colormap = default_colormap(field_mode);
    out->iso_level = 0.5f;
    out->wall_opacity = 0.55f;
    out->samples_per_cell = 1.5f;
    out->exposure = 1.5f;
    out->field_gain = default_gain(field_mode);
    out->absorption = 12.0f;
    out->show_iso = 1;
    out->show_volume = 1;
    out->show_box = 1;
    out->background = 1.0f;
    out->clip_axis = -1;
    out->clip_pos = 0.0f;
    return BS_OK;
}

int bs_render_view_make(const uint32_t grid[3], const bs_render_settings *st, uint32_t width, uint32_t height,
                        uint32_t frame, bs_render_view *out) {
    if (!grid || !st || !out) return set_error(BS_ERR_INVALID, "null argument");
    if (width == 0 || height == 0) return set_error(BS_ERR_INVALID, "the frame must be at least 1x1 pixels");
    if (grid[0] == 0 || grid[1] == 0 || grid[2] == 0) return set_error(BS_ERR_INVALID, "empty grid");
    if (st->field_mode < 0 || st->field_mode > 3 || st->colormap < 0 || st->colormap > 2 || st->clip_axis < -1 ||
        st->clip_axis > 2)
        return set_error(BS_ERR_INVALID, "render settings out of range (field mode, colour map or clip axis)");
    // box_dimensions (render.rs:363-371): the longest axis spans 1.0
    const float n_max = (float)std::max(grid[0], std::max(grid[1], grid[2]));
    const Vec3 half{0.5f * (float)grid[0] / n_max, 0.5f * (float)grid[1] / n_max, 0.5f * (float)grid[2] / n_max};
    const float cell = 1.0f / n_max;
    // OrbitCamera::default, then frame_box (camera.rs:24-34, 98-101)
    const float yaw = 0.6f, pitch = 0.42f, fov_y = 50.0f * 3.14159265358979323846f / 180.0f;
    const float znear = 0.002f, zfar = 100.0f;
    const Vec3 target{0.0f, 0.0f, 0.0f};
    const float distance = std::sqrt(dot(half, half)) * 2.4f;
    const Vec3 offset = Vec3{std::cos(pitch) * std::sin(yaw), std::sin(pitch), std::cos(pitch) * std::cos(yaw)} * distance;
    const Vec3 eye = target + offset;
    const Mat4 vp = mul(perspective_rh(fov_y, std::max((float)width / (float)height, 1e-3f), znear, zfar),
                        look_at_rh(eye, target, Vec3{0.0f, 1.0f, 0.0f}));
    Mat4 inv;
    if (!inverse(vp, inv)) return set_error(BS_ERR_INVALID, "singular view-projection matrix");
    std::memcpy(out->inv_view_proj, inv.m, sizeof inv.m);
    const float cam[4] = {eye.x, eye.y, eye.z, 0.0f};
    const float bh[4] = {half.x, half.y, half.z, cell};
    const float gr[4] = {(float)grid[0], (float)grid[1], (float)grid[2], 0.0f};
    const float r0[4] = {st->iso_level, st->wall_opacity, st->samples_per_cell, st->exposure};
    const float r1[4] = {(float)st->field_mode, (float)st->clip_axis, st->clip_pos, (float)st->colormap};
    const float r2[4] = {st->field_gain, st->absorption, st->show_iso ? 1.0f : 0.0f, st->show_volume ? 1.0f : 0.0f};
    const float r3[4] = {(float)(frame % 64) * 0.618034f, st->show_box ? 1.0f : 0.0f, st->background, 0.0f};
    std::memcpy(out->cam_pos, cam, sizeof cam);
    std::memcpy(out->box_half, bh, sizeof bh);
    std::memcpy(out->grid, gr, sizeof gr);
    std::memcpy(out->render0, r0, sizeof r0);
    std::memcpy(out->render1, r1, sizeof r1);
    std::memcpy(out->render2, r2, sizeof r2);
    std::memcpy(out->render3, r3, sizeof r3);
    return BS_OK;
}

int bs_png_write(const char *path, const uint8_t *rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba) return set_error(BS_ERR_INVALID, "null argument");
    if (width == 0 || height == 0) return set_error(BS_ERR_INVALID, "the image must be at least 1x1 pixels");
    // scanlines, each led by filter type 0 (none)
    const size_t row = (size_t)width * 4;
    std::vector<unsigned char> raw((row + 1) * height);
    for (uint32_t y = 0; y < height; ++y) {
        raw[y * (row + 1)] = 0;
        std::memcpy(&raw[y * (row + 1) + 1], rgba + y * row, row);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return set_error(BS_ERR_INVALID, "zlib compression of the image failed");
    std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    const unsigned char rest[5] = {8, 6, 0, 0, 0};  // 8 bits, RGBA, deflate, no filter set, no interlace
    ihdr.insert(ihdr.end(), rest, rest + 5);
    put_chunk(png, "IHDR", ihdr.data(), ihdr.size());
    put_chunk(png, "IDAT", z.data(), zlen);
    put_chunk(png, "IEND", nullptr, 0);
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(BS_ERR_INVALID, std::string("cannot create ") + path);
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    if (std::fclose(f) != 0 || !ok) return set_error(BS_ERR_INVALID, std::string("failed to write ") + path);
    return BS_OK;
}

}  // extern "C"
