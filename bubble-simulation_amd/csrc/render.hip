// render.hip -- offscreen volume rendering of the visualisation texels: the
// ray march of volume.wgsl:203-381 (fs_main) as a HIP kernel, one thread per
// pixel on 8x8 pixel tiles (one wave each, so the rays of a wave stay close
// and their texel fetches share cache lines).
//
// The texels are the rgba16f cells bs_sim_record_visualisation packs (x
// fastest).  The reference samples them through a hardware trilinear filter
// with clamp-to-edge addressing (render.rs:184-196); here that filter is
// written out -- texel centres at (i + 1/2) / n, the lower corner
// floor(u n - 1/2), both corners clamped into the grid, weights from the
// unclamped coordinate, lerps along x, then y, then z -- in f32, so every
// sample is a fixed sequence of IEEE operations.  Compiled without FMA
// contraction: the oracle restates the march operation for operation, and
// the images agree up to the last bit of a few pixels (pow and exp).
#include <hip/hip_runtime.h>

#include "launch.h"

namespace bs {

namespace {

constexpr int kMaxSteps = 2048;  // volume.wgsl:35

struct F3 {
    float x, y, z;
};
__device__ inline F3 f3(float a, float b, float c) { return F3{a, b, c}; }
__device__ inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline F3 operator*(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ inline F3 operator/(F3 a, F3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ inline F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
__device__ inline F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 normalize(F3 a) { return a / sqrtf(dot(a, a)); }
__device__ inline F3 fabs3(F3 a) { return f3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ inline F3 fmin3(F3 a, F3 b) { return f3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ inline F3 fmax3(F3 a, F3 b) { return f3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
// WGSL mix: e1 (1 - e3) + e2 e3
__device__ inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ inline F3 mix3(F3 a, F3 b, float t) { return f3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline float fractf(float x) { return x - floorf(x); }
__device__ inline float smoothstepf(float lo, float hi, float x) {
    const float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// Colour maps, 11 control points each (volume.wgsl:41-79)
__constant__ float kRamps[3][11][3] = {
    {{0.267f, 0.005f, 0.329f}, {0.283f, 0.141f, 0.458f}, {0.254f, 0.265f, 0.530f}, {0.207f, 0.372f, 0.553f},
     {0.164f, 0.471f, 0.558f}, {0.128f, 0.567f, 0.551f}, {0.135f, 0.659f, 0.518f}, {0.267f, 0.749f, 0.441f},
     {0.478f, 0.821f, 0.318f}, {0.741f, 0.873f, 0.150f}, {0.993f, 0.906f, 0.144f}},
    {{0.001f, 0.000f, 0.014f}, {0.078f, 0.044f, 0.214f}, {0.225f, 0.036f, 0.388f}, {0.373f, 0.074f, 0.432f},
     {0.522f, 0.128f, 0.420f}, {0.665f, 0.182f, 0.370f}, {0.798f, 0.280f, 0.270f}, {0.902f, 0.413f, 0.144f},
     {0.969f, 0.579f, 0.032f}, {0.988f, 0.762f, 0.135f}, {0.988f, 0.998f, 0.645f}},
    {{0.30f, 0.55f, 1.00f}, {0.22f, 0.42f, 0.85f}, {0.15f, 0.30f, 0.66f}, {0.09f, 0.19f, 0.45f},
     {0.04f, 0.09f, 0.24f}, {0.02f, 0.02f, 0.03f}, {0.26f, 0.09f, 0.04f}, {0.48f, 0.17f, 0.06f},
     {0.72f, 0.29f, 0.08f}, {0.92f, 0.48f, 0.14f}, {1.00f, 0.72f, 0.32f}},
};

__device__ inline F3 ramp(float t, int which) {
    const float x = clampf(t, 0.0f, 1.0f) * 10.0f;
    const int i0 = (int)floorf(x);
    const int i1 = min(i0 + 1, 10);
    const float f = x - floorf(x);
    const int w = which == 0 ? 0 : (which == 1 ? 1 : 2);
    const float *a = kRamps[w][i0], *b = kRamps[w][i1];
    return mix3(f3(a[0], a[1], a[2]), f3(b[0], b[1], b[2]), f);
}

__device__ inline float safe_inv(float x) {
    const float s = x < 0.0f ? -1.0f : 1.0f;
    return s / fmaxf(fabsf(x), 1e-9f);
}

struct F4 {
    float r, g, b, a;
};

__device__ inline float half_to_float(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// The view uniform and the texels of one frame.
struct Frame {
    bs_render_view V;
    const uint2 *tex;  // one rgba16f texel (8 bytes) per cell
    int nx, ny, nz;
};

__device__ inline F4 texel(const Frame &F, int x, int y, int z) {
    const uint2 t = F.tex[((int64_t)z * F.ny + y) * F.nx + x];
    return F4{half_to_float((uint16_t)(t.x & 0xffffu)), half_to_float((uint16_t)(t.x >> 16)),
              half_to_float((uint16_t)(t.y & 0xffffu)), half_to_float((uint16_t)(t.y >> 16))};
}
__device__ inline F4 mix4(F4 a, F4 b, float t) {
    return F4{mixf(a.r, b.r, t), mixf(a.g, b.g, t), mixf(a.b, b.b, t), mixf(a.a, b.a, t)};
}
__device__ inline void corner(float u, int n, int &i0, int &i1, float &w) {
    const float c = u * (float)n - 0.5f;
    const float f = floorf(c);
    w = c - f;
    const int i = (int)f;
    i0 = min(max(i, 0), n - 1);
    i1 = min(max(i + 1, 0), n - 1);
}
// textureSampleLevel(vis_tex, vis_samp, to_texcoord(p), 0) (volume.wgsl:122-128)
__device__ inline F4 sample_state(const Frame &F, F3 p) {
    const F3 uvw = (p / f3(F.V.box_half[0], F.V.box_half[1], F.V.box_half[2])) * 0.5f + f3(0.5f, 0.5f, 0.5f);
    int x0, x1, y0, y1, z0, z1;
    float wx, wy, wz;
    corner(uvw.x, F.nx, x0, x1, wx);
    corner(uvw.y, F.ny, y0, y1, wy);
    corner(uvw.z, F.nz, z0, z1, wz);
    const F4 c00 = mix4(texel(F, x0, y0, z0), texel(F, x1, y0, z0), wx);
    const F4 c10 = mix4(texel(F, x0, y1, z0), texel(F, x1, y1, z0), wx);
    const F4 c01 = mix4(texel(F, x0, y0, z1), texel(F, x1, y0, z1), wx);
    const F4 c11 = mix4(texel(F, x0, y1, z1), texel(F, x1, y1, z1), wx);
    return mix4(mix4(c00, c10, wy), mix4(c01, c11, wy), wz);
}

// gradient of the order parameter in world units (volume.wgsl:131-141)
__device__ inline F3 grad_phi(const Frame &F, F3 p) {
    const float h = F.V.box_half[3];
    const F3 dx = f3(h, 0.0f, 0.0f), dy = f3(0.0f, h, 0.0f), dz = f3(0.0f, 0.0f, h);
    return f3(sample_state(F, p + dx).r - sample_state(F, p - dx).r,
              sample_state(F, p + dy).r - sample_state(F, p - dy).r,
              sample_state(F, p + dz).r - sample_state(F, p - dz).r) /
           (2.0f * h);
}

__device__ inline float hash12(float px, float py) {
    F3 p3 = f3(fractf(px * 0.1031f), fractf(py * 0.1031f), fractf(px * 0.1031f));
    const float d = dot(p3, f3(p3.y, p3.z, p3.x) + f3(33.33f, 33.33f, 33.33f));
    p3 = p3 + f3(d, d, d);
    return fractf((p3.x + p3.y) * p3.z);
}

// map_field (volume.wgsl:156-186): ramp position, intensity, ramp
__device__ inline void map_field(const Frame &F, const F4 &s, float &t, float &intensity, int &ramp_id) {
    const int mode = (int)F.V.render1[0];
    const float gain = F.V.render2[0];
    ramp_id = (int)F.V.render1[3];
    if (mode == 0) {
        const float x = clampf(s.r, 0.0f, 1.0f);
        t = x;
        intensity = x * gain;
    } else if (mode == 1) {
        const float x = clampf(s.g * gain, 0.0f, 1.0f);
        t = x;
        intensity = x;
    } else if (mode == 2) {
        const float x = clampf(s.b * gain, -1.0f, 1.0f);
        t = 0.5f + 0.5f * x;
        intensity = fabsf(x);
        ramp_id = 2;
    } else {
        const float x = clampf(s.a * gain, 0.0f, 1.0f);
        t = x;
        intensity = x;
    }
}

__device__ inline uint8_t unorm8(float c) { return (uint8_t)(clampf(c, 0.0f, 1.0f) * 255.0f + 0.5f); }

}  // namespace

__global__ void __launch_bounds__(64) k_render_volume(Frame F, uint32_t width, uint32_t height, uchar4 *out) {
#pragma clang fp contract(off)
    const uint32_t px = blockIdx.x * 8 + threadIdx.x % 8;
    const uint32_t py = blockIdx.y * 8 + threadIdx.x / 8;
    if (px >= width || py >= height) return;
    const bs_render_view &V = F.V;
    // fragment position (pixel centre) and the interpolated NDC of the
    // full-screen triangle (volume.wgsl:188-199; NDC y points up, rows down)
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float ndc_x = fx / (float)width * 2.0f - 1.0f;
    const float ndc_y = 1.0f - fy / (float)height * 2.0f;
    const F3 ro = f3(V.cam_pos[0], V.cam_pos[1], V.cam_pos[2]);
    const float *M = V.inv_view_proj;  // column-major
    const float far_x = M[0] * ndc_x + M[4] * ndc_y + M[8] * 1.0f + M[12] * 1.0f;
    const float far_y = M[1] * ndc_x + M[5] * ndc_y + M[9] * 1.0f + M[13] * 1.0f;
    const float far_z = M[2] * ndc_x + M[6] * ndc_y + M[10] * 1.0f + M[14] * 1.0f;
    const float far_w = M[3] * ndc_x + M[7] * ndc_y + M[11] * 1.0f + M[15] * 1.0f;
    const F3 rd = normalize(f3(far_x, far_y, far_z) / far_w - ro);

    const float horizon = clampf(rd.y * 0.5f + 0.5f, 0.0f, 1.0f);
    F3 color = mix3(f3(0.015f, 0.017f, 0.028f), f3(0.045f, 0.055f, 0.085f), horizon) * V.render3[2];

    const F3 half = f3(V.box_half[0], V.box_half[1], V.box_half[2]);
    const F3 inv_rd = f3(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
    const F3 t0 = (neg(half) - ro) * inv_rd;
    const F3 t1 = (half - ro) * inv_rd;
    const F3 tmin = fmin3(t0, t1), tmax = fmax3(t0, t1);
    float t_near = fmaxf(fmaxf(fmaxf(tmin.x, tmin.y), tmin.z), 0.0f);
    float t_far = fminf(fminf(tmax.x, tmax.y), tmax.z);
    bool hit = t_near < t_far;

    // cutaway against an axis-aligned half-space (volume.wgsl:223-245)
    const int clip_axis = (int)V.render1[1];
    if (hit && clip_axis >= 0) {
        const F3 n_axis = f3(clip_axis == 0 ? 1.0f : 0.0f, clip_axis == 1 ? 1.0f : 0.0f, clip_axis == 2 ? 1.0f : 0.0f);
        const float plane = V.render1[2] * V.box_half[clip_axis];
        const float denom = dot(rd, n_axis);
        const float dist = plane - dot(ro, n_axis);
        if (fabsf(denom) < 1e-6f) {
            if (dist < 0.0f) hit = false;
        } else {
            const float t_plane = dist / denom;
            if (denom > 0.0f) t_far = fminf(t_far, t_plane);
            else t_near = fmaxf(t_near, t_plane);
            if (t_near >= t_far) hit = false;
        }
    }

    if (hit) {
        // box wireframe (volume.wgsl:247-259)
        if (V.render3[1] > 0.5f) {
            const float w = 0.004f * V.box_half[3] * V.grid[0];
            float edge = 0.0f;
            for (int k = 0; k < 2; ++k) {
                const F3 p = ro + rd * (k == 0 ? t_near : t_far);
                const F3 d = half - fabs3(p);
                const float median = fmaxf(fminf(d.x, d.y), fminf(fmaxf(d.x, d.y), d.z));
                edge = fmaxf(edge, 1.0f - smoothstepf(0.0f, w, median));
            }
            color = color + f3(0.20f, 0.26f, 0.34f) * edge;
        }

        // march (volume.wgsl:261-369)
        const float cell = V.box_half[3];
        const float samples_per_cell = fmaxf(V.render0[2], 0.25f);
        float ds = cell / samples_per_cell;
        const float span = t_far - t_near;
        if (span / ds > (float)kMaxSteps) ds = span / (float)kMaxSteps;
        const int n_steps = min((int)ceilf(span / ds), kMaxSteps);
        const float jitter = hash12(fx + V.render3[0], fy + V.render3[0]);
        const float iso = V.render0[0];
        const float wall_opacity = V.render0[1];
        const float exposure = V.render0[3];
        const float absorption = V.render2[1];
        const bool show_iso = V.render2[2] > 0.5f;
        const bool show_vol = V.render2[3] > 0.5f;

        F3 accum = f3(0.0f, 0.0f, 0.0f);
        float alpha = 0.0f;
        float t = t_near + jitter * ds;
        float prev_phi = sample_state(F, ro + rd * t_near).r;
        float prev_t = t_near;
        for (int i = 0; i < n_steps; ++i) {
            if (t > t_far || alpha > 0.995f) break;
            const F3 p = ro + rd * t;
            const F4 s = sample_state(F, p);
            // bubble wall isosurface
            const float d_prev = prev_phi - iso;
            const float d_cur = s.r - iso;
            if (show_iso && d_prev != 0.0f && d_prev * d_cur <= 0.0f) {
                const float denom = s.r - prev_phi;
                const float frac = fabsf(denom) > 1e-6f ? (iso - prev_phi) / denom : 0.5f;
                const float t_hit = mixf(prev_t, t, clampf(frac, 0.0f, 1.0f));
                const F3 p_hit = ro + rd * t_hit;
                const F3 g = grad_phi(F, p_hit);
                const F3 n = normalize(neg(g) - rd * 1e-6f);
                const F4 hs = sample_state(F, p_hit);
                const float drive = clampf(hs.g * 4.0f, 0.0f, 1.0f);
                const F3 base = mix3(f3(0.16f, 0.62f, 0.72f), f3(1.00f, 0.72f, 0.28f), drive);
                const F3 l1 = normalize(f3(0.55f, 0.72f, 0.42f));
                const F3 l2 = normalize(f3(-0.5f, -0.2f, -0.7f));
                const F3 view = neg(rd);
                const float diff = fmaxf(dot(n, l1), 0.0f) + 0.35f * fmaxf(dot(n, l2), 0.0f);
                const F3 h = normalize(l1 + view + f3(0.0f, 1e-5f, 0.0f));
                const float spec = powf(fmaxf(dot(n, h), 0.0f), 48.0f);
                const float fres = powf(1.0f - fabsf(dot(n, view)), 3.0f);
                const F3 shaded = base * (0.22f + 0.85f * diff) + f3(1.0f, 1.0f, 1.0f) * spec * 0.35f +
                                  base * fres * 0.9f;
                const float a = wall_opacity;
                accum = accum + (shaded * (1.0f - alpha)) * a;
                alpha = alpha + (1.0f - alpha) * a;
            }
            // emissive plasma
            if (show_vol) {
                float mt, intensity;
                int ramp_id;
                map_field(F, s, mt, intensity, ramp_id);
                if (intensity > 1e-4f) {
                    const float weight = intensity * intensity;
                    const float a = 1.0f - expf(-absorption * weight * ds);
                    const F3 emit = ramp(mt, ramp_id) * exposure;
                    accum = accum + (emit * (1.0f - alpha)) * a;
                    alpha = alpha + (1.0f - alpha) * a;
                }
            }
            prev_phi = s.r;
            prev_t = t;
            t = t + ds;
        }
        color = color * (1.0f - alpha) + accum;
        // filmic tone map and gamma (volume.wgsl:372-377); a ray that misses
        // the box or is clipped away returns the background before them
        color = color / (f3(1.0f, 1.0f, 1.0f) + color);
        color = f3(powf(fmaxf(color.x, 0.0f), 1.0f / 2.2f), powf(fmaxf(color.y, 0.0f), 1.0f / 2.2f),
                   powf(fmaxf(color.z, 0.0f), 1.0f / 2.2f));
    }
    out[(size_t)py * width + px] = make_uchar4(unorm8(color.x), unorm8(color.y), unorm8(color.z), 255);
}

hipError_t launch_render_volume(const void *texels, const uint32_t grid[3], const bs_render_view &view,
                                uint32_t width, uint32_t height, void *rgba_out, hipStream_t st) {
    Frame F;
    F.V = view;
    F.tex = (const uint2 *)texels;
    F.nx = (int)grid[0];
    F.ny = (int)grid[1];
    F.nz = (int)grid[2];
    hipLaunchKernelGGL(k_render_volume, dim3((width + 7) / 8, (height + 7) / 8), dim3(64), 0, st, F, width, height,
                       (uchar4 *)rgba_out);
    return hipGetLastError();
}

}  // namespace bs
