// kernels.h -- HIP kernels of the lattice solver for gfx950 (wave64).
//
// Kernel set (reference entry point each replaces, shaders/*.wgsl):
//   k_init            init.wgsl:18-26
//   k_stamp           nucleate.wgsl:59-85 (bubble batch passed by value)
//   k_prim            primitives.wgsl:16-23        (two-pass path only)
//   k_stage_twopass   step.wgsl:92-174             (two-pass path only)
//   k_stage_fused     primitives + step fused, one launch per RK stage
//   k_reduce1/2       reduce.wgsl:81-140, FP64 compensated
//
// State is AoS: field V2 (phi, pi), fluid V4 (E, Zx, Zy, Zz), x fastest,
// i = (z*ny + y)*nx + x (common.wgsl:91-95).
#pragma once

#include "scheme.h"


namespace bs {

// One lattice box as the kernels see it.  `field[s]`, `fluid[s]` are the
// three RK slots; `cur` is the slot holding U^n (sim.rs:180-242 `cur`).
template <class R> struct Box {
    Params<R> P;
    V2<R> *field[3];
    V4<R> *fluid[3];
    int cur;
    int pad[3];
};

// stage_slots (sim.rs:1046-1051) for stage s: a=cur, b=cur+1, c=cur+2;
// (in, out) = (a,b), (b,c), (c,b); U^n is always a.
__host__ __device__ inline void stage_slot_ids(int cur, int s, int &in, int &out) {
    const int a = cur, b = (cur + 1) % 3, c = (cur + 2) % 3;
    in = s == 0 ? a : (s == 1 ? b : c);
    out = s == 1 ? c : b;
}

// The value that lane l - 1 of the wave holds (0 in lane 0): DPP wave_shr:1.
__device__ __forceinline__ float lane_from_left(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ V4<float> lane_from_left(const V4<float> &v) {
    return V4<float>{lane_from_left(v.x), lane_from_left(v.y), lane_from_left(v.z), lane_from_left(v.w)};
}

template <class R> struct RK3c;
template <> struct RK3c<float> {
    // RK3 (sim.rs:135)
    __host__ __device__ static inline void get(int s, float &a0, float &a1) {
        a0 = s == 0 ? 0.0f : (s == 1 ? 0.75f : 1.0f / 3.0f);
        a1 = s == 0 ? 1.0f : (s == 1 ? 0.25f : 2.0f / 3.0f);
    }
};
template <> struct RK3c<double> {
    // the f32 coefficients, widened (the FP64 scheme shares every constant
    // that the reference stores as f32)
    __host__ __device__ static inline void get(int s, double &a0, double &a1) {
        a0 = s == 0 ? 0.0 : (s == 1 ? 0.75 : (double)(1.0f / 3.0f));
        a1 = s == 0 ? 1.0 : (s == 1 ? 0.25 : (double)(2.0f / 3.0f));
    }
};

// ---------------------------------------------------------------------------
//  init (init.wgsl:18-26)
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(256) k_init(V2<R> *field, V4<R> *fluid, int64_t n, R e_ref) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    field[i] = V2<R>{R(0), R(0)};
    fluid[i] = V4<R>{e_ref, R(0), R(0), R(0)};
}

// ---------------------------------------------------------------------------
//  nucleate (nucleate.wgsl:59-85): phi = max(phi, profile_b) over a batch of
//  bubbles; pi untouched.
// ---------------------------------------------------------------------------
constexpr int kStampBatch = 32;
template <class R> struct StampBatch {
    int count;
    int pad[3];
    R b[kStampBatch][4];  // cx, cy, cz, r0 (f32 values, widened for R = double)
};

template <class R>
__global__ void __launch_bounds__(256) k_stamp(V2<R> *field, Params<R> P, StampBatch<R> batch) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % P.nx);
    const int y = (int)((i / P.nx) % P.ny);
    const int z = (int)(i / ((int64_t)P.nx * P.ny));
    V2<R> f = field[i];
    R phi = f.x;
    for (int b = 0; b < batch.count; ++b) {
        const R prof = bubble_profile<R>(x, y, z, batch.b[b][0], batch.b[b][1], batch.b[b][2], batch.b[b][3], P);
        phi = ExactMath<R>::max(phi, prof);
    }
    field[i] = V2<R>{phi, f.y};
}

// ---------------------------------------------------------------------------
//  Visualisation pack (vis.wgsl:32-55, after the primitives pass that
//  Simulation::record_visualisation runs first, sim.rs:920-942): one
//  rgba16f texel per cell, x fastest,
//    r = phi / phi_b, g = |v|, b = T / T_ref - 1, a = (E + p) |v|^2 / e_ref.
//  Channels are formed in the lattice precision, narrowed to f32 (the
//  reference's f32(...)) and stored as IEEE half, round to nearest even.
// ---------------------------------------------------------------------------
struct alignas(8) Half4 {
    uint16_t r, g, b, a;
};
__device__ inline uint16_t half_bits(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

template <class R, class E>
__global__ void __launch_bounds__(256) k_vis_pack(const V2<R> *__restrict__ field, const V4<R> *__restrict__ fluid,
                                                  Params<R> P, Half4 *__restrict__ out) {
    typedef ExactMath<R> M;
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R phi = field[i].x;
    const V4<R> u = fluid[i];
    const V4<R> q = eos_prim_from_cons<M, R, E>(u, phi, P);  // primitives.wgsl:21
    const R e_lab = M::max(u.x, P.e_floor);
    const R v2 = dot3(q.x, q.y, q.z, q.x, q.y, q.z);
    const R speed = M::sqrt(v2);
    const R temp = eos_temperature<M>(q.w, phi, P);
    const R kinetic = (e_lab + q.w) * v2;
    out[i] = Half4{half_bits((float)(phi / P.phi_b)), half_bits((float)speed),
                   half_bits((float)(temp / P.t_ref - R(1))), half_bits((float)(kinetic / P.e_ref))};
}

// ---------------------------------------------------------------------------
//  Two-pass path: primitives (primitives.wgsl:16-23), then the stage
//  (step.wgsl:92-174) reading the primitive cache.  Mirrors the reference's
//  dispatch structure; faces are evaluated from both sides as there.
// ---------------------------------------------------------------------------
template <class M, class R, class E>
__global__ void __launch_bounds__(256) k_prim(const V2<R> *__restrict__ field, const V4<R> *__restrict__ fluid,
                                              V4<R> *__restrict__ prim, Params<R> P) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    prim[i] = eos_prim_from_cons<M, R, E>(fluid[i], field[i].x, P);
}

template <int A, class M, class R, class E>
__device__ inline void twopass_axis(const V2<R> *__restrict__ fin, const V4<R> *__restrict__ prim, const Params<R> &P,
                                    int x, int y, int z, R phi0, const V4<R> &q0, R &lap_sum, R &g, V4<R> &flux_div) {
    auto idx = [&](int dd) -> int64_t {
        const int xx = A == 0 ? wrapi(x + dd, P.nx) : x;
        const int yy = A == 1 ? wrapi(y + dd, P.ny) : y;
        const int zz = A == 2 ? wrapi(z + dd, P.nz) : z;
        return ((int64_t)zz * P.ny + yy) * P.nx + xx;
    };
    const int64_t im1 = idx(-1), ip1 = idx(1), im2 = idx(-2), ip2 = idx(2);
    const R phi_m = fin[im1].x;
    const R phi_p = fin[ip1].x;
    lap_sum = lap_sum + (phi_m + phi_p - R(2) * phi0);
    g = R(0.5) * (phi_p - phi_m) * P.inv_dx;
    const V4<R> qm2 = sanitise<M>(prim[im2], P);
    const V4<R> qm1 = sanitise<M>(prim[im1], P);
    const V4<R> qp1 = sanitise<M>(prim[ip1], P);
    const V4<R> qp2 = sanitise<M>(prim[ip2], P);
    const V4<R> h_m1 = half_slope4<M>(sub4(qm1, qm2), sub4(q0, qm1));
    const V4<R> h_0 = half_slope4<M>(sub4(q0, qm1), sub4(qp1, q0));
    const V4<R> h_p1 = half_slope4<M>(sub4(qp1, q0), sub4(qp2, qp1));
    const V4<R> f_minus = face_flux_h<A, M, R, E>(qm1, h_m1, q0, h_0, phi_m, phi0, P);
    const V4<R> f_plus = face_flux_h<A, M, R, E>(q0, h_0, qp1, h_p1, phi0, phi_p, P);
    flux_div.x = flux_div.x + (f_plus.x - f_minus.x) * P.inv_dx;
    flux_div.y = flux_div.y + (f_plus.y - f_minus.y) * P.inv_dx;
    flux_div.z = flux_div.z + (f_plus.z - f_minus.z) * P.inv_dx;
    flux_div.w = flux_div.w + (f_plus.w - f_minus.w) * P.inv_dx;
}

template <class M, class R, class E>
__global__ void __launch_bounds__(256) k_stage_twopass(const V2<R> *__restrict__ field_n, const V4<R> *__restrict__ fluid_n,
                                                       const V2<R> *__restrict__ field_in, const V4<R> *__restrict__ fluid_in,
                                                       const V4<R> *__restrict__ prim, V2<R> *__restrict__ field_out,
                                                       V4<R> *__restrict__ fluid_out, Params<R> P, R a0, R a1) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 >= n) return;
    const int x = (int)(i0 % P.nx);
    const int y = (int)((i0 / P.nx) % P.ny);
    const int z = (int)(i0 / ((int64_t)P.nx * P.ny));
    const V2<R> f0 = field_in[i0];
    const R phi0 = f0.x, pi0 = f0.y;
    const V4<R> q0 = sanitise<M>(prim[i0], P);
    R lap_sum = R(0), gx, gy, gz;
    V4<R> flux_div{R(0), R(0), R(0), R(0)};
    twopass_axis<0, M, R, E>(field_in, prim, P, x, y, z, phi0, q0, lap_sum, gx, flux_div);
    twopass_axis<1, M, R, E>(field_in, prim, P, x, y, z, phi0, q0, lap_sum, gy, flux_div);
    twopass_axis<2, M, R, E>(field_in, prim, P, x, y, z, phi0, q0, lap_sum, gz, flux_div);
    const V2<R> fnn = field_n[i0];
    R phi_o, pi_o;
    V4<R> fo;
    stage_update<M>(phi0, pi0, q0, lap_sum, gx, gy, gz, flux_div, fluid_in[i0], fnn.x, fnn.y, fluid_n[i0], a0, a1,
                    P.dt, P, phi_o, pi_o, fo);
    field_out[i0] = V2<R>{phi_o, pi_o};
    fluid_out[i0] = fo;
}

// ---------------------------------------------------------------------------
//  Fused stage: one workgroup owns a TX x TY column of cells over a z range
//  [zb, ze) and marches up in z.  Per plane:
//    - the sanitised primitives of the plane, with a 2-cell halo in x and y,
//      sit in LDS (interior ones come from registers, where they were
//      computed two planes earlier; halo ones from a thread-private LDS
//      ring, recovered two planes earlier together with the column load of
//      that plane, while the neighbouring tiles load the same plane);
//    - every x- and y-face of the tile is evaluated once (one face per
//      thread and axis, then one round for the far-edge faces) and stored
//      in LDS;
//    - each thread evaluates its own +z face from a register ring of the
//      column's primitives q(z..z+2), half slope(z) and carries the -z face;
//    - the divergence, friction and RK combination are applied per cell.
//  The result is bitwise that of k_prim + k_stage_twopass under ExactMath:
//  every face sees the same operands in the same order.
// ---------------------------------------------------------------------------
template <int TX, int TY, class R> struct FusedLds {
    V4<R> q[TY + 4][TX + 4];  // sanitised primitives of plane z, 2-cell halo
    R phi[TY + 2][TX + 2];    // phi of plane z, 1-cell halo
    V4<R> fx[TY][TX + 1];     // flux through the face left of tile column c
    V4<R> fy[TY + 1][TX];     // flux through the face below tile row r
    V4<R> hq[3][4 * (TX + TY)];  // halo primitives of planes z .. z+2 (each slot private to its thread)
    R hph[3][4 * (TX + TY)];     // their phi
};

template <class R> __device__ inline int64_t cidx(const Params<R> &P, int x, int y, int z) {
    return ((int64_t)z * P.ny + y) * P.nx + x;
}

// A pointer that is the same in every lane of the wave, moved into scalar
// registers and tagged as global memory.  The compiler cannot see through
// readfirstlane, so it keeps the z-march's plane bases as scalar values
// instead of rewriting every access into a per-lane 64-bit pointer advanced
// by a vector add each plane.
template <class T> using global_t = __attribute__((address_space(1))) T;
__device__ __forceinline__ global_t<char> *uniform_ptr(const void *p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return reinterpret_cast<global_t<char> *>(((uint64_t)hi << 32) | lo);
}

// Element `idx` of a plane (the plane's base is the same in every lane),
// through a raw buffer resource: the plane's base and size sit in four
// scalar registers (rebuilt per plane by a few scalar instructions) and each
// lane supplies only a 32-bit byte offset that is fixed for the whole
// z-march, so no vector instruction is spent on addresses inside the loop
// (a global pointer with a scalar base spent 8 64-bit vector adds per plane:
// 0.5 % slower at 512^3, 7 % with the gamma-law EOS).  Spec validation keeps
// an x-y plane below 4 GiB.  Offsets at or past the plane's size read zero
// and drop stores, the hardware's range check.  Loads and stores move one
// dwordx2 / dwordx4 per element; the streaming forms carry cache policy 2,
// the non-temporal hint, for data used once per stage -- U^n and the output
// slot -- so they do not evict from L2 the U_in planes that the halo reads
// come back to (measured +3 % at 512^3).
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kRawBufferWord3 = 0x00020000;  // raw 32-bit data, no swizzle
constexpr int kCacheDefault = 0, kCacheStream = 2;

template <class E> struct PlaneRef {
    __amdgpu_buffer_rsrc_t r;
};
template <class E> __device__ __forceinline__ PlaneRef<E> plane_ref(const E *base, unsigned elems) {
    return PlaneRef<E>{__builtin_amdgcn_make_buffer_rsrc((void *)uniform_ptr(base), 0, (int)(elems * (unsigned)sizeof(E)),
                                                         kRawBufferWord3)};
}
template <int AUX> __device__ __forceinline__ u32x4 buf_load16(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, AUX);
}
template <int AUX, class E> __device__ __forceinline__ E buf_load(const PlaneRef<E> &p, unsigned idx) {
    const unsigned off = idx * (unsigned)sizeof(E);
    if constexpr (sizeof(E) == 8) {
        return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(p.r, (int)off, 0, AUX));
    } else if constexpr (sizeof(E) == 16) {
        return __builtin_bit_cast(E, buf_load16<AUX>(p.r, off));
    } else {
        static_assert(sizeof(E) == 32, "plane elements are 8, 16 or 32 bytes");
        struct W { u32x4 a, b; };
        return __builtin_bit_cast(E, W{buf_load16<AUX>(p.r, off), buf_load16<AUX>(p.r, off + 16)});
    }
}
template <int AUX, class E> __device__ __forceinline__ void buf_store(const PlaneRef<E> &p, unsigned idx, const E &v) {
    const unsigned off = idx * (unsigned)sizeof(E);
    if constexpr (sizeof(E) == 8) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), p.r, (int)off, 0, AUX);
    } else if constexpr (sizeof(E) == 16) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), p.r, (int)off, 0, AUX);
    } else {
        static_assert(sizeof(E) == 32, "plane elements are 8, 16 or 32 bytes");
        struct W { u32x4 a, b; };
        const W w = __builtin_bit_cast(W, v);
        __builtin_amdgcn_raw_buffer_store_b128(w.a, p.r, (int)off, 0, AUX);
        __builtin_amdgcn_raw_buffer_store_b128(w.b, p.r, (int)(off + 16), 0, AUX);
    }
}
template <class E> __device__ __forceinline__ E plane_load(const PlaneRef<E> &p, unsigned idx) {
    return buf_load<kCacheDefault>(p, idx);
}
template <class E> __device__ __forceinline__ E stream_load(const PlaneRef<E> &p, unsigned idx) {
    return buf_load<kCacheStream>(p, idx);
}
template <class E> __device__ __forceinline__ void stream_store(const PlaneRef<E> &p, unsigned idx, const E &v) {
    buf_store<kCacheStream>(p, idx, v);
}

template <int TX, int TY, class M, class R, bool S0, class E>
__device__ __forceinline__ void fused_tile(FusedLds<TX, TY, R> &L, const Params<R> &P,
                                           const V2<R> *__restrict__ field_n, const V4<R> *__restrict__ fluid_n,
                                           const V2<R> *__restrict__ field_in, const V4<R> *__restrict__ fluid_in,
                                           V2<R> *__restrict__ field_out, V4<R> *__restrict__ fluid_out,
                                           int x0, int y0, int zb, int ze, int stage) {
    constexpr int NT = TX * TY;
    const int tid = threadIdx.x;
    // halo items are numbered from the second wave on, so the first wave,
    // which also evaluates the far-edge faces, carries no halo item (the
    // waves of a block meet at two barriers per plane, so the heaviest
    // wave sets the pace; measured 0.6 % faster at 512^3)
    const int htid = NT % 64 == 0 && NT > 64 ? (tid >= 64 ? tid - 64 : tid - 64 + NT) : tid;
    const int tx = tid % TX;
    const int ty = tid / TX;
    const int gx = wrapi(x0 + tx, P.nx);
    const int gy = wrapi(y0 + ty, P.ny);
    const bool valid = (x0 + tx < P.nx) && (y0 + ty < P.ny);
    R a0, a1;
    RK3c<R>::get(stage, a0, a1);
    const R dt = P.dt;

    // offsets within a z plane, fixed for the whole march
    const int64_t plane = (int64_t)P.nx * P.ny;
    const unsigned col = (unsigned)(gy * P.nx + gx);
    // halo items of a plane: 4 TY x-halo cells, then 4 TX y-halo cells
    constexpr int NH = 4 * (TX + TY);
    constexpr int HPT = (NH + NT - 1) / NT;
    unsigned hoff[HPT];
    int hlds[HPT];  // (hy + 2) * (TX + 4) + (hx + 2), or -1
    int hphi[HPT];  // (hy + 1) * (TX + 2) + (hx + 1) for the 1-cell halo, or -1
#pragma unroll
    for (int k = 0; k < HPT; ++k) {
        const int h = htid + k * NT;
        hlds[k] = -1;
        hphi[k] = -1;
        hoff[k] = 0;
        if (h < NH) {
            int hx, hy;
            if (h < 4 * TY) {
                const int r = h / 4, c = h % 4;
                hy = r;
                hx = c < 2 ? c - 2 : TX + (c - 2);
            } else {
                const int k2 = h - 4 * TY;
                const int r = k2 / TX, c = k2 % TX;
                hx = c;
                hy = r < 2 ? r - 2 : TY + (r - 2);
            }
            hoff[k] = (unsigned)(wrapi(y0 + hy, P.ny) * P.nx + wrapi(x0 + hx, P.nx));
            hlds[k] = (hy + 2) * (TX + 4) + (hx + 2);
            if (hx >= -1 && hx <= TX && hy >= -1 && hy <= TY) hphi[k] = (hy + 1) * (TX + 2) + (hx + 1);
        }
    }
    // faces: round 1 gives every thread the x face left of its cell, round 2
    // the y face below it; the TY + TX faces on the far edges (x = TX, y = TY)
    // form one more round, one uniform code path for both directions.  Under
    // ExactMath the y faces are evaluated as x faces of the component-swapped
    // states (vx <-> vy): the swap commutes with every operation of the face
    // (the sums of squares add the same two products in either order), so
    // the result is the y face's, bit for bit.  FastMath takes the direction
    // as a per-lane value instead (a select for the normal velocity and one
    // for the pressure term, where the swap moved every component).
    // FastMath shares each cell's x half slope with the face on its left:
    // the cell forms q - h (the right state of its left face) and q + h (the
    // left state of its right face), and the lane to its right takes the
    // latter over DPP, so each x slope is formed once instead of twice.  The
    // faces at x = 0 have no lane on their left and join the far-edge round.
    constexpr bool XSHARE = M::fast;
    constexpr int NTAIL = TX + TY + (XSHARE ? TY : 0);
    const bool tail = tid < NTAIL;

    // sanitised primitives of a cell.  eos_prim_from_cons already floors p
    // and caps |Z| <= z_max E, so |v| = |Z| / (E + p) < z_max: sanitise is
    // the identity up to rounding, and FastMath leaves it out.
    auto prim_of = [&](const V2<R> &f, const V4<R> &u) -> V4<R> {
        if constexpr (M::fast) return eos_prim_from_cons<M, R, E>(u, f.x, P);
        else return sanitise<M>(eos_prim_from_cons<M, R, E>(u, f.x, P), P);
    };

    // ---- prologue: column state for zb-2 .. zb+1, the face zb-1/2 ----
    // (phi of plane z-1, U_in of planes z, z+1 and the primitives of z, z+1
    // stay in registers, so every input plane is read from memory once per
    // column).  Each ring has three slots, addressed by the plane's phase
    // (z - zb) mod 3.
    V2<R> fr[3], ftmp;  // field of planes z, z+1, z+2 (slot I2 holds z-1 until z+2 arrives)
    V4<R> ur[3], utmp;  // U_in of the same planes
    V4<R> qr[3];        // their primitives
    V4<R> hr[3];        // z half slopes of planes z, z+1
    V4<R> fzr[3];       // z faces z-1/2 and z+1/2
    auto field_at = [&](const V2<R> *a, int zz) { return plane_ref(a + (int64_t)zz * plane, (unsigned)plane); };
    auto fluid_at = [&](const V4<R> *a, int zz) { return plane_ref(a + (int64_t)zz * plane, (unsigned)plane); };
    ftmp = plane_load(field_at(field_in, wrapi(zb - 2, P.nz)), col);
    utmp = plane_load(fluid_at(fluid_in, wrapi(zb - 2, P.nz)), col);
    const V4<R> qa = prim_of(ftmp, utmp);
    ftmp = plane_load(field_at(field_in, wrapi(zb - 1, P.nz)), col);
    utmp = plane_load(fluid_at(fluid_in, wrapi(zb - 1, P.nz)), col);
    const V4<R> qb = prim_of(ftmp, utmp);
    fr[0] = plane_load(field_at(field_in, wrapi(zb, P.nz)), col);
    ur[0] = plane_load(fluid_at(fluid_in, wrapi(zb, P.nz)), col);
    qr[0] = prim_of(fr[0], ur[0]);
    fr[1] = plane_load(field_at(field_in, wrapi(zb + 1, P.nz)), col);
    ur[1] = plane_load(fluid_at(fluid_in, wrapi(zb + 1, P.nz)), col);
    qr[1] = prim_of(fr[1], ur[1]);
    const V4<R> h_b = half_slope4<M>(sub4(qb, qa), sub4(qr[0], qb));
    hr[0] = half_slope4<M>(sub4(qr[0], qb), sub4(qr[1], qr[0]));
    fzr[0] = face_flux_h<2, M, R, E>(qb, h_b, qr[0], hr[0], ftmp.x, fr[0].x, P);
    fr[2] = ftmp;  // plane zb-1: its phi is the first plane's phi(z-1)

    V4<R> *qflat = &L.q[0][0];
    R *phiflat = &L.phi[0][0];
    // halo cells are recovered two planes ahead, together with the column
    // load of the same plane, and wait in a thread-private LDS ring
    V2<R> hfr[HPT];  // raw halo states of the plane being fetched
    V4<R> hur[HPT];
    auto halo_load = [&](int zz) {
        const PlaneRef<V2<R>> fp = field_at(field_in, zz);
        const PlaneRef<V4<R>> up = fluid_at(fluid_in, zz);
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hlds[k] >= 0) {
                hfr[k] = plane_load(fp, hoff[k]);
                hur[k] = plane_load(up, hoff[k]);
            }
        }
    };
    auto halo_store = [&](int slot) {
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hlds[k] >= 0) {
                const int h = htid + k * NT;
                L.hq[slot][h] = prim_of(hfr[k], hur[k]);
                L.hph[slot][h] = hfr[k].x;
            }
        }
    };
    auto halo_ahead = [&](int zz, int slot) {
        halo_load(zz);
        halo_store(slot);
    };
    halo_ahead(zb, 0);
    if (zb + 1 < ze) halo_ahead(wrapi(zb + 1, P.nz), 1);

    // One plane of the march.  The column rings hold planes z, z+1, z+2 in
    // slots I0, I1, I2 = (z - zb) mod 3 and its successors, compile-time
    // constants in the march below (unrolled by three), so that advancing
    // the rings renames registers instead of moving them; I0 is also the
    // halo ring slot of plane z.
    auto march = [&](const int z, auto phase) {
        constexpr int I0 = decltype(phase)::value, I1 = (I0 + 1) % 3, I2 = (I0 + 2) % 3;
        const int zp2 = z + 2 < P.nz ? z + 2 : z + 2 - P.nz;
        const R ph_m = fr[I2].x;  // plane z-1, before plane z+2 takes its slot
        const V2<R> f0 = fr[I0], f1 = fr[I1];
        const V4<R> qz = qr[I0], qz1 = qr[I1], hz = hr[I0], fzm = fzr[I0];
        // the update's operand U^n(z) of this cell, loaded first so that it
        // arrives while the faces are computed (U_in(z) is in the ring)
        const R pi0 = f0.y;
        const V4<R> uin = ur[I0];
        V2<R> fnn;
        V4<R> un;
        if constexpr (S0) {
            fnn = V2<R>{f0.x, pi0};  // U_in is U^n at stage 0 (same slot)
            un = uin;
        } else {
            fnn = stream_load(field_at(field_n, z), col);
            un = stream_load(fluid_at(fluid_n, z), col);
        }
        // (a) plane z+2 of this column (for the +z face)
        fr[I2] = plane_load(field_at(field_in, zp2), col);
        ur[I2] = plane_load(fluid_at(fluid_in, zp2), col);
        const V4<R> q2 = prim_of(fr[I2], ur[I2]);
        qr[I2] = q2;
        // halo of plane z+2: loads issued here, recovered after the x and y
        // faces, so the loads are in flight across the LDS fill, the barrier
        // and the faces (a barrier waits for LDS, not for memory)
        const bool halo_next = z + 2 < ze;
        if (halo_next) halo_load(zp2);

        // (b) plane z into LDS: interior from registers, halo from the ring
        L.q[ty + 2][tx + 2] = qz;
        L.phi[ty + 1][tx + 1] = f0.x;
#pragma unroll
        for (int k = 0; k < HPT; ++k) {
            if (hlds[k] >= 0) {
                const int h = htid + k * NT;
                qflat[hlds[k]] = L.hq[I0][h];
                if (hphi[k] >= 0) phiflat[hphi[k]] = L.hph[I0][h];
            }
        }
        __syncthreads();

        // (c) x and y faces, slopes formed inside each face from four cells
        if constexpr (XSHARE) {
            const V4<R> h = half_slope4<M>(sub4(qz, L.q[ty + 2][tx + 1]), sub4(L.q[ty + 2][tx + 3], qz));
            R v2r, v2l;
            V4<R> right = add4(qz, h);  // left state of the face at x + 1/2
            V4<R> b = sub4(qz, h);      // right state of the face at x - 1/2
            sanitise_pair_v2<M>(right, v2r, b, v2l, P);
            const V4<R> a = lane_from_left(right);
            const V4<R> fl = hll_flux_v2<0, M, R, E>(a, lane_from_left(v2r), L.phi[ty + 1][tx], b, v2l, f0.x, P);
            if (tx != 0) L.fx[ty][tx] = fl;
        } else {
            L.fx[ty][tx] = face_flux<0, M, R, E>(L.q[ty + 2][tx], L.q[ty + 2][tx + 1], qz, L.q[ty + 2][tx + 3],
                                                 L.phi[ty + 1][tx], f0.x, P);
        }
        L.fy[ty][tx] = face_flux<1, M, R, E>(L.q[ty][tx + 2], L.q[ty + 1][tx + 2], qz, L.q[ty + 3][tx + 2],
                                             L.phi[ty][tx + 1], f0.x, P);
        if (tail) {
            // far-edge face `tid`: x faces x = TX of rows 0..TY-1, y faces
            // y = TY of columns 0..TX-1, then (XSHARE) x faces x = 0 of rows
            // 0..TY-1 (addresses formed here, in the wave that runs this
            // round, to keep them out of registers)
            const bool ya = tid >= TY && tid < TY + TX;
            const bool xl = tid >= TY + TX;
            const int c = ya ? tid - TY : (xl ? tid - (TY + TX) : tid);
            const int xe = xl ? 0 : TX;
            const int q0 = ya ? TY * (TX + 4) + (c + 2) : (c + 2) * (TX + 4) + xe;
            const int qs = ya ? TX + 4 : 1;
            const int p0 = ya ? TY * (TX + 2) + (c + 1) : (c + 1) * (TX + 2) + xe;
            const int ps = ya ? TX + 2 : 1;
            V4<R> c0 = qflat[q0], c1 = qflat[q0 + qs], c2 = qflat[q0 + 2 * qs], c3 = qflat[q0 + 3 * qs];
            V4<R> ft;
            if constexpr (M::fast) {
                // the face direction as a per-lane value: the normal velocity
                // and the pressure term take a select each
                ft = face_flux_fast_xy<M, R, E>(c0, c1, c2, c3, phiflat[p0], phiflat[p0 + ps], P, ya ? 1 : 0);
            } else {
                if (ya) {
                    c0 = swap_xy(c0); c1 = swap_xy(c1); c2 = swap_xy(c2); c3 = swap_xy(c3);
                }
                ft = face_flux<0, M, R, E>(c0, c1, c2, c3, phiflat[p0], phiflat[p0 + ps], P);
                if (ya) ft = swap_yz(ft);
            }
            if (ya) {
                L.fy[TY][c] = ft;
            } else {
                L.fx[c][xe] = ft;
            }
        }
        if (halo_next) halo_store(I2);
        // own +z face
        const V4<R> hz1 = half_slope4<M>(sub4(qz1, qz), sub4(q2, qz1));
        const V4<R> fzp = face_flux_h<2, M, R, E>(qz, hz, qz1, hz1, f0.x, f1.x, P);
        hr[I1] = hz1;
        fzr[I1] = fzp;

        // field stencil from LDS / registers (step.wgsl:118-122)
        const R pxm = L.phi[ty + 1][tx], pxp = L.phi[ty + 1][tx + 2];
        const R pym = L.phi[ty][tx + 1], pyp = L.phi[ty + 2][tx + 1];
        __syncthreads();

        // (d) divergence, sources, RK combination
        if (valid) {
            const R ph_0 = f0.x, ph_1 = f1.x;
            R lap_sum, gxv, gyv, gzv;
            if constexpr (M::fast) {
                // the same sums regrouped; the central differences are
                // scaled by 1/(2 dx) inside the update
                lap_sum = ((pxm + pxp) + (pym + pyp)) + (ph_m + ph_1) - R(6) * ph_0;
                gxv = pxp - pxm;
                gyv = pyp - pym;
                gzv = ph_1 - ph_m;
            } else {
                // the reference accumulates from zero (step.wgsl:117-145)
                lap_sum = R(0) + (pxm + pxp - R(2) * ph_0);
                gxv = R(0.5) * (pxp - pxm) * P.inv_dx;
                lap_sum = lap_sum + (pym + pyp - R(2) * ph_0);
                gyv = R(0.5) * (pyp - pym) * P.inv_dx;
                lap_sum = lap_sum + (ph_m + ph_1 - R(2) * ph_0);
                gzv = R(0.5) * (ph_1 - ph_m) * P.inv_dx;
            }
            const V4<R> fxm = L.fx[ty][tx], fxp = L.fx[ty][tx + 1];
            const V4<R> fym = L.fy[ty][tx], fyp = L.fy[ty + 1][tx];
            R phi_o, pi_o;
            V4<R> fo;
            if constexpr (M::fast) {
                // the three flux differences summed unscaled; 1/dx is applied
                // with dt in the update
                const V4<R> fs{(fxp.x - fxm.x) + (fyp.x - fym.x) + (fzp.x - fzm.x),
                               (fxp.y - fxm.y) + (fyp.y - fym.y) + (fzp.y - fzm.y),
                               (fxp.z - fxm.z) + (fyp.z - fym.z) + (fzp.z - fzm.z),
                               (fxp.w - fxm.w) + (fyp.w - fym.w) + (fzp.w - fzm.w)};
                stage_update<M, R, S0, true>(ph_0, pi0, qz, lap_sum, gxv, gyv, gzv, fs, uin, fnn.x, fnn.y, un, a0, a1,
                                             dt, P, phi_o, pi_o, fo);
            } else {
                V4<R> fd{(fxp.x - fxm.x) * P.inv_dx, (fxp.y - fxm.y) * P.inv_dx, (fxp.z - fxm.z) * P.inv_dx,
                         (fxp.w - fxm.w) * P.inv_dx};
                fd = V4<R>{R(0) + fd.x, R(0) + fd.y, R(0) + fd.z, R(0) + fd.w};
                fd.x = fd.x + (fyp.x - fym.x) * P.inv_dx;
                fd.y = fd.y + (fyp.y - fym.y) * P.inv_dx;
                fd.z = fd.z + (fyp.z - fym.z) * P.inv_dx;
                fd.w = fd.w + (fyp.w - fym.w) * P.inv_dx;
                fd.x = fd.x + (fzp.x - fzm.x) * P.inv_dx;
                fd.y = fd.y + (fzp.y - fzm.y) * P.inv_dx;
                fd.z = fd.z + (fzp.z - fzm.z) * P.inv_dx;
                fd.w = fd.w + (fzp.w - fzm.w) * P.inv_dx;
                stage_update<M, R, S0>(ph_0, pi0, qz, lap_sum, gxv, gyv, gzv, fd, uin, fnn.x, fnn.y, un, a0, a1, dt, P,
                                       phi_o, pi_o, fo);
            }
            stream_store(field_at(field_out, z), col, V2<R>{phi_o, pi_o});
            stream_store(fluid_at(fluid_out, z), col, fo);
        }

    };
    int z = zb;
    for (; z + 3 <= ze; z += 3) {
        march(z, std::integral_constant<int, 0>{});
        march(z + 1, std::integral_constant<int, 1>{});
        march(z + 2, std::integral_constant<int, 2>{});
    }
    if (z < ze) march(z, std::integral_constant<int, 0>{});
    if (z + 1 < ze) march(z + 1, std::integral_constant<int, 1>{});
}

// Minimum waves per SIMD the register allocator must leave room for: the
// fast f32 stage, with its march unrolled by three (so the column rings
// advance by register renaming), needs 117-127 VGPRs on the tiles that
// carry one halo item per thread (every fast tile), which 4 waves allow
// without spilling (at 5 waves, 96 VGPRs, it spills); it runs about 0.5 %
// faster than the rolled march at 5 waves.  The exact and f64
// instantiations keep the compiler's choice.
template <class M> struct Occupancy {
    static constexpr int waves = 1;
};
template <> struct Occupancy<FastMathF32> {
    static constexpr int waves = 4;
};

// Single-box fused stage.  Blocks enumerate (tile_x, tile_y, z_chunk) with
// tile_x fastest, in dispatch order: consecutive blocks land on different
// XCDs (round-robin), so all eight XCDs stream neighbouring rows of the same
// planes.  The first RK stage remaps the tiles of each z chunk so that XCD
// k takes the k-th eighth of them, a band of whole tile rows: x neighbours
// then run on one XCD and their halo re-reads hit its L2, while the eight
// XCDs still stream the same planes (at 512^3 its FETCH_SIZE per launch
// drops from 7.90 to 4.15 GB, 1.29x the algorithmic read; the launch runs
// 0.7-0.8 % faster from step 300 on, as fast before).  Only tiles of 16
// columns or fewer, whose x halo is the larger share of the fetch, are
// banded: a first stage on 32x8 tiles (the ExactMath and f64 ones; measured
// with the gamma-law EOS) ran 0.2-0.4 % slower so (0.4-0.5 % in vertical
// strips), and keeps the plain order.
// Stages 2 and 3, which also stream U^n, want the eight XCDs reading
// neighbouring rows (banded, they fetch 8.94 -> 7.67 GB but run up to 1 %
// slower before step 300; a contiguous run of whole z chunks per XCD ran
// them 4 % slower): when the tile columns divide by 4 and the tile rows by
// 2, the chunk's tiles form a 4 x 2 grid of rectangles, one per XCD, and
// the eight walk their rectangles row by row together, so that most of a
// tile's x and y neighbours share its L2 while the XCDs read two bands of
// rows at a time (8.94 -> 7.48-7.55 GB per launch at 512^3 and 0.3-0.6 %
// faster, 384^3 2.2 % and 256^3 0.4-1.1 % faster; eight vertical strips
// fetched 7.91 GB and gained 0.1-0.6 %, a 2 x 4 grid ran 3-4 % slower).
template <int TX, int TY, class M, class R, bool S0, class E>
__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(Occupancy<M>::waves))) k_stage_fused(Box<R> box, int stage, int ntx, int nty, int zc) {
    __shared__ FusedLds<TX, TY, R> L;
    const int b = blockIdx.x;
    const int T = ntx * nty;
    const int zc_i = b / T;
    int r = b - zc_i * T;                                       // tile within the chunk, dispatch order
    if (S0 && TX <= 16 && (T & 7) == 0) r = (r & 7) * (T >> 3) + (r >> 3);  // XCD r & 7 takes the (r & 7)-th eighth
    int tx_i, ty_i;
    if (!S0 && (ntx & 3) == 0 && (nty & 1) == 0) {
        // the chunk's tiles as a 4 x 2 grid of rectangles: XCD k = r & 7
        // takes rectangle (k % 4, k / 4) and walks it row by row, in step
        // with the other seven
        const int k = r & 7, slot = r >> 3, wx = ntx >> 2, hy = nty >> 1;
        tx_i = (k & 3) * wx + slot % wx;
        ty_i = (k >> 2) * hy + slot / wx;
    } else {
        tx_i = r % ntx;
        ty_i = r / ntx;
    }
    const int zb = zc_i * zc;
    const int ze = min(zb + zc, box.P.nz);
    int in, out;
    stage_slot_ids(box.cur, stage, in, out);
    const int n = box.cur;
    fused_tile<TX, TY, M, R, S0, E>(L, box.P, box.field[n], box.fluid[n], box.field[in], box.fluid[in], box.field[out],
                             box.fluid[out], tx_i * TX, ty_i * TY, zb, ze, stage);
}

// Multi-box fused stage: one task per block from a task table
// (box index, x0, y0, zb, ze); each box reads its own `cur` and progress
// from device memory so a captured graph replays correctly.
struct Task {
    int box, x0, y0, zb, ze, pad[3];
};
// Device-side progress and nucleation table of one box in a batch: events
// sorted by the step before which they are stamped.
struct BoxEvents {
    const uint64_t *step;   // stamping step of each event
    const float *bubble;    // 4 floats per event: cx, cy, cz, r0
    int count;
    int next;               // first event not yet stamped
    uint64_t steps_done;    // steps completed by this box
    uint64_t steps_target;  // the box stops advancing at this count
    int active;
    int pad;
};
template <int TX, int TY, class M, class R, bool S0, class E>
__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(Occupancy<M>::waves))) k_stage_fused_multi(const Box<R> *__restrict__ boxes,
                                                              const BoxEvents *__restrict__ events,
                                                              const Task *__restrict__ tasks, int stage, int ntasks) {
    __shared__ FusedLds<TX, TY, R> L;
    int b = blockIdx.x;
    if ((ntasks & 7) == 0) b = (b & 7) * (ntasks >> 3) + (b >> 3);
    const Task t = tasks[b];
    if (t.zb >= t.ze) return;  // padding task (the list is padded to a multiple of 8)
    const BoxEvents &ev = events[t.box];
    if (!ev.active || ev.steps_done >= ev.steps_target) return;  // uniform per block
    const Box<R> &bx = boxes[t.box];
    int in, out;
    stage_slot_ids(bx.cur, stage, in, out);
    const int n = bx.cur;
    fused_tile<TX, TY, M, R, S0, E>(L, bx.P, bx.field[n], bx.fluid[n], bx.field[in], bx.fluid[in], bx.field[out],
                             bx.fluid[out], t.x0, t.y0, t.zb, t.ze, stage);
}

// ---------------------------------------------------------------------------
//  Diagnostics (reduce.wgsl:56-140), accumulated in FP64 with Neumaier
//  compensation per thread, TwoSum-combined across the wave (shuffles),
//  the block (LDS) and the fixed-order second pass.  The grid of pass 1 is
//  fixed (kReduceBlocks), so the summation order depends only on the box.
// ---------------------------------------------------------------------------
constexpr int kReduceBlocks = 512;
constexpr int kReduceThreads = 256;

struct CSum {
    double s, c;
};
__device__ inline void neumaier_add(CSum &a, double x) {
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
__device__ inline CSum csum_merge(CSum a, CSum b) {
    // TwoSum of the heads, tails added
    const double s = a.s + b.s;
    const double bb = s - a.s;
    const double err = (a.s - (s - bb)) + (b.s - bb);
    return CSum{s, a.c + b.c + err};
}
struct ReducePartial {
    CSum v[4];
    double vmax;
    double nonfinite;
};

template <class R, class E>
__global__ void __launch_bounds__(kReduceThreads) k_reduce1(const V2<R> *__restrict__ field, const V4<R> *__restrict__ fluid,
                                                            Params<R> P, ReducePartial *__restrict__ partials) {
    __shared__ ReducePartial sh[kReduceThreads / 64];
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    CSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double vmax = 0.0;
    double nonfinite = 0.0;
    const int64_t stride = (int64_t)kReduceBlocks * kReduceThreads;
    for (int64_t i = (int64_t)blockIdx.x * kReduceThreads + threadIdx.x; i < n; i += stride) {
        const int x = (int)(i % P.nx);
        const int y = (int)((i / P.nx) % P.ny);
        const int z = (int)(i / ((int64_t)P.nx * P.ny));
        const V2<R> f = field[i];
        const R pxm = field[cidx(P, wrapi(x - 1, P.nx), y, z)].x, pxp = field[cidx(P, wrapi(x + 1, P.nx), y, z)].x;
        const R pym = field[cidx(P, x, wrapi(y - 1, P.ny), z)].x, pyp = field[cidx(P, x, wrapi(y + 1, P.ny), z)].x;
        const R pzm = field[cidx(P, x, y, wrapi(z - 1, P.nz))].x, pzp = field[cidx(P, x, y, wrapi(z + 1, P.nz))].x;
        const CellDiag<ExactMath<R>, R> c =
            cell_contribution<ExactMath<R>, R, E>(f.x, f.y, pxm, pxp, pym, pyp, pzm, pzp, fluid[i], P);
        neumaier_add(acc[0], (double)c.e_total);
        neumaier_add(acc[1], (double)c.broken);
        neumaier_add(acc[2], (double)c.kinetic);
        neumaier_add(acc[3], (double)c.e_lab);
        const double va = (double)c.vabs;
        vmax = va > vmax ? va : vmax;
        if (!isfinite(c.e_total) || !isfinite(c.kinetic) || !isfinite(c.vabs)) nonfinite = 1.0;
    }
    // wave reduction
    for (int off = 32; off >= 1; off >>= 1) {
        for (int k = 0; k < 4; ++k) {
            CSum o{__shfl_xor(acc[k].s, off), __shfl_xor(acc[k].c, off)};
            acc[k] = csum_merge(acc[k], o);
        }
        const double ov = __shfl_xor(vmax, off);
        vmax = ov > vmax ? ov : vmax;
        const double on = __shfl_xor(nonfinite, off);
        nonfinite = on > nonfinite ? on : nonfinite;
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        for (int k = 0; k < 4; ++k) sh[wave].v[k] = acc[k];
        sh[wave].vmax = vmax;
        sh[wave].nonfinite = nonfinite;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ReducePartial r = sh[0];
        for (int w = 1; w < kReduceThreads / 64; ++w) {
            for (int k = 0; k < 4; ++k) r.v[k] = csum_merge(r.v[k], sh[w].v[k]);
            r.vmax = sh[w].vmax > r.vmax ? sh[w].vmax : r.vmax;
            r.nonfinite = sh[w].nonfinite > r.nonfinite ? sh[w].nonfinite : r.nonfinite;
        }
        partials[blockIdx.x] = r;
    }
}

struct ReduceResult {
    double sums[4];
    double vmax;
    double nonfinite;
};

__global__ void __launch_bounds__(64) k_reduce2(const ReducePartial *__restrict__ partials, ReduceResult *__restrict__ out);

}  // namespace bs
