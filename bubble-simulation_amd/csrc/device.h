// device.h -- device-side state shared by the single-box simulation, the
// batch and the server: one box's three RK slots and its parameter block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "bubblesim.h"
#include "internal.h"
#include "launch.h"

namespace bs {

#define BS_HIP_TRY(expr)                                                                          \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::bs::set_error(e_ == hipErrorOutOfMemory ? BS_ERR_OOM : BS_ERR_DEVICE,        \
                                   std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

// f32 parameter block -> lattice-precision Params (exact widening for f64)
template <class R> inline Params<R> widen_params(const bs_sim_params &p) {
    Params<R> q;
    q.nx = (int)p.nx; q.ny = (int)p.ny; q.nz = (int)p.nz; q.eos = kEosBag;
    q.dx = (R)p.dx; q.inv_dx = (R)p.inv_dx; q.dt = (R)p.dt; q.eta = (R)p.eta;
    q.lambda = (R)p.lambda; q.eps = (R)p.eps; q.phi_b = (R)p.phi_b; q.m2 = (R)p.m2;
    q.delta = (R)p.delta; q.a_rad = (R)p.a_rad; q.wall_width = (R)p.wall_width; q.p_floor = (R)p.p_floor;
    q.e_floor = (R)p.e_floor; q.z_max = (R)p.z_max; q.e_ref = (R)p.e_ref; q.t_ref = (R)p.t_ref;
    q.eos_gamma = (R)(4.0f / 3.0f);
    q.dof_ratio = R(1);
    return q;
}

inline size_t real_size(int precision) { return precision == BS_PRECISION_F64 ? 8 : 4; }

// Three RK slots of one box in device memory.
struct BoxState {
    int precision = BS_PRECISION_F32;
    uint32_t grid[3] = {0, 0, 0};
    uint64_t n_cells = 0;
    void *field[3] = {nullptr, nullptr, nullptr};
    void *fluid[3] = {nullptr, nullptr, nullptr};
    bs_sim_params params{};
    int cur = 0;
    uint64_t bytes = 0;
    bool async_alloc = false;
    int eos = BS_EOS_BAG;               // equation of state (bs_spec.eos)
    float eos_gamma = 4.0f / 3.0f;
    float dof_ratio = 1.0f;
    void set_eos(const bs_spec &s) { eos = s.eos; eos_gamma = s.eos_gamma; dof_ratio = s.dof_ratio; }

    // allocate the slots (stream-ordered when `stream` is given)
    int alloc(int prec, const uint32_t g[3], hipStream_t stream, bool async);
    void release(hipStream_t stream);
    int init(hipStream_t stream);  // init.wgsl on slot cur
    template <class R> Box<R> box() const {
        Box<R> b;
        b.P = widen_params<R>(params);
        b.P.eos = eos == BS_EOS_GAMMA_LAW ? kEosGammaLaw : kEosBag;
        b.P.eos_gamma = (R)eos_gamma;
        b.P.dof_ratio = (R)dof_ratio;
        for (int s = 0; s < 3; ++s) {
            b.field[s] = (V2<R> *)field[s];
            b.fluid[s] = (V4<R> *)fluid[s];
        }
        b.cur = cur;
        b.pad[0] = b.pad[1] = b.pad[2] = 0;
        return b;
    }
};

// stamp a list of bubbles (nucleate.wgsl) on slot `cur`, in batches
int stamp_bubbles(BoxState &bx, const std::vector<bs_event> &events, hipStream_t stream);

// one RK stage of a single box (fused or two-pass)
int run_stage(BoxState &bx, int stage, int math, int kernel, TileShape tile, int zc, void *prim, hipStream_t stream);

// choose the z chunk for a box and tile so the launch has enough workgroups
int default_zc(const uint32_t g[3], TileShape t);

// FP64 compensated diagnostics of slot `cur` (blocking)
struct ReduceScratch {
    ReducePartial *partials = nullptr;
    ReduceResult *dres = nullptr;
    ReduceResult *hres = nullptr;  // pinned
    int alloc();
    void release();
};
int launch_diagnostics(BoxState &bx, ReduceScratch &rs, hipStream_t stream);
void diagnostics_from_result(const ReduceResult &r, uint64_t n_cells, bs_diagnostics *out);

// visualisation pack (vis.wgsl) of slot `cur` into `out` (8 B per cell)
int launch_vis(BoxState &bx, Half4 *out, hipStream_t stream);

// FNV-1a over the downloaded state of slot cur (blocking)
int state_checksum(BoxState &bx, hipStream_t stream, uint64_t *out);
int download_state(BoxState &bx, hipStream_t stream, void *field, void *fluid);

// tile shape of a launch: BS_TILE ("32x16") when set, else `dflt`;
// BS_ERR_INVALID when BS_TILE names a shape not built for this arithmetic
int tile_from_env(TileShape dflt, int math, int precision, TileShape *out);

}  // namespace bs
