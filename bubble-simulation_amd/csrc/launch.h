// launch.h -- host-callable launchers of the kernels in kernels.h.
//
// Defined in kernels_exact.hip (ExactMath, any precision; compiled without
// FMA contraction) and kernels_fast.hip (FastMathF32; compiled with
// contraction).  `math` selects between them for f32; f64 is always exact.
#pragma once

#include <hip/hip_runtime.h>

#include "bubblesim.h"
#include "kernels.h"

namespace bs {

struct TileShape {
    int tx, ty;
};

// Tile shapes the fused kernel is instantiated for (x width, y height).
// (16, 15): 255 x-faces and 256 y-faces per plane fill whole waves of 64
constexpr int kNumTiles = 8;
constexpr TileShape kTiles[kNumTiles] = {{32, 16}, {64, 8}, {32, 8}, {16, 16}, {64, 4}, {16, 8}, {32, 4}, {16, 15}};
int tile_index(TileShape t);  // -1 if not instantiated

// Halo cells each thread of a TX x TY tile loads per plane: the 4 (TX + TY)
// cells of the 2-cell cross halo spread over the TX * TY threads.
constexpr int halo_items(TileShape t) { return (4 * (t.tx + t.ty) + t.tx * t.ty - 1) / (t.tx * t.ty); }
// The FastMath kernel is built for the tiles with one halo item per thread:
// a second item (64x4, 32x4) needs 128 VGPRs and spills at the 4 waves per
// SIMD the fast kernel is compiled for.  The exact kernels take every tile.
constexpr bool tile_built(TileShape t, int math, int precision) {
    return precision == BS_PRECISION_F64 || math == BS_MATH_EXACT || halo_items(t) == 1;
}

template <class R> hipError_t launch_init(V2<R> *field, V4<R> *fluid, int64_t n, R e_ref, hipStream_t st);
template <class R> hipError_t launch_stamp(V2<R> *field, const Params<R> &P, const StampBatch<R> &b, hipStream_t st);
template <class R>
hipError_t launch_vis_pack(const V2<R> *field, const V4<R> *fluid, const Params<R> &P, Half4 *out, hipStream_t st);
template <class R>
hipError_t launch_reduce(const V2<R> *field, const V4<R> *fluid, const Params<R> &P, ReducePartial *partials,
                         ReduceResult *out, hipStream_t st);

// one fused RK stage over a whole box; zc = planes per z chunk
template <class R>
hipError_t launch_stage_fused(int math, const Box<R> &box, int stage, TileShape t, int zc, hipStream_t st);
// one fused RK stage over a task table spanning many boxes, all of one
// equation of state (kEosBag / kEosGammaLaw); the gamma-law tasks of a plan
// are cut and launched with eos_tile(kEosGammaLaw, plan tile)
template <class R>
hipError_t launch_stage_fused_multi(int math, int eos, const Box<R> *dboxes, const BoxEvents *devents, const Task *dtasks,
                                    int ntasks, int stage, TileShape t, hipStream_t st);
// The tile a launch of the given equation of state runs for a requested
// tile: the bag model is built for every tile; the gamma-law EOS for the
// two default shapes (16x16, 32x8), other requests run on 32x8.
constexpr TileShape eos_tile(int eos, TileShape t) {
    return eos != kEosGammaLaw || (t.tx == 16 && t.ty == 16) || (t.tx == 32 && t.ty == 8) ? t : TileShape{32, 8};
}
// primitive pass + stage pass (reference dispatch structure)
template <class R>
hipError_t launch_stage_twopass(int math, const Box<R> &box, int stage, V4<R> *prim, hipStream_t st);

// f32 fast-math variants (kernels_fast.hip)
hipError_t launch_stage_fused_fast(const Box<float> &box, int stage, TileShape t, int zc, hipStream_t st);
hipError_t launch_stage_fused_multi_fast(int eos, const Box<float> *dboxes, const BoxEvents *devents, const Task *dtasks,
                                         int ntasks, int stage, TileShape t, hipStream_t st);
hipError_t launch_stage_twopass_fast(const Box<float> &box, int stage, V4<float> *prim, hipStream_t st);

// batch bookkeeping kernels (kernels_exact.hip)
template <class R>
hipError_t launch_stamp_multi(Box<R> *dboxes, BoxEvents *devents, int nboxes, hipStream_t st);
template <class R> hipError_t launch_advance_multi(Box<R> *dboxes, BoxEvents *devents, int nboxes, hipStream_t st);
// volume.wgsl's ray march over rgba16f texels (render.hip): width x height
// RGBA8 pixels, row 0 at the top, into device memory
hipError_t launch_render_volume(const void *texels, const uint32_t grid[3], const bs_render_view &view,
                                uint32_t width, uint32_t height, void *rgba_out, hipStream_t st);
// flags[box] = tick for every box whose slot cur holds a non-finite value
template <class R>
hipError_t launch_finite_multi(const Box<R> *dboxes, int nboxes, uint32_t *flags, uint32_t tick, hipStream_t st);

}  // namespace bs
