// kernels_fast.hip -- FastMathF32 instantiations of the stage kernels.
// Built with -ffp-contract=fast-honor-pragmas: the compiler fuses
// multiply-adds except in the RK stage combination (scheme.h, stage_update).
// Only the tiles with one halo item per thread are built (tile_built in
// launch.h).
#include "launch.h"

namespace bs {

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

template <int TX, int TY, class E = BagEos>
static hipError_t fused_one(const Box<float> &box, int stage, int zc, hipStream_t st) {
    const int ntx = (box.P.nx + TX - 1) / TX;
    const int nty = (box.P.ny + TY - 1) / TY;
    const int nzc = (box.P.nz + zc - 1) / zc;
    const int nblocks = ntx * nty * nzc;
    hipLaunchKernelGGL((stage == 0 ? k_stage_fused<TX, TY, FastMathF32, float, true, E>
                                   : k_stage_fused<TX, TY, FastMathF32, float, false, E>),
                       dim3(nblocks), dim3(TX * TY), 0, st, box, stage, ntx, nty, zc);
    return hipGetLastError();
}

template <int TX, int TY, class E = BagEos>
static hipError_t fused_multi_one(const Box<float> *dboxes, const BoxEvents *devents, const Task *dtasks, int ntasks, int stage, hipStream_t st) {
    hipLaunchKernelGGL((stage == 0 ? k_stage_fused_multi<TX, TY, FastMathF32, float, true, E>
                                   : k_stage_fused_multi<TX, TY, FastMathF32, float, false, E>),
                       dim3(ntasks), dim3(TX * TY), 0, st, dboxes, devents, dtasks, stage, ntasks);
    return hipGetLastError();
}

hipError_t launch_stage_fused_fast(const Box<float> &box, int stage, TileShape t, int zc, hipStream_t st) {
    if (box.P.eos == kEosGammaLaw) {
        if (tile_index(t) == 3) return fused_one<16, 16, GammaLawEos>(box, stage, zc, st);
        return fused_one<32, 8, GammaLawEos>(box, stage, zc, st);
    }
    switch (tile_index(t)) {
    case 0: return fused_one<32, 16>(box, stage, zc, st);
    case 1: return fused_one<64, 8>(box, stage, zc, st);
    case 2: return fused_one<32, 8>(box, stage, zc, st);
    case 3: return fused_one<16, 16>(box, stage, zc, st);
    case 5: return fused_one<16, 8>(box, stage, zc, st);
    case 7: return fused_one<16, 15>(box, stage, zc, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stage_fused_multi_fast(int eos, const Box<float> *dboxes, const BoxEvents *devents, const Task *dtasks,
                                         int ntasks, int stage, TileShape t, hipStream_t st) {
    if (eos == kEosGammaLaw) {
        switch (tile_index(t)) {
        case 2: return fused_multi_one<32, 8, GammaLawEos>(dboxes, devents, dtasks, ntasks, stage, st);
        case 3: return fused_multi_one<16, 16, GammaLawEos>(dboxes, devents, dtasks, ntasks, stage, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (tile_index(t)) {
    case 0: return fused_multi_one<32, 16>(dboxes, devents, dtasks, ntasks, stage, st);
    case 1: return fused_multi_one<64, 8>(dboxes, devents, dtasks, ntasks, stage, st);
    case 2: return fused_multi_one<32, 8>(dboxes, devents, dtasks, ntasks, stage, st);
    case 3: return fused_multi_one<16, 16>(dboxes, devents, dtasks, ntasks, stage, st);
    case 5: return fused_multi_one<16, 8>(dboxes, devents, dtasks, ntasks, stage, st);
    case 7: return fused_multi_one<16, 15>(dboxes, devents, dtasks, ntasks, stage, st);
    default: return hipErrorInvalidValue;
    }
}

template <class E>
static hipError_t twopass_fast(const Box<float> &box, int stage, V4<float> *prim, hipStream_t st) {
    const int64_t n = (int64_t)box.P.nx * box.P.ny * box.P.nz;
    int in, out;
    stage_slot_ids(box.cur, stage, in, out);
    float a0, a1;
    RK3c<float>::get(stage, a0, a1);
    hipLaunchKernelGGL((k_prim<FastMathF32, float, E>), dim3(grid_for(n, 256)), dim3(256), 0, st, box.field[in],
                       box.fluid[in], prim, box.P);
    hipLaunchKernelGGL((k_stage_twopass<FastMathF32, float, E>), dim3(grid_for(n, 256)), dim3(256), 0, st,
                       box.field[box.cur], box.fluid[box.cur], box.field[in], box.fluid[in], prim, box.field[out],
                       box.fluid[out], box.P, a0, a1);
    return hipGetLastError();
}
hipError_t launch_stage_twopass_fast(const Box<float> &box, int stage, V4<float> *prim, hipStream_t st) {
    if (box.P.eos == kEosGammaLaw) return twopass_fast<GammaLawEos>(box, stage, prim, st);
    return twopass_fast<BagEos>(box, stage, prim, st);
}

}  // namespace bs
