// kernels_exact.hip -- ExactMath instantiations (f32 and f64) and the
// precision-independent kernels.  Built with -ffp-contract=off: every
// multiply and add is rounded on its own, as in the reference's expressions.
#include "launch.h"

namespace bs {

int tile_index(TileShape t) {
    for (int i = 0; i < kNumTiles; ++i)
        if (kTiles[i].tx == t.tx && kTiles[i].ty == t.ty) return i;
    return -1;
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

template <class R> hipError_t launch_init(V2<R> *field, V4<R> *fluid, int64_t n, R e_ref, hipStream_t st) {
    hipLaunchKernelGGL(k_init<R>, dim3(grid_for(n, 256)), dim3(256), 0, st, field, fluid, n, e_ref);
    return hipGetLastError();
}

template <class R> hipError_t launch_stamp(V2<R> *field, const Params<R> &P, const StampBatch<R> &b, hipStream_t st) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(k_stamp<R>, dim3(grid_for(n, 256)), dim3(256), 0, st, field, P, b);
    return hipGetLastError();
}

template <class R>
hipError_t launch_vis_pack(const V2<R> *field, const V4<R> *fluid, const Params<R> &P, Half4 *out, hipStream_t st) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    if (P.eos == kEosGammaLaw)
        hipLaunchKernelGGL((k_vis_pack<R, GammaLawEos>), dim3(grid_for(n, 256)), dim3(256), 0, st, field, fluid, P, out);
    else
        hipLaunchKernelGGL((k_vis_pack<R, BagEos>), dim3(grid_for(n, 256)), dim3(256), 0, st, field, fluid, P, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(64) k_reduce2(const ReducePartial *__restrict__ partials, ReduceResult *__restrict__ out) {
    // one wave folds the kReduceBlocks partials: lane l takes l, l+64, ... in
    // order, then a fixed xor-butterfly merges the lanes
    CSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double vmax = 0.0, nonfinite = 0.0;
    for (int i = threadIdx.x; i < kReduceBlocks; i += 64) {
        const ReducePartial &p = partials[i];
        for (int k = 0; k < 4; ++k) acc[k] = csum_merge(acc[k], p.v[k]);
        vmax = p.vmax > vmax ? p.vmax : vmax;
        nonfinite = p.nonfinite > nonfinite ? p.nonfinite : nonfinite;
    }
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
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) out->sums[k] = acc[k].s + acc[k].c;
        out->vmax = vmax;
        out->nonfinite = nonfinite;
    }
}

template <class R>
hipError_t launch_reduce(const V2<R> *field, const V4<R> *fluid, const Params<R> &P, ReducePartial *partials,
                         ReduceResult *out, hipStream_t st) {
    if (P.eos == kEosGammaLaw)
        hipLaunchKernelGGL((k_reduce1<R, GammaLawEos>), dim3(kReduceBlocks), dim3(kReduceThreads), 0, st, field, fluid, P,
                           partials);
    else
        hipLaunchKernelGGL((k_reduce1<R, BagEos>), dim3(kReduceBlocks), dim3(kReduceThreads), 0, st, field, fluid, P,
                           partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(64), 0, st, partials, out);
    return hipGetLastError();
}

template <int TX, int TY, class M, class R, class E>
static hipError_t fused_one(const Box<R> &box, int stage, int zc, hipStream_t st) {
    const int ntx = (box.P.nx + TX - 1) / TX;
    const int nty = (box.P.ny + TY - 1) / TY;
    const int nzc = (box.P.nz + zc - 1) / zc;
    const int nblocks = ntx * nty * nzc;
    hipLaunchKernelGGL((stage == 0 ? k_stage_fused<TX, TY, M, R, true, E> : k_stage_fused<TX, TY, M, R, false, E>),
                       dim3(nblocks), dim3(TX * TY), 0, st, box, stage, ntx, nty, zc);
    return hipGetLastError();
}

template <int TX, int TY, class M, class R, class E = BagEos>
static hipError_t fused_multi_one(const Box<R> *dboxes, const BoxEvents *devents, const Task *dtasks, int ntasks, int stage, hipStream_t st) {
    hipLaunchKernelGGL((stage == 0 ? k_stage_fused_multi<TX, TY, M, R, true, E>
                                   : k_stage_fused_multi<TX, TY, M, R, false, E>),
                       dim3(ntasks), dim3(TX * TY), 0, st, dboxes, devents, dtasks, stage, ntasks);
    return hipGetLastError();
}

// The bag model is instantiated for every tile shape; the gamma-law EOS for
// the two default shapes (32x8, 16x16), other requests run on 32x8.
template <class M, class R>
static hipError_t fused_dispatch(const Box<R> &box, int stage, TileShape t, int zc, hipStream_t st) {
    if (box.P.eos == kEosGammaLaw) {
        if (tile_index(t) == 3) return fused_one<16, 16, M, R, GammaLawEos>(box, stage, zc, st);
        return fused_one<32, 8, M, R, GammaLawEos>(box, stage, zc, st);
    }
    switch (tile_index(t)) {
    case 0: return fused_one<32, 16, M, R, BagEos>(box, stage, zc, st);
    case 1: return fused_one<64, 8, M, R, BagEos>(box, stage, zc, st);
    case 2: return fused_one<32, 8, M, R, BagEos>(box, stage, zc, st);
    case 3: return fused_one<16, 16, M, R, BagEos>(box, stage, zc, st);
    case 4: return fused_one<64, 4, M, R, BagEos>(box, stage, zc, st);
    case 5: return fused_one<16, 8, M, R, BagEos>(box, stage, zc, st);
    case 6: return fused_one<32, 4, M, R, BagEos>(box, stage, zc, st);
    case 7: return fused_one<16, 15, M, R, BagEos>(box, stage, zc, st);
    default: return hipErrorInvalidValue;
    }
}
template <class M, class R>
static hipError_t fused_multi_dispatch(int eos, const Box<R> *dboxes, const BoxEvents *devents, const Task *dtasks, int ntasks,
                                       int stage, TileShape t, hipStream_t st) {
    if (eos == kEosGammaLaw) {
        switch (tile_index(t)) {
        case 2: return fused_multi_one<32, 8, M, R, GammaLawEos>(dboxes, devents, dtasks, ntasks, stage, st);
        case 3: return fused_multi_one<16, 16, M, R, GammaLawEos>(dboxes, devents, dtasks, ntasks, stage, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (tile_index(t)) {
    case 0: return fused_multi_one<32, 16, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 1: return fused_multi_one<64, 8, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 2: return fused_multi_one<32, 8, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 3: return fused_multi_one<16, 16, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 4: return fused_multi_one<64, 4, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 5: return fused_multi_one<16, 8, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 6: return fused_multi_one<32, 4, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    case 7: return fused_multi_one<16, 15, M, R>(dboxes, devents, dtasks, ntasks, stage, st);
    default: return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_stage_fused<float>(int math, const Box<float> &box, int stage, TileShape t, int zc, hipStream_t st) {
    if (math == BS_MATH_FAST) return launch_stage_fused_fast(box, stage, t, zc, st);
    return fused_dispatch<ExactMath<float>, float>(box, stage, t, zc, st);
}
template <>
hipError_t launch_stage_fused<double>(int, const Box<double> &box, int stage, TileShape t, int zc, hipStream_t st) {
    return fused_dispatch<ExactMath<double>, double>(box, stage, t, zc, st);
}
template <>
hipError_t launch_stage_fused_multi<float>(int math, int eos, const Box<float> *dboxes, const BoxEvents *devents,
                                           const Task *dtasks, int ntasks, int stage, TileShape t, hipStream_t st) {
    if (math == BS_MATH_FAST) return launch_stage_fused_multi_fast(eos, dboxes, devents, dtasks, ntasks, stage, t, st);
    return fused_multi_dispatch<ExactMath<float>, float>(eos, dboxes, devents, dtasks, ntasks, stage, t, st);
}
template <>
hipError_t launch_stage_fused_multi<double>(int, int eos, const Box<double> *dboxes, const BoxEvents *devents,
                                            const Task *dtasks, int ntasks, int stage, TileShape t, hipStream_t st) {
    return fused_multi_dispatch<ExactMath<double>, double>(eos, dboxes, devents, dtasks, ntasks, stage, t, st);
}

template <class M, class R, class E>
static hipError_t twopass_e(const Box<R> &box, int stage, V4<R> *prim, hipStream_t st) {
    const int64_t n = (int64_t)box.P.nx * box.P.ny * box.P.nz;
    int in, out;
    stage_slot_ids(box.cur, stage, in, out);
    R a0, a1;
    RK3c<R>::get(stage, a0, a1);
    hipLaunchKernelGGL((k_prim<M, R, E>), dim3(grid_for(n, 256)), dim3(256), 0, st, box.field[in], box.fluid[in], prim,
                       box.P);
    hipLaunchKernelGGL((k_stage_twopass<M, R, E>), dim3(grid_for(n, 256)), dim3(256), 0, st, box.field[box.cur],
                       box.fluid[box.cur], box.field[in], box.fluid[in], prim, box.field[out], box.fluid[out], box.P, a0,
                       a1);
    return hipGetLastError();
}
template <class M, class R>
static hipError_t twopass(const Box<R> &box, int stage, V4<R> *prim, hipStream_t st) {
    if (box.P.eos == kEosGammaLaw) return twopass_e<M, R, GammaLawEos>(box, stage, prim, st);
    return twopass_e<M, R, BagEos>(box, stage, prim, st);
}
template <>
hipError_t launch_stage_twopass<float>(int math, const Box<float> &box, int stage, V4<float> *prim, hipStream_t st) {
    if (math == BS_MATH_FAST) return launch_stage_twopass_fast(box, stage, prim, st);
    return twopass<ExactMath<float>, float>(box, stage, prim, st);
}
template <>
hipError_t launch_stage_twopass<double>(int, const Box<double> &box, int stage, V4<double> *prim, hipStream_t st) {
    return twopass<ExactMath<double>, double>(box, stage, prim, st);
}

// ---------------------------------------------------------------------------
//  Batch bookkeeping: device-driven nucleation and the per-step advance, so
//  a captured step graph can be replayed without host involvement.
// ---------------------------------------------------------------------------
// Nucleation of a batch: the blocks walk the boxes in order and stamp the
// events due before the current step of each box, every block taking its
// share of that box's cells, so a step on which no box has an event due
// costs one pass over the box table per block.
constexpr int kStampBlocks = 256;

template <class R>
__global__ void __launch_bounds__(256) k_stamp_multi(Box<R> *boxes, const BoxEvents *events, int nboxes) {
    for (int bi = 0; bi < nboxes; ++bi) {
        const BoxEvents &ev = events[bi];
        if (!ev.active || ev.steps_done >= ev.steps_target) continue;
        if (ev.next >= ev.count || ev.step[ev.next] > ev.steps_done) continue;
        int last = ev.next;
        while (last < ev.count && ev.step[last] <= ev.steps_done) ++last;
        const Box<R> &b = boxes[bi];
        const Params<R> &P = b.P;
        V2<R> *field = b.field[b.cur];
        const int64_t n = (int64_t)P.nx * P.ny * P.nz;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kStampBlocks * 256) {
            const int x = (int)(i % P.nx);
            const int y = (int)((i / P.nx) % P.ny);
            const int z = (int)(i / ((int64_t)P.nx * P.ny));
            V2<R> f = field[i];
            R phi = f.x;
            for (int k = ev.next; k < last; ++k) {
                const float *bb = ev.bubble + 4 * k;
                const R prof = bubble_profile<R>(x, y, z, (R)bb[0], (R)bb[1], (R)bb[2], (R)bb[3], P);
                phi = ExactMath<R>::max(phi, prof);
            }
            field[i] = V2<R>{phi, f.y};
        }
    }
}

template <class R> __global__ void k_advance_multi(Box<R> *boxes, BoxEvents *events, int nboxes) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nboxes) return;
    BoxEvents &ev = events[bi];
    if (!ev.active || ev.steps_done >= ev.steps_target) return;
    while (ev.next < ev.count && ev.step[ev.next] <= ev.steps_done) ++ev.next;
    ev.steps_done += 1;
    boxes[bi].cur = (boxes[bi].cur + 1) % 3;
}

template <class R> hipError_t launch_stamp_multi(Box<R> *dboxes, BoxEvents *devents, int nboxes, hipStream_t st) {
    hipLaunchKernelGGL(k_stamp_multi<R>, dim3(kStampBlocks), dim3(256), 0, st, dboxes, devents, nboxes);
    return hipGetLastError();
}

// Non-finite check of a set of boxes: kFiniteBlocksPerBox blocks per box
// stream slot cur (24 B per cell in f32); any block that meets a NaN or
// infinity writes the tick into the box's flag (the same value from every
// writer, so the plain stores need no atomics).
constexpr int kFiniteBlocksPerBox = 32;

template <class R>
__global__ void __launch_bounds__(256) k_finite_multi(const Box<R> *boxes, int nboxes, uint32_t *flags, uint32_t tick) {
    const int bi = blockIdx.x / kFiniteBlocksPerBox;
    const int part = blockIdx.x % kFiniteBlocksPerBox;
    if (bi >= nboxes) return;
    const Box<R> &b = boxes[bi];
    const V2<R> *field = b.field[b.cur];
    const V4<R> *fluid = b.fluid[b.cur];
    const int64_t n = (int64_t)b.P.nx * b.P.ny * b.P.nz;
    bool bad = false;
    for (int64_t i = (int64_t)part * 256 + threadIdx.x; i < n; i += (int64_t)kFiniteBlocksPerBox * 256) {
        const V2<R> f = field[i];
        const V4<R> u = fluid[i];
        const bool ok = isfinite(f.x) && isfinite(f.y) && isfinite(u.x) && isfinite(u.y) && isfinite(u.z) &&
                        isfinite(u.w);
        bad = bad || !ok;
    }
    if (bad) flags[bi] = tick;
}

template <class R>
hipError_t launch_finite_multi(const Box<R> *dboxes, int nboxes, uint32_t *flags, uint32_t tick, hipStream_t st) {
    hipLaunchKernelGGL(k_finite_multi<R>, dim3(nboxes * kFiniteBlocksPerBox), dim3(256), 0, st, dboxes, nboxes, flags,
                       tick);
    return hipGetLastError();
}
template <class R> hipError_t launch_advance_multi(Box<R> *dboxes, BoxEvents *devents, int nboxes, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_multi<R>, dim3((nboxes + 63) / 64), dim3(64), 0, st, dboxes, devents, nboxes);
    return hipGetLastError();
}

template hipError_t launch_init<float>(V2<float> *, V4<float> *, int64_t, float, hipStream_t);
template hipError_t launch_init<double>(V2<double> *, V4<double> *, int64_t, double, hipStream_t);
template hipError_t launch_stamp<float>(V2<float> *, const Params<float> &, const StampBatch<float> &, hipStream_t);
template hipError_t launch_vis_pack<float>(const V2<float> *, const V4<float> *, const Params<float> &, Half4 *,
                                           hipStream_t);
template hipError_t launch_vis_pack<double>(const V2<double> *, const V4<double> *, const Params<double> &, Half4 *,
                                            hipStream_t);
template hipError_t launch_stamp<double>(V2<double> *, const Params<double> &, const StampBatch<double> &, hipStream_t);
template hipError_t launch_reduce<float>(const V2<float> *, const V4<float> *, const Params<float> &, ReducePartial *,
                                         ReduceResult *, hipStream_t);
template hipError_t launch_reduce<double>(const V2<double> *, const V4<double> *, const Params<double> &,
                                          ReducePartial *, ReduceResult *, hipStream_t);
template hipError_t launch_stamp_multi<float>(Box<float> *, BoxEvents *, int, hipStream_t);
template hipError_t launch_stamp_multi<double>(Box<double> *, BoxEvents *, int, hipStream_t);
template hipError_t launch_advance_multi<float>(Box<float> *, BoxEvents *, int, hipStream_t);
template hipError_t launch_advance_multi<double>(Box<double> *, BoxEvents *, int, hipStream_t);
template hipError_t launch_finite_multi<float>(const Box<float> *, int, uint32_t *, uint32_t, hipStream_t);
template hipError_t launch_finite_multi<double>(const Box<double> *, int, uint32_t *, uint32_t, hipStream_t);

}  // namespace bs
