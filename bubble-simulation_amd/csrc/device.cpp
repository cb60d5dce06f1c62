// device.cpp -- BoxState memory, stamping, stage dispatch and diagnostics.
#include "device.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace bs {

int BoxState::alloc(int prec, const uint32_t g[3], hipStream_t stream, bool async) {
    precision = prec;
    grid[0] = g[0]; grid[1] = g[1]; grid[2] = g[2];
    n_cells = (uint64_t)g[0] * g[1] * g[2];
    async_alloc = async;
    const size_t r = real_size(prec);
    bytes = 0;
    for (int s = 0; s < 3; ++s) {
        hipError_t e1, e2;
        if (async) {
            e1 = hipMallocAsync(&field[s], n_cells * 2 * r, stream);
            e2 = e1 == hipSuccess ? hipMallocAsync(&fluid[s], n_cells * 4 * r, stream) : e1;
        } else {
            e1 = hipMalloc(&field[s], n_cells * 2 * r);
            e2 = e1 == hipSuccess ? hipMalloc(&fluid[s], n_cells * 4 * r) : e1;
        }
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (e1 != hipSuccess) field[s] = nullptr;
            if (e2 != hipSuccess) fluid[s] = nullptr;
            release(stream);
            (void)hipGetLastError();
            char buf[200];
            std::snprintf(buf, sizeof buf, "out of device memory allocating a %ux%ux%u lattice (%.2f GB)", g[0], g[1],
                          g[2], (double)(n_cells * 6 * r * 3) / 1e9);
            return set_error(BS_ERR_OOM, buf);
        }
        bytes += n_cells * 6 * r;
    }
    return BS_OK;
}

void BoxState::release(hipStream_t stream) {
    for (int s = 0; s < 3; ++s) {
        if (field[s]) (void)(async_alloc ? hipFreeAsync(field[s], stream) : hipFree(field[s]));
        if (fluid[s]) (void)(async_alloc ? hipFreeAsync(fluid[s], stream) : hipFree(fluid[s]));
        field[s] = fluid[s] = nullptr;
    }
    bytes = 0;
}

int BoxState::init(hipStream_t stream) {
    if (precision == BS_PRECISION_F64) {
        BS_HIP_TRY(launch_init<double>((V2<double> *)field[cur], (V4<double> *)fluid[cur], (int64_t)n_cells,
                                       (double)params.e_ref, stream));
    } else {
        BS_HIP_TRY(launch_init<float>((V2<float> *)field[cur], (V4<float> *)fluid[cur], (int64_t)n_cells,
                                      params.e_ref, stream));
    }
    return BS_OK;
}

template <class R> static int stamp_t(BoxState &bx, const std::vector<bs_event> &events, hipStream_t stream) {
    const Params<R> P = widen_params<R>(bx.params);
    size_t i = 0;
    while (i < events.size()) {
        StampBatch<R> b;
        std::memset(&b, 0, sizeof b);
        while (i < events.size() && b.count < kStampBatch) {
            const bs_event &e = events[i++];
            b.b[b.count][0] = (R)e.pos[0];
            b.b[b.count][1] = (R)e.pos[1];
            b.b[b.count][2] = (R)e.pos[2];
            b.b[b.count][3] = (R)e.radius;
            ++b.count;
        }
        BS_HIP_TRY(launch_stamp<R>((V2<R> *)bx.field[bx.cur], P, b, stream));
    }
    return BS_OK;
}

int stamp_bubbles(BoxState &bx, const std::vector<bs_event> &events, hipStream_t stream) {
    if (events.empty()) return BS_OK;
    return bx.precision == BS_PRECISION_F64 ? stamp_t<double>(bx, events, stream) : stamp_t<float>(bx, events, stream);
}

int run_stage(BoxState &bx, int stage, int math, int kernel, TileShape tile, int zc, void *prim, hipStream_t stream) {
    if (bx.precision == BS_PRECISION_F64) {
        const Box<double> b = bx.box<double>();
        if (kernel == BS_KERNEL_TWO_PASS) BS_HIP_TRY(launch_stage_twopass<double>(math, b, stage, (V4<double> *)prim, stream));
        else BS_HIP_TRY(launch_stage_fused<double>(math, b, stage, tile, zc, stream));
    } else {
        const Box<float> b = bx.box<float>();
        if (kernel == BS_KERNEL_TWO_PASS) BS_HIP_TRY(launch_stage_twopass<float>(math, b, stage, (V4<float> *)prim, stream));
        else BS_HIP_TRY(launch_stage_fused<float>(math, b, stage, tile, zc, stream));
    }
    return BS_OK;
}

int default_zc(const uint32_t g[3], TileShape t) {
    // aim for 16384 workgroups (64 per CU on 256 CUs) with chunks of 24 to
    // 32 planes: shorter ones pay the 4-plane prologue too often, longer
    // ones let neighbouring tiles drift apart in z so that their halo
    // re-reads miss L2.  Measured per launch (steps 205-1105): 256^3 and
    // 320^3 run 2.2 % and 1.5 % faster at 24 planes than at 16 (and than
    // at 32), 384^3 1.1 % faster at 26-32 than at 16, 448^3 1 % slower at
    // 56 than at 32, 512^3 0.3 % slower at 24 than at 32, within +-1 % at
    // 64 (its stage 2/3 fetch 0.6 % higher), 3.5 % slower at 128.
    const int64_t tiles = (int64_t)((g[0] + t.tx - 1) / t.tx) * ((g[1] + t.ty - 1) / t.ty);
    int64_t chunks = (16384 + tiles - 1) / tiles;
    int zc = (int)((g[2] + chunks - 1) / chunks);
    if (zc > 32) zc = 32;
    if (zc < 24) zc = 24;
    if (zc > (int)g[2]) zc = (int)g[2];
    // A box with too few tiles to give every CU 8 workgroups at 16 planes
    // is bound by the march's latency, not by the device: shorter chunks
    // trade prologue planes for parallelism.  Each chunk length is costed as
    // (rounds of 1024 resident workgroups) x (planes + 3 for the prologue);
    // measured on the plain path: 64^3 steps 3x faster with 2-plane chunks,
    // 96^3 1.5x with 4, 128^3 3 % with 8, 192^3 best at 16.
    auto wgs = [&](int z) { return tiles * (int64_t)((g[2] + z - 1) / z); };
    if (wgs(16) < 2048) {
        int64_t best = -1;
        for (int z : {16, 8, 4, 2}) {
            if (z > (int)g[2] && z != 16) continue;
            const int zz = z < (int)g[2] ? z : (int)g[2];
            const int64_t cost = ((wgs(zz) + 1023) / 1024) * (zz + 3);
            if (best < 0 || cost < best) {
                best = cost;
                zc = zz;
            }
        }
    }
    if (const char *e = std::getenv("BS_ZC")) {
        const int v = std::atoi(e);
        if (v > 0) zc = v;
    }
    return zc;
}

int tile_from_env(TileShape dflt, int math, int precision, TileShape *out) {
    TileShape t = dflt;
    if (const char *e = std::getenv("BS_TILE")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%dx%d", &a, &b) != 2 || tile_index(TileShape{a, b}) < 0)
            return set_error(BS_ERR_INVALID, std::string("BS_TILE=") + e + " is not an instantiated tile shape");
        t = TileShape{a, b};
    }
    if (!tile_built(t, math, precision))
        return set_error(BS_ERR_INVALID, "tile " + std::to_string(t.tx) + "x" + std::to_string(t.ty) +
                                             " carries two halo cells per thread and is built for exact arithmetic only");
    *out = t;
    return BS_OK;
}

int ReduceScratch::alloc() {
    BS_HIP_TRY(hipMalloc(&partials, sizeof(ReducePartial) * kReduceBlocks));
    BS_HIP_TRY(hipMalloc(&dres, sizeof(ReduceResult)));
    BS_HIP_TRY(hipHostMalloc(&hres, sizeof(ReduceResult), hipHostMallocDefault));
    return BS_OK;
}

void ReduceScratch::release() {
    if (partials) (void)hipFree(partials);
    if (dres) (void)hipFree(dres);
    if (hres) (void)hipHostFree(hres);
    partials = nullptr;
    dres = nullptr;
    hres = nullptr;
}

int launch_vis(BoxState &bx, Half4 *out, hipStream_t stream) {
    if (bx.precision == BS_PRECISION_F64) {
        const Box<double> b = bx.box<double>();
        BS_HIP_TRY(launch_vis_pack<double>(b.field[b.cur], b.fluid[b.cur], b.P, out, stream));
    } else {
        const Box<float> b = bx.box<float>();
        BS_HIP_TRY(launch_vis_pack<float>(b.field[b.cur], b.fluid[b.cur], b.P, out, stream));
    }
    return BS_OK;
}

int launch_diagnostics(BoxState &bx, ReduceScratch &rs, hipStream_t stream) {
    if (bx.precision == BS_PRECISION_F64) {
        const Box<double> b = bx.box<double>();
        BS_HIP_TRY(launch_reduce<double>(b.field[b.cur], b.fluid[b.cur], b.P, rs.partials, rs.dres, stream));
    } else {
        const Box<float> b = bx.box<float>();
        BS_HIP_TRY(launch_reduce<float>(b.field[b.cur], b.fluid[b.cur], b.P, rs.partials, rs.dres, stream));
    }
    BS_HIP_TRY(hipMemcpyAsync(rs.hres, rs.dres, sizeof(ReduceResult), hipMemcpyDeviceToHost, stream));
    return BS_OK;
}

void diagnostics_from_result(const ReduceResult &r, uint64_t n_cells, bs_diagnostics *out) {
    std::memset(out, 0, sizeof(*out));
    const double n = (double)n_cells;
    out->mean_energy = r.sums[0] / n;
    out->broken_fraction = r.sums[1] / n;
    out->mean_kinetic = r.sums[2] / n;
    out->mean_fluid_energy = r.sums[3] / n;
    out->max_velocity = r.vmax;
    // Diagnostics::kinetic_fraction (sim.rs:160-166)
    out->kinetic_fraction = std::fabs(out->mean_energy) < 1e-20 ? 0.0 : out->mean_kinetic / out->mean_energy;
    out->cells = n_cells;
    out->nonfinite = (r.nonfinite != 0.0 || !std::isfinite(out->mean_energy) || !std::isfinite(out->max_velocity)) ? 1u : 0u;
}

int download_state(BoxState &bx, hipStream_t stream, void *field, void *fluid) {
    const size_t r = real_size(bx.precision);
    if (field)
        BS_HIP_TRY(hipMemcpyAsync(field, bx.field[bx.cur], bx.n_cells * 2 * r, hipMemcpyDeviceToHost, stream));
    if (fluid)
        BS_HIP_TRY(hipMemcpyAsync(fluid, bx.fluid[bx.cur], bx.n_cells * 4 * r, hipMemcpyDeviceToHost, stream));
    BS_HIP_TRY(hipStreamSynchronize(stream));
    return BS_OK;
}

int state_checksum(BoxState &bx, hipStream_t stream, uint64_t *out) {
    const size_t r = real_size(bx.precision);
    std::vector<unsigned char> buf;
    try {
        buf.resize(bx.n_cells * 6 * r);
    } catch (const std::bad_alloc &) {
        return set_error(BS_ERR_OOM, "no host memory for the state checksum");
    }
    int rc = download_state(bx, stream, buf.data(), buf.data() + bx.n_cells * 2 * r);
    if (rc) return rc;
    uint64_t h = 1469598103934665603ULL;
    for (unsigned char c : buf) {
        h ^= c;
        h *= 1099511628211ULL;
    }
    *out = h;
    return BS_OK;
}

}  // namespace bs
