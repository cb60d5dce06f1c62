// sim.cpp -- one lattice box: the Simulation API (sim.rs:180-1026) over the
// HIP kernels.  One HIP stream per handle; all state stays on the device.
#include <chrono>
#include <cstring>

#include "device.h"

using namespace bs;

struct bs_sim {
    bs_spec spec{};
    bs_model model{};
    float time = 0.0f;
    uint64_t steps = 0;
    bool nucleated_this_frame = false;
    std::vector<bs_event> schedule;
    size_t next_event = 0;
    hipStream_t stream = nullptr;
    BoxState bx;
    void *prim = nullptr;  // two-pass primitive cache
    Half4 *vis = nullptr;  // visualisation texels, allocated on first use
    ReduceScratch rs;
    hipEvent_t diag_event = nullptr;
    bool diag_pending = false;
    TileShape tile{32, 8};  // stages 2 and 3
    int zc = 16;
    TileShape tile0{32, 8};  // stage 1 (U^n is U_in: no U^n stream)
    int zc0 = 16;
    // per-stage timing (bs_sim_advance_timed)
    bool timing = false;
    std::vector<hipEvent_t> tev;
    double timed_ms = 0.0;
};

namespace {

int sim_alloc(bs_sim *s) {
    // tile shapes first: a BS_TILE the arithmetic has no kernel for is
    // rejected before anything is allocated
    const int math = s->spec.precision == BS_PRECISION_F64 ? BS_MATH_EXACT : s->spec.math;
    // Stages 2 and 3 take 32x8 tiles, except on a box below 2^22 cells whose
    // x edge is an odd multiple of 16: there 32-wide tiles run half-empty at
    // the far edge, and 16x16 measured faster (80^3 +10 %, 112^3 +13 %,
    // 144^3 +16 %; at 176^3 and 208^3 the 16x16 tiles' extra halo fetch
    // costs more than the idle lanes, -2.5 % and -4.5 %)
    const uint64_t cells = (uint64_t)s->spec.grid[0] * s->spec.grid[1] * s->spec.grid[2];
    const bool odd16 = s->spec.grid[0] % 32 != 0 && s->spec.grid[0] % 16 == 0 && cells < (uint64_t(1) << 22);
    int rc = tile_from_env(odd16 ? TileShape{16, 16} : TileShape{32, 8}, math, s->spec.precision, &s->tile);
    if (rc) return rc;
    s->zc = default_zc(s->spec.grid, s->tile);
    // The first stage reads half the bytes of the other two and is bound by
    // instruction issue, where 16x16 tiles (fewer halo and far-edge faces per
    // cell) run 5 % faster at 512^3 (the gamma-law EOS, whose halo cells
    // each take eight Newton steps, 0.5-0.9 % on the whole step); stages 2
    // and 3 are bound by HBM, where 32x8 fetches less.  An explicit BS_TILE
    // applies to every stage.
    const bool fast_f32 = s->spec.math == BS_MATH_FAST && s->spec.precision == BS_PRECISION_F32 &&
                          s->spec.kernel == BS_KERNEL_FUSED;
    rc = tile_from_env(fast_f32 || odd16 ? TileShape{16, 16} : TileShape{32, 8}, math, s->spec.precision, &s->tile0);
    if (rc) return rc;
    s->zc0 = default_zc(s->spec.grid, s->tile0);
    rc = s->bx.alloc(s->spec.precision, s->spec.grid, s->stream, false);
    if (rc) return rc;
    s->bx.set_eos(s->spec);
    if (s->spec.kernel == BS_KERNEL_TWO_PASS) {
        const size_t bytes = s->bx.n_cells * 4 * real_size(s->spec.precision);
        hipError_t e = hipMalloc(&s->prim, bytes);
        if (e != hipSuccess) {
            s->prim = nullptr;
            (void)hipGetLastError();
            return set_error(BS_ERR_OOM, "out of device memory allocating the primitive cache");
        }
    }
    return BS_OK;
}

void sim_free(bs_sim *s) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->bx.release(s->stream);
    if (s->prim) (void)hipFree(s->prim);
    s->prim = nullptr;
    if (s->vis) (void)hipFree(s->vis);
    s->vis = nullptr;
}

// Simulation::reset (sim.rs:790-821)
int sim_reset(bs_sim *s) {
    s->time = 0.0f;
    s->steps = 0;
    s->bx.cur = 0;
    s->next_event = 0;
    sim_params_fill(s->model, s->spec.grid, 0.0f, &s->bx.params);
    return s->bx.init(s->stream);
}

int record_step(bs_sim *s) {
    for (int st = 0; st < 3; ++st) {
        if (s->timing) BS_HIP_TRY(hipEventRecord(s->tev[2 * st], s->stream));
        int rc = run_stage(s->bx, st, s->spec.math, s->spec.kernel, st == 0 ? s->tile0 : s->tile,
                           st == 0 ? s->zc0 : s->zc, s->prim, s->stream);
        if (rc) return rc;
        if (s->timing) BS_HIP_TRY(hipEventRecord(s->tev[2 * st + 1], s->stream));
    }
    if (s->timing) {
        for (int st = 0; st < 3; ++st) {
            BS_HIP_TRY(hipEventSynchronize(s->tev[2 * st + 1]));
            float ms = 0.0f;
            BS_HIP_TRY(hipEventElapsedTime(&ms, s->tev[2 * st], s->tev[2 * st + 1]));
            s->timed_ms += ms;
        }
    }
    s->bx.cur = (s->bx.cur + 1) % 3;
    s->time += model_dt(s->model);
    s->steps += 1;
    return BS_OK;
}

// record_nucleation (sim.rs:843-879): pop events with time <= self.time
int record_nucleation(bs_sim *s) {
    std::vector<bs_event> batch;
    while (s->next_event < s->schedule.size() && s->schedule[s->next_event].time <= s->time && batch.size() < 1024) {
        batch.push_back(s->schedule[s->next_event]);
        ++s->next_event;
    }
    if (batch.empty()) return BS_OK;
    sim_params_fill(s->model, s->spec.grid, s->time, &s->bx.params);
    int rc = stamp_bubbles(s->bx, batch, s->stream);
    if (rc) return rc;
    s->nucleated_this_frame = true;
    return BS_OK;
}

// record_steps (sim.rs:829-841)
int record_steps(bs_sim *s, uint32_t n) {
    if (!s->bx.field[0] || (s->spec.kernel == BS_KERNEL_TWO_PASS && !s->prim))
        return set_error(BS_ERR_INVALID, "the simulation holds no lattice (a resize failed to allocate)");
    s->nucleated_this_frame = false;
    sim_params_fill(s->model, s->spec.grid, s->time, &s->bx.params);
    int rc = record_nucleation(s);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        rc = record_step(s);
        if (rc) return rc;
    }
    return BS_OK;
}

}  // namespace

extern "C" {

int bs_sim_create(const bs_spec *spec, bs_sim **out) {
    if (!spec || !out) return set_error(BS_ERR_INVALID, "null argument");
    *out = nullptr;
    std::string err;
    int rc = spec_validate(*spec, err);
    if (rc) return set_error(rc, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return set_error(BS_ERR_DEVICE, "no HIP device available");
    }
    bs_sim *s = new bs_sim();
    s->spec = *spec;
    s->model = spec->model;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return set_error(BS_ERR_DEVICE, "hipStreamCreate failed");
    }
    rc = spec_schedule(*spec, s->schedule, err);
    if (rc) {
        bs_sim_destroy(s);
        return set_error(rc, err);
    }
    rc = sim_alloc(s);
    if (!rc) rc = s->rs.alloc();
    if (!rc && hipEventCreateWithFlags(&s->diag_event, hipEventDisableTiming) != hipSuccess)
        rc = set_error(BS_ERR_DEVICE, "hipEventCreate failed");
    if (!rc) rc = sim_reset(s);
    if (rc) {
        const std::string msg = bs_last_error();
        bs_sim_destroy(s);
        return set_error(rc, msg);
    }
    *out = s;
    return BS_OK;
}

void bs_sim_destroy(bs_sim *s) {
    if (!s) return;
    sim_free(s);
    s->rs.release();
    if (s->diag_event) (void)hipEventDestroy(s->diag_event);
    for (hipEvent_t e : s->tev) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int bs_sim_reset(bs_sim *s, const bs_event *schedule, uint32_t n) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    if (schedule) s->schedule.assign(schedule, schedule + n);
    return sim_reset(s);
}

int bs_sim_set_model(bs_sim *s, const bs_model *m) {
    if (!s || !m) return set_error(BS_ERR_INVALID, "null argument");
    std::string err;
    int rc = model_validate(*m, err);
    if (rc) return set_error(rc, err);  // ignored, as set_model does (sim.rs:663-670)
    s->model = *m;
    s->spec.model = *m;
    sim_params_fill(s->model, s->spec.grid, s->time, &s->bx.params);
    return BS_OK;
}

int bs_sim_record_steps(bs_sim *s, uint32_t n) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    return record_steps(s, n);
}

int bs_sim_advance(bs_sim *s, uint64_t n) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    for (uint64_t i = 0; i < n; ++i) {
        int rc = record_steps(s, 1);
        if (rc) return rc;
    }
    return BS_OK;
}

int bs_sim_advance_timed(bs_sim *s, uint64_t n, double *stage_kernel_ms, double *wall_ms) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    if (s->tev.empty()) {
        s->tev.resize(6);
        for (auto &e : s->tev) BS_HIP_TRY(hipEventCreate(&e));
    }
    BS_HIP_TRY(hipStreamSynchronize(s->stream));
    s->timing = true;
    s->timed_ms = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = BS_OK;
    for (uint64_t i = 0; i < n && !rc; ++i) rc = record_steps(s, 1);
    s->timing = false;
    if (rc) return rc;
    BS_HIP_TRY(hipStreamSynchronize(s->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if (stage_kernel_ms) *stage_kernel_ms = s->timed_ms;
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return BS_OK;
}

int bs_sim_request_diagnostics(bs_sim *s) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    if (s->diag_pending) return BS_OK;  // the previous readback is not consumed yet (sim.rs:951-954)
    int rc = launch_diagnostics(s->bx, s->rs, s->stream);
    if (rc) return rc;
    BS_HIP_TRY(hipEventRecord(s->diag_event, s->stream));
    s->diag_pending = true;
    return BS_OK;
}

int bs_sim_poll_diagnostics(bs_sim *s, bs_diagnostics *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    if (!s->diag_pending) return set_error(BS_ERR_NOT_READY, "no diagnostics requested");
    hipError_t q = hipEventQuery(s->diag_event);
    if (q == hipErrorNotReady) return BS_ERR_NOT_READY;
    if (q != hipSuccess) return set_error(BS_ERR_DEVICE, std::string("hipEventQuery: ") + hipGetErrorString(q));
    diagnostics_from_result(*s->rs.hres, s->bx.n_cells, out);
    s->diag_pending = false;
    return BS_OK;
}

int bs_sim_diagnostics(bs_sim *s, bs_diagnostics *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    if (s->diag_pending) {
        BS_HIP_TRY(hipEventSynchronize(s->diag_event));
        s->diag_pending = false;
    }
    int rc = bs_sim_request_diagnostics(s);
    if (rc) return rc;
    BS_HIP_TRY(hipEventSynchronize(s->diag_event));
    return bs_sim_poll_diagnostics(s, out);
}

int bs_sim_diagnostics_raw(bs_sim *s, bs_diagnostics_gpu *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    bs_diagnostics d;
    int rc = bs_sim_diagnostics(s, &d);
    if (rc) return rc;
    const ReduceResult &r = *s->rs.hres;
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < 4; ++k) out->sums[k] = (float)r.sums[k];
    const float vm = (float)r.vmax;
    std::memcpy(&out->max_v_bits, &vm, 4);
    return BS_OK;
}

int bs_sim_synchronize(bs_sim *s) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    BS_HIP_TRY(hipStreamSynchronize(s->stream));
    return BS_OK;
}

int bs_sim_resize(bs_sim *s, const uint32_t grid[3]) {
    if (!s || !grid) return set_error(BS_ERR_INVALID, "null argument");
    bs_spec ns = s->spec;
    ns.grid[0] = grid[0]; ns.grid[1] = grid[1]; ns.grid[2] = grid[2];
    std::string err;
    int rc = spec_validate(ns, err);  // validate before releasing anything (sim.rs:743-752)
    if (rc) return set_error(rc, err);
    uint64_t fr = 0, tot = 0;
    if (hipMemGetInfo((size_t *)&fr, (size_t *)&tot) == hipSuccess) {
        const uint64_t need = bytes_per_cell(ns.precision, ns.kernel) * grid[0] * grid[1] * grid[2];
        const uint64_t have = s->bx.bytes + (s->prim ? s->bx.n_cells * 4 * real_size(ns.precision) : 0) + fr;
        if (need > have)
            return set_error(BS_ERR_OOM, "a " + std::to_string(grid[0]) + "x" + std::to_string(grid[1]) + "x" +
                                             std::to_string(grid[2]) + " lattice does not fit in device memory");
    }
    std::vector<bs_event> sched;
    rc = spec_schedule(ns, sched, err);
    if (rc) return set_error(rc, err);
    sim_free(s);  // release first: peak memory is max(old, new)
    s->spec = ns;
    rc = sim_alloc(s);
    if (rc) {
        // Keep the handle usable: fall back to the minimum 8^3 lattice with
        // the same parameters, which is where the reference's resize stands
        // when its final allocation fails (sim.rs:756-760).
        const std::string msg = bs_last_error();
        sim_free(s);
        s->spec.grid[0] = s->spec.grid[1] = s->spec.grid[2] = 8;
        if (spec_schedule(s->spec, s->schedule, err) == BS_OK && sim_alloc(s) == BS_OK) (void)sim_reset(s);
        return set_error(rc, msg);
    }
    s->schedule = sched;
    return sim_reset(s);
}

int bs_sim_nucleate_now(bs_sim *s, const float pos[3]) {
    if (!s || !pos) return set_error(BS_ERR_INVALID, "null argument");
    bs_event e;
    e.time = s->time;
    e.pos[0] = pos[0]; e.pos[1] = pos[1]; e.pos[2] = pos[2];
    e.radius = model_seed_radius(s->model);
    s->schedule.insert(s->schedule.begin() + (long)s->next_event, e);
    return BS_OK;
}

uint64_t bs_sim_cell_count(const bs_sim *s) { return s ? s->bx.n_cells : 0; }
uint32_t bs_sim_remaining_bubbles(const bs_sim *s) {
    return s ? (uint32_t)(s->schedule.size() - s->next_event) : 0;
}
float bs_sim_time(const bs_sim *s) { return s ? s->time : 0.0f; }
uint64_t bs_sim_steps(const bs_sim *s) { return s ? s->steps : 0; }
int32_t bs_sim_nucleated_this_frame(const bs_sim *s) { return s && s->nucleated_this_frame ? 1 : 0; }
uint64_t bs_sim_device_memory_bytes(const bs_sim *s) {
    if (!s) return 0;
    uint64_t b = s->bx.bytes;
    if (s->prim) b += s->bx.n_cells * 4 * real_size(s->spec.precision);
    if (s->vis) b += s->bx.n_cells * sizeof(Half4);
    b += sizeof(ReducePartial) * kReduceBlocks + sizeof(ReduceResult);
    return b;
}
int bs_sim_grid(const bs_sim *s, uint32_t g[3]) {
    if (!s || !g) return set_error(BS_ERR_INVALID, "null argument");
    g[0] = s->spec.grid[0]; g[1] = s->spec.grid[1]; g[2] = s->spec.grid[2];
    return BS_OK;
}
int bs_sim_model(const bs_sim *s, bs_model *m) {
    if (!s || !m) return set_error(BS_ERR_INVALID, "null argument");
    *m = s->model;
    return BS_OK;
}
int bs_sim_schedule(const bs_sim *s, bs_event *out, uint32_t cap, uint32_t *n_out) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    if (n_out) *n_out = (uint32_t)s->schedule.size();
    if (out) {
        const size_t n = std::min<size_t>(cap, s->schedule.size());
        if (n) std::memcpy(out, s->schedule.data(), n * sizeof(bs_event));
    }
    return BS_OK;
}

// Simulation::record_visualisation (sim.rs:920-942): pack slot `cur` into
// the texel buffer (allocated on the first call, 8 B per cell), on the
// simulation's stream.
int bs_sim_record_visualisation(bs_sim *s) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    if (!s->vis) {
        hipError_t e = hipMalloc((void **)&s->vis, s->bx.n_cells * sizeof(Half4));
        if (e != hipSuccess) {
            s->vis = nullptr;
            (void)hipGetLastError();
            return set_error(BS_ERR_OOM, "out of device memory allocating the visualisation texels");
        }
    }
    return launch_vis(s->bx, s->vis, s->stream);
}

int bs_sim_visualisation_buffer(bs_sim *s, void **device_ptr, uint64_t *bytes) {
    if (!s || !device_ptr) return set_error(BS_ERR_INVALID, "null argument");
    if (!s->vis) return set_error(BS_ERR_NOT_READY, "no visualisation recorded yet");
    *device_ptr = s->vis;
    if (bytes) *bytes = s->bx.n_cells * sizeof(Half4);
    return BS_OK;
}

int bs_sim_download_visualisation(bs_sim *s, uint16_t *out, uint64_t n_halfs) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    if (!s->vis) return set_error(BS_ERR_NOT_READY, "no visualisation recorded yet");
    if (n_halfs < 4 * s->bx.n_cells) return set_error(BS_ERR_INVALID, "output holds fewer than 4 halves per cell");
    BS_HIP_TRY(hipMemcpyAsync(out, s->vis, s->bx.n_cells * sizeof(Half4), hipMemcpyDeviceToHost, s->stream));
    BS_HIP_TRY(hipStreamSynchronize(s->stream));
    return BS_OK;
}

// volume.wgsl fs_main over the recorded texels (render.hip)
int bs_sim_render(bs_sim *s, const bs_render_view *view, uint32_t width, uint32_t height, uint8_t *rgba_out) {
    if (!s || !view || !rgba_out) return set_error(BS_ERR_INVALID, "null argument");
    if (!s->vis) return set_error(BS_ERR_NOT_READY, "no visualisation recorded yet");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return set_error(BS_ERR_INVALID, "the frame must be 1x1 to 16384x16384 pixels");
    const size_t bytes = (size_t)width * height * 4;
    void *d_out = nullptr;
    BS_HIP_TRY(hipMalloc(&d_out, bytes));
    hipError_t e = launch_render_volume(s->vis, s->bx.grid, *view, width, height, d_out, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgba_out, d_out, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return set_error(BS_ERR_DEVICE, std::string("volume render: ") + hipGetErrorString(e));
    return BS_OK;
}

// capture::render_to_png (capture.rs:19-117): the visualisation of the
// current state, frame 1 of a fresh renderer, to an RGBA PNG
int bs_sim_screenshot(bs_sim *s, const bs_render_settings *settings, uint32_t width, uint32_t height,
                      const char *path) {
    if (!s || !settings || !path) return set_error(BS_ERR_INVALID, "null argument");
    // the frame limit of bs_sim_render, before the host buffer is sized
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return set_error(BS_ERR_INVALID, "the frame must be 1x1 to 16384x16384 pixels");
    bs_render_view view;
    int rc = bs_render_view_make(s->bx.grid, settings, width, height, 1, &view);
    if (rc) return rc;
    rc = bs_sim_record_visualisation(s);
    if (rc) return rc;
    std::vector<uint8_t> px;
    try {
        px.resize((size_t)width * height * 4);
    } catch (const std::bad_alloc &) {
        return set_error(BS_ERR_OOM, "no host memory for the frame");
    }
    rc = bs_sim_render(s, &view, width, height, px.data());
    if (rc) return rc;
    return bs_png_write(path, px.data(), width, height);
}

int bs_sim_download(bs_sim *s, void *field, void *fluid) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    return download_state(s->bx, s->stream, field, fluid);
}

int bs_sim_upload(bs_sim *s, const void *field, const void *fluid) {
    if (!s) return set_error(BS_ERR_INVALID, "null simulation");
    const size_t r = real_size(s->spec.precision);
    BS_HIP_TRY(hipStreamSynchronize(s->stream));
    if (field) BS_HIP_TRY(hipMemcpy(s->bx.field[s->bx.cur], field, s->bx.n_cells * 2 * r, hipMemcpyHostToDevice));
    if (fluid) BS_HIP_TRY(hipMemcpy(s->bx.fluid[s->bx.cur], fluid, s->bx.n_cells * 4 * r, hipMemcpyHostToDevice));
    return BS_OK;
}

}  // extern "C"
