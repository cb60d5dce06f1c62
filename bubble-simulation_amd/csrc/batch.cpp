// batch.cpp -- bs_batch: a parameter sweep of independent boxes advanced
// together (one launch per RK stage across all boxes, steps captured into a
// hipGraph).  Each box follows the single-box semantics exactly: events are
// stamped before the step whose f32 clock first reaches their time.
#include <cstring>

#include "multibox.h"

using namespace bs;

// steps captured per graph (bs_batch_advance replays it, and steps the
// remainder eagerly)
constexpr int kBatchGraphSteps = 25;

struct bs_batch {
    int precision = BS_PRECISION_F32;
    int math = BS_MATH_FAST;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Entry>> entries;
    Engine eng;
    ReduceScratch rs;
    bool progress_valid = true;  // host mirrors match the device
};

static int batch_sync_progress(bs_batch *b) {
    if (b->progress_valid) return BS_OK;
    int rc = b->eng.download_progress();
    if (rc) return rc;
    b->progress_valid = true;
    return BS_OK;
}

extern "C" {

int bs_batch_create(const bs_spec *specs, uint32_t n, bs_batch **out) {
    if (!specs || !out || n == 0) return set_error(BS_ERR_INVALID, "need at least one box");
    *out = nullptr;
    for (uint32_t i = 1; i < n; ++i) {
        if (specs[i].precision != specs[0].precision || specs[i].math != specs[0].math)
            return set_error(BS_ERR_INVALID, "all boxes of a batch must share precision and math");
    }
    if (bs_device_count() == 0) return set_error(BS_ERR_DEVICE, "no HIP device available");
    bs_batch *b = new bs_batch();
    b->precision = specs[0].precision;
    b->math = specs[0].math;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return set_error(BS_ERR_DEVICE, "hipStreamCreate failed");
    }
    b->eng.precision = b->precision;
    b->eng.math = b->math;
    b->eng.stream = b->stream;
    int rc = b->rs.alloc();
    for (uint32_t i = 0; i < n && !rc; ++i) {
        std::unique_ptr<Entry> e(new Entry());
        rc = e->create(specs[i], b->stream, false);
        if (rc) {
            const std::string msg = bs_last_error();
            e->release(b->stream);
            rc = set_error(rc, "box " + std::to_string(i) + ": " + msg);
            break;
        }
        b->entries.push_back(std::move(e));
    }
    if (!rc) {
        for (auto &e : b->entries) b->eng.entries.push_back(e.get());
        rc = autotune(b->eng);
    }
    if (rc) {
        const std::string msg = bs_last_error();
        bs_batch_destroy(b);
        return set_error(rc, msg);
    }
    *out = b;
    return BS_OK;
}

void bs_batch_destroy(bs_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->eng.entries.clear();
    for (auto &e : b->entries) e->release(b->stream);
    b->entries.clear();
    b->rs.release();
    b->eng.drop_graph();
    b->eng.reap_graphs(true);
    b->eng.stream = nullptr;  // destroyed below, before the engine
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int bs_batch_advance(bs_batch *b, uint64_t n, int32_t use_graph) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    int rc = batch_sync_progress(b);
    if (rc) return rc;
    b->progress_valid = false;
    return b->eng.run(n, use_graph != 0, kBatchGraphSteps);
}

int bs_batch_synchronize(bs_batch *b) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    BS_HIP_TRY(hipStreamSynchronize(b->stream));
    return BS_OK;
}

uint32_t bs_batch_size(const bs_batch *b) { return b ? (uint32_t)b->entries.size() : 0; }

uint64_t bs_batch_total_cells(const bs_batch *b) {
    uint64_t c = 0;
    if (b)
        for (auto &e : b->entries) c += e->bx.n_cells;
    return c;
}

int bs_batch_diagnostics(bs_batch *b, uint32_t box, bs_diagnostics *out) {
    if (!b || !out) return set_error(BS_ERR_INVALID, "null argument");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(b);
    if (rc) return rc;
    Entry *e = b->entries[box].get();
    rc = launch_diagnostics(e->bx, b->rs, b->stream);
    if (rc) return rc;
    BS_HIP_TRY(hipStreamSynchronize(b->stream));
    diagnostics_from_result(*b->rs.hres, e->bx.n_cells, out);
    return BS_OK;
}

int bs_batch_download(bs_batch *b, uint32_t box, void *field, void *fluid) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(b);
    if (rc) return rc;
    return download_state(b->entries[box]->bx, b->stream, field, fluid);
}

int bs_batch_upload(bs_batch *b, uint32_t box, const void *field, const void *fluid) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(b);  // the slot holding U^n of that box
    if (rc) return rc;
    BoxState &bx = b->entries[box]->bx;
    const size_t r = real_size(bx.precision);
    BS_HIP_TRY(hipStreamSynchronize(b->stream));
    if (field) BS_HIP_TRY(hipMemcpy(bx.field[bx.cur], field, bx.n_cells * 2 * r, hipMemcpyHostToDevice));
    if (fluid) BS_HIP_TRY(hipMemcpy(bx.fluid[bx.cur], fluid, bx.n_cells * 4 * r, hipMemcpyHostToDevice));
    return BS_OK;
}

int bs_batch_steps(const bs_batch *b, uint32_t box, uint64_t *steps_out) {
    if (!b || !steps_out) return set_error(BS_ERR_INVALID, "null argument");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(const_cast<bs_batch *>(b));
    if (rc) return rc;
    *steps_out = b->entries[box]->steps_done;
    return BS_OK;
}

int bs_batch_tile(const bs_batch *b, uint32_t *tx, uint32_t *ty) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    if (tx) *tx = (uint32_t)b->eng.plan[1].tile.tx;
    if (ty) *ty = (uint32_t)b->eng.plan[1].tile.ty;
    return BS_OK;
}

int bs_batch_plan(const bs_batch *b, uint32_t out[6]) {
    if (!b || !out) return set_error(BS_ERR_INVALID, "null argument");
    for (int c = 0; c < 2; ++c) {
        out[3 * c + 0] = (uint32_t)b->eng.plan[c].tile.tx;
        out[3 * c + 1] = (uint32_t)b->eng.plan[c].tile.ty;
        out[3 * c + 2] = (uint32_t)b->eng.plan[c].zc;
    }
    return BS_OK;
}

uint64_t bs_batch_graph_steps(const bs_batch *b) { return b ? b->eng.graph_replayed : 0; }

}  // extern "C"
