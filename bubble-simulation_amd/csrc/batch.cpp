// batch.cpp -- bs_batch: a parameter sweep of independent boxes advanced
// together.  Each box follows the single-box semantics exactly: events are
// stamped before the step whose f32 clock first reaches their time.
//
// The boxes are split into groups of about equal cell count, each advanced
// by its own engine on its own stream: one fused launch per RK stage across
// the group's boxes, steps captured into a hipGraph.  The groups' launches
// run concurrently, so while one group's stage launch drains its last
// workgroups, the other groups' work fills the compute units.  A box's
// arithmetic does not depend on its group (tasks never span boxes, and every
// group uses the one tile plan the autotuner chose for the whole batch).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "multibox.h"

using namespace bs;

// steps captured per graph (bs_batch_advance replays it, and steps the
// remainder eagerly)
constexpr int kBatchGraphSteps = 25;
// stream groups of a batch (BS_BATCH_STREAMS overrides, 1..4: the device
// keeps 4 hardware queues per process)
constexpr int kBatchGroups = 2;

namespace {
struct Group {
    hipStream_t stream = nullptr;
    std::unique_ptr<Engine> eng;
};
}  // namespace

struct bs_batch {
    int precision = BS_PRECISION_F32;
    int math = BS_MATH_FAST;
    std::vector<std::unique_ptr<Entry>> entries;
    std::vector<Group> groups;
    std::vector<int> group_of;  // box -> group
    ReduceScratch rs;
    bool progress_valid = true;  // host mirrors match the device
};

static int batch_sync_progress(bs_batch *b) {
    if (b->progress_valid) return BS_OK;
    for (Group &g : b->groups) {
        int rc = g.eng->download_progress();
        if (rc) return rc;
    }
    b->progress_valid = true;
    return BS_OK;
}

static hipStream_t box_stream(const bs_batch *b, uint32_t box) { return b->groups[b->group_of[box]].stream; }

static int group_count(uint32_t n) {
    int k = kBatchGroups;
    if (const char *e = std::getenv("BS_BATCH_STREAMS")) k = std::atoi(e);
    k = std::max(1, std::min(k, 4));
    return (int)std::min<uint32_t>((uint32_t)k, n);
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
    const int ng = group_count(n);
    for (int g = 0; g < ng; ++g) {
        Group gr;
        if (hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking) != hipSuccess) {
            bs_batch_destroy(b);
            return set_error(BS_ERR_DEVICE, "hipStreamCreate failed");
        }
        gr.eng.reset(new Engine());
        gr.eng->precision = b->precision;
        gr.eng->math = b->math;
        gr.eng->stream = gr.stream;
        b->groups.push_back(std::move(gr));
    }
    hipStream_t s0 = b->groups[0].stream;
    int rc = b->rs.alloc();
    for (uint32_t i = 0; i < n && !rc; ++i) {
        std::unique_ptr<Entry> e(new Entry());
        rc = e->create(specs[i], s0, false);
        if (rc) {
            const std::string msg = bs_last_error();
            e->release(s0);
            rc = set_error(rc, "box " + std::to_string(i) + ": " + msg);
            break;
        }
        b->entries.push_back(std::move(e));
    }
    if (!rc) {
        // one tile plan for the whole batch, timed over every box
        Engine tune;
        tune.precision = b->precision;
        tune.math = b->math;
        tune.stream = s0;
        for (auto &e : b->entries) tune.entries.push_back(e.get());
        rc = autotune(tune);
        if (!rc) {
            // boxes to groups, largest first onto the group with the fewest cells
            std::vector<uint32_t> order(n);
            for (uint32_t i = 0; i < n; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
                return b->entries[a]->bx.n_cells > b->entries[c]->bx.n_cells;
            });
            std::vector<uint64_t> load(ng, 0);
            b->group_of.assign(n, 0);
            for (uint32_t i : order) {
                const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                b->group_of[i] = g;
                load[g] += b->entries[i]->bx.n_cells;
            }
            for (uint32_t i = 0; i < n; ++i) b->groups[b->group_of[i]].eng->entries.push_back(b->entries[i].get());
            // the entries were initialised on the first group's stream
            if (hipStreamSynchronize(s0) != hipSuccess) rc = set_error(BS_ERR_DEVICE, "hipStreamSynchronize failed");
            for (Group &g : b->groups) {
                if (rc) break;
                g.eng->plan[0].tile = tune.plan[0].tile;
                g.eng->plan[0].zc = tune.plan[0].zc;
                g.eng->plan[1].tile = tune.plan[1].tile;
                g.eng->plan[1].zc = tune.plan[1].zc;
                rc = g.eng->upload();
            }
            for (Group &g : b->groups)
                if (!rc && hipStreamSynchronize(g.stream) != hipSuccess)
                    rc = set_error(BS_ERR_DEVICE, "hipStreamSynchronize failed");
        }
        tune.entries.clear();
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
    for (Group &g : b->groups)
        if (g.stream) (void)hipStreamSynchronize(g.stream);
    hipStream_t s0 = b->groups.empty() ? nullptr : b->groups[0].stream;
    for (Group &g : b->groups) g.eng->entries.clear();
    for (auto &e : b->entries) e->release(s0);
    b->entries.clear();
    b->rs.release();
    for (Group &g : b->groups) {
        g.eng->drop_graph();
        g.eng->reap_graphs(true);
        g.eng->stream = nullptr;  // destroyed below, before the engine
        g.eng.reset();
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    delete b;
}

int bs_batch_advance(bs_batch *b, uint64_t n, int32_t use_graph) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    int rc = batch_sync_progress(b);
    if (rc) return rc;
    // every group's launches are enqueued before any is waited for; run()
    // keeps the host mirrors in step, so the next call needs no read-back
    for (Group &g : b->groups) {
        rc = g.eng->run(n, use_graph != 0, kBatchGraphSteps);
        if (rc) {
            b->progress_valid = false;  // some launches may be queued
            return rc;
        }
    }
    return BS_OK;
}

int bs_batch_synchronize(bs_batch *b) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    for (Group &g : b->groups) BS_HIP_TRY(hipStreamSynchronize(g.stream));
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
    hipStream_t st = box_stream(b, box);
    rc = launch_diagnostics(e->bx, b->rs, st);
    if (rc) return rc;
    BS_HIP_TRY(hipStreamSynchronize(st));
    diagnostics_from_result(*b->rs.hres, e->bx.n_cells, out);
    return BS_OK;
}

int bs_batch_download(bs_batch *b, uint32_t box, void *field, void *fluid) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(b);
    if (rc) return rc;
    return download_state(b->entries[box]->bx, box_stream(b, box), field, fluid);
}

int bs_batch_upload(bs_batch *b, uint32_t box, const void *field, const void *fluid) {
    if (!b) return set_error(BS_ERR_INVALID, "null batch");
    if (box >= b->entries.size()) return set_error(BS_ERR_NOT_FOUND, "no such box");
    int rc = batch_sync_progress(b);  // the slot holding U^n of that box
    if (rc) return rc;
    BoxState &bx = b->entries[box]->bx;
    const size_t r = real_size(bx.precision);
    hipStream_t st = box_stream(b, box);
    if (field) BS_HIP_TRY(hipMemcpyAsync(bx.field[bx.cur], field, bx.n_cells * 2 * r, hipMemcpyHostToDevice, st));
    if (fluid) BS_HIP_TRY(hipMemcpyAsync(bx.fluid[bx.cur], fluid, bx.n_cells * 4 * r, hipMemcpyHostToDevice, st));
    BS_HIP_TRY(hipStreamSynchronize(st));
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
    const Plan &p = b->groups[0].eng->plan[1];
    if (tx) *tx = (uint32_t)p.tile.tx;
    if (ty) *ty = (uint32_t)p.tile.ty;
    return BS_OK;
}

int bs_batch_plan(const bs_batch *b, uint32_t out[6]) {
    if (!b || !out) return set_error(BS_ERR_INVALID, "null argument");
    for (int c = 0; c < 2; ++c) {
        const Plan &p = b->groups[0].eng->plan[c];
        out[3 * c + 0] = (uint32_t)p.tile.tx;
        out[3 * c + 1] = (uint32_t)p.tile.ty;
        out[3 * c + 2] = (uint32_t)p.zc;
    }
    return BS_OK;
}

uint64_t bs_batch_graph_steps(const bs_batch *b) {
    if (!b || b->groups.empty()) return 0;
    uint64_t m = UINT64_MAX;
    for (const Group &g : b->groups) m = std::min(m, g.eng->graph_replayed);
    return m;
}

uint64_t bs_batch_stamps_skipped(const bs_batch *b) {
    if (!b || b->groups.empty()) return 0;
    uint64_t m = UINT64_MAX;
    for (const Group &g : b->groups) m = std::min(m, g.eng->stamps_skipped);
    return m;
}

uint32_t bs_batch_streams(const bs_batch *b) { return b ? (uint32_t)b->groups.size() : 0; }

}  // extern "C"
