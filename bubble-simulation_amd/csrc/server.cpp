// server.cpp -- bs_server: continuous batching of sweep jobs.
//
// Two host threads share the device.  The worker owns the stepping streams,
// one per size class of job (below 128^3 cells, and the rest); each
// scheduling tick it
//   1. admits queued jobs in arrival order while their lattice fits the HBM
//      budget (and the active-slot cap), allocating their state
//      stream-ordered on their class's stream and initialising it there;
//   2. advances each class's active jobs with one fused launch per RK stage
//      across all of them (a captured hipGraph of `chunk_steps` steps when
//      the class's set did not change since the previous tick): the large
//      class by one chunk, the small class by the square root of the ratio
//      of the two classes' cells in chunks (1 to 8), so that small jobs
//      progress faster than the large ones they would otherwise wait on;
//   3. checks every active job's state for NaN or infinity, takes the energy
//      baseline of the jobs whose last bubble has been stamped (the first
//      report after the last nucleation, headless.rs:54-78), and hands the
//      jobs that reached their step count, or went non-finite, to the retire
//      thread.
// The retire thread owns a stream of its own.  For each handed-over job it
// waits for the job's last step (an event on its class's stream), runs the final
// diagnostics, reads back the state when the job asked for it (checksum,
// keep_state), returns the device memory and publishes the result, while
// the worker keeps the other jobs stepping.
// A job's arithmetic depends only on its own spec: tasks never span boxes,
// the tile shapes are fixed for the server and a job's z chunks follow from
// its grid, so a job gives the same bits alone and in any mix.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "multibox.h"

using namespace bs;

namespace {

struct Job {
    uint64_t id = 0;
    bs_job job{};
    std::vector<unsigned char> init;  // initial field then fluid bytes, when given
    double submit_s = 0, admit_s = 0;
    std::unique_ptr<Entry> entry;
    int status = BS_OK;
    // energy baseline: reduced on the stepping stream into the job's own
    // scratch when its last event has been stamped, read back at retirement
    ReducePartial *d_partials = nullptr;
    ReduceResult *d_base = nullptr;
    bool base_taken = false;
    uint64_t base_step = 0;
    hipEvent_t last_step = nullptr;  // recorded after the job's last step
    int group = 0;                   // size class: the set and stream it steps in
};

// Size classes.  Each class steps as a set of its own on its own stream, so
// a small job is not held to the step time of a set that holds large ones:
// every tick the large class runs one chunk of steps and the small class
// sqrt(cells of the large class / cells of the small class) chunks (1 to
// kSmallChunksMax), the two streams sharing the device.  (Chunks in the
// plain ratio of the cells, an equal share of the device, cut the median
// latency of the 6 Hz sweep stream further, 0.95 -> 0.62 s, but nearly
// doubled its 99th percentile; the square root takes it to 0.71 s at 2.30 s
// against 2.19.)  The tile shapes are the same in both classes and a small
// job's z chunks follow from its own grid (below), so a job's bits still
// depend only on its own spec.
constexpr int kServerGroups = 2;
constexpr uint64_t kSmallJobCells = uint64_t(1) << 21;  // below 128^3 cells: the small class
constexpr uint64_t kSmallChunksMax = 8;

struct Group {
    Engine eng;
    hipStream_t stream = nullptr;                // stepping (worker)
    std::vector<std::unique_ptr<Job>> active;    // worker only
    bool set_changed = true;
    uint32_t *h_flags = nullptr;                 // pinned copy of the engine's non-finite flags
    size_t h_flags_cap = 0;
    std::vector<char> bad;                       // per active job: found non-finite this tick
};

struct Done {
    bs_job_result res{};
    std::vector<unsigned char> state;  // field then fluid, when keep_state
    uint64_t field_bytes = 0;
};

}  // namespace

struct bs_server {
    bs_server_config cfg{};
    uint64_t budget = 0;
    Group groups[kServerGroups];    // stepping, one set and stream per size class
    hipStream_t rstream = nullptr;  // retirement (retire thread)
    std::thread worker, retirer;
    std::mutex mu;
    std::condition_variable cv;   // results published, queue changed
    std::condition_variable rcv;  // retire queue changed
    bool stop = false;          // worker: stop stepping
    bool retire_stop = false;   // retire thread: stop once its queue is empty
    uint32_t n_active = 0;      // jobs admitted and not yet handed over
    std::deque<std::unique_ptr<Job>> queue;
    std::deque<std::unique_ptr<Job>> retiring;  // worker -> retire thread
    std::map<uint64_t, Done> done;
    uint64_t next_id = 1;
    bs_server_stats stats{};
    std::chrono::steady_clock::time_point t0;
    ReduceScratch rs;  // retire thread
    int worker_error = 0;
    std::string worker_msg;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // jobs being stepped, over both classes (worker only)
    size_t active_count() const {
        size_t n = 0;
        for (const Group &g : groups) n += g.active.size();
        return n;
    }
};

static uint64_t job_bytes(const bs_job &j) {
    return bytes_per_cell(j.spec.precision, BS_KERNEL_FUSED) * (uint64_t)j.spec.grid[0] * j.spec.grid[1] *
           j.spec.grid[2];
}

// ---- retire thread ----------------------------------------------------------

static void release_job(Job &j, hipStream_t st) {
    if (j.entry) j.entry->release(st);
    if (j.d_partials) (void)hipFreeAsync(j.d_partials, st);
    if (j.d_base) (void)hipFreeAsync(j.d_base, st);
    j.d_partials = nullptr;
    j.d_base = nullptr;
    if (j.last_step) (void)hipEventDestroy(j.last_step);
    j.last_step = nullptr;
}

static void retire_one(bs_server *s, Job &j) {
    Done d;
    d.res.job_id = j.id;
    d.res.status = j.status;
    d.res.submit_s = j.submit_s;
    d.res.admit_s = j.admit_s;
    Entry *e = j.entry.get();
    if (e && j.last_step) (void)hipStreamWaitEvent(s->rstream, j.last_step, 0);
    if (e && (j.status == BS_OK || j.status == BS_ERR_NONFINITE)) {
        d.res.steps = e->steps_done;
        bool ok = launch_diagnostics(e->bx, s->rs, s->rstream) == BS_OK;
        ReduceResult base{};
        if (ok && j.base_taken)
            ok = hipMemcpyAsync(&base, j.d_base, sizeof base, hipMemcpyDeviceToHost, s->rstream) == hipSuccess;
        if (ok && hipStreamSynchronize(s->rstream) == hipSuccess) {
            diagnostics_from_result(*s->rs.hres, e->bx.n_cells, &d.res.final_diag);
            if (j.base_taken) {
                diagnostics_from_result(base, e->bx.n_cells, &d.res.baseline_diag);
                d.res.baseline_step = j.base_step;
            }
        }
        const size_t r = real_size(e->spec.precision);
        std::vector<unsigned char> buf;
        if (j.status == BS_OK && (j.job.keep_state || j.job.checksum)) {
            try {
                buf.resize(e->bx.n_cells * 6 * r);
            } catch (const std::bad_alloc &) {
                j.status = BS_ERR_OOM;  // no host memory for the state: the job fails, the server goes on
                d.res.status = j.status;
            }
        }
        if (!buf.empty() &&
            download_state(e->bx, s->rstream, buf.data(), buf.data() + e->bx.n_cells * 2 * r) == BS_OK) {
            uint64_t h = 1469598103934665603ULL;
            for (unsigned char c : buf) {
                h ^= c;
                h *= 1099511628211ULL;
            }
            d.res.state_checksum = h;
            if (j.job.keep_state) {
                d.field_bytes = e->bx.n_cells * 2 * r;
                d.state.swap(buf);
            }
        }
    }
    release_job(j, s->rstream);
    d.res.finish_s = s->now();
    d.res.latency_s = d.res.finish_s - d.res.submit_s;
    std::lock_guard<std::mutex> lk(s->mu);
    if (j.status == BS_OK) {
        s->stats.completed += 1;
        if (e) s->stats.site_updates += e->bx.n_cells * e->steps_done;
    } else {
        s->stats.failed += 1;
        if (j.status == BS_ERR_NONFINITE) s->stats.nonfinite += 1;
    }
    s->stats.bytes_in_use -= std::min<uint64_t>(s->stats.bytes_in_use, job_bytes(j.job));
    s->done[j.id] = std::move(d);
    s->cv.notify_all();
}

static void retire_loop(bs_server *s) {
    (void)hipSetDevice(0);
    while (true) {
        std::unique_ptr<Job> j;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->rcv.wait(lk, [&] { return !s->retiring.empty() || s->retire_stop; });
            if (s->retiring.empty()) break;  // stopping, nothing left
            j = std::move(s->retiring.front());
            s->retiring.pop_front();
        }
        retire_one(s, *j);
    }
    (void)hipStreamSynchronize(s->rstream);
}

// ---- worker -----------------------------------------------------------------

// hand a job over to the retire thread; it is no longer stepped
static void hand_over(bs_server *s, std::unique_ptr<Job> j, int status, bool was_active) {
    if (j->status == BS_OK) j->status = status;
    const hipStream_t st = s->groups[j->group].stream;
    if (j->entry && hipEventCreateWithFlags(&j->last_step, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(j->last_step, st) != hipSuccess) {
            (void)hipEventDestroy(j->last_step);
            j->last_step = nullptr;
            (void)hipStreamSynchronize(st);
        }
    } else if (j->entry) {
        j->last_step = nullptr;
        (void)hipStreamSynchronize(st);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (was_active) s->n_active -= 1;
    s->retiring.push_back(std::move(j));
    s->rcv.notify_all();
}

static int admit(bs_server *s, Job &j) {
    const hipStream_t st = s->groups[j.group].stream;
    j.entry.reset(new Entry());
    int rc = j.entry->create(j.job.spec, st, true);
    if (!rc) {
        hipError_t e = hipMallocAsync((void **)&j.d_partials, sizeof(ReducePartial) * kReduceBlocks, st);
        if (e == hipSuccess) e = hipMallocAsync((void **)&j.d_base, sizeof(ReduceResult), st);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            rc = set_error(e == hipErrorOutOfMemory ? BS_ERR_OOM : BS_ERR_DEVICE,
                           std::string("job scratch: ") + hipGetErrorString(e));
        }
    }
    if (!rc && !j.init.empty()) {
        // over the initialised slot, before anything reads it; the host
        // bytes stay alive with the job until the copy has run
        BoxState &bx = j.entry->bx;
        const size_t fb = bx.n_cells * 2 * real_size(bx.precision);
        hipError_t e = hipMemcpyAsync(bx.field[bx.cur], j.init.data(), fb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(bx.fluid[bx.cur], j.init.data() + fb, j.init.size() - fb, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) rc = set_error(BS_ERR_DEVICE, std::string("initial state: ") + hipGetErrorString(e));
    }
    j.entry->steps_target = j.job.steps;
    return rc;
}

// the energy baseline of a job: once its last event has been stamped
static int take_baseline(bs_server *s, Job &j) {
    Entry *e = j.entry.get();
    if (j.base_taken || e->steps_done <= e->last_event_step()) return BS_OK;
    const hipStream_t st = s->groups[j.group].stream;
    if (e->bx.precision == BS_PRECISION_F64) {
        const Box<double> b = e->bx.box<double>();
        BS_HIP_TRY(launch_reduce<double>(b.field[b.cur], b.fluid[b.cur], b.P, j.d_partials, j.d_base, st));
    } else {
        const Box<float> b = e->bx.box<float>();
        BS_HIP_TRY(launch_reduce<float>(b.field[b.cur], b.fluid[b.cur], b.P, j.d_partials, j.d_base, st));
    }
    j.base_taken = true;
    j.base_step = e->steps_done;
    return BS_OK;
}

static int ensure_flags(Group &g, size_t n) {
    if (n <= g.h_flags_cap) return BS_OK;
    if (g.h_flags) (void)hipHostFree(g.h_flags);
    g.h_flags = nullptr;
    g.h_flags_cap = 0;
    const size_t cap = std::max<size_t>(n, 64);
    BS_HIP_TRY(hipHostMalloc((void **)&g.h_flags, cap * sizeof(uint32_t), hipHostMallocDefault));
    g.h_flags_cap = cap;
    return BS_OK;
}

// one tick over the active sets: advance, non-finite check, baselines; fills
// each class's `bad` with its jobs found non-finite
static int tick(bs_server *s, uint32_t tick_id) {
    int rc = BS_OK;
    uint64_t cells[kServerGroups] = {0, 0};
    for (int c = 0; c < kServerGroups; ++c)
        for (auto &j : s->groups[c].active) cells[c] += j->entry->bx.n_cells;
    const uint64_t chunk = s->cfg.chunk_steps;
    // every launch of the tick is enqueued on both streams before any wait
    for (int c = 0; c < kServerGroups; ++c) {
        Group &g = s->groups[c];
        if (g.active.empty()) continue;
        if (g.set_changed) {
            g.eng.entries.clear();
            for (auto &j : g.active) g.eng.entries.push_back(j->entry.get());
            rc = g.eng.upload();
            if (rc) return rc;
        }
        const bool reuse = !g.set_changed;
        g.set_changed = false;
        uint64_t chunks = 1;
        if (c == 0 && cells[0] && cells[1]) {
            const double r = std::sqrt((double)cells[1] / (double)cells[0]);
            chunks = std::max<uint64_t>(1, std::min<uint64_t>(kSmallChunksMax, (uint64_t)r));
        }
        rc = g.eng.run(chunk * chunks, reuse, (int)chunk);  // advances the host mirrors too
        if (rc) return rc;
    }
    for (Group &g : s->groups) {
        if (g.active.empty()) continue;
        rc = g.eng.check_finite(tick_id);
        if (rc) return rc;
        const size_t n = g.active.size();
        rc = ensure_flags(g, n);
        if (rc) return rc;
        BS_HIP_TRY(hipMemcpyAsync(g.h_flags, g.eng.d_flags, n * sizeof(uint32_t), hipMemcpyDeviceToHost, g.stream));
        for (auto &j : g.active) {
            rc = take_baseline(s, *j);
            if (rc) return rc;
        }
    }
    for (Group &g : s->groups) {
        g.bad.assign(g.active.size(), 0);
        if (g.active.empty()) continue;
        BS_HIP_TRY(hipStreamSynchronize(g.stream));
        for (size_t i = 0; i < g.active.size(); ++i) g.bad[i] = g.h_flags[i] == tick_id;
    }
    return BS_OK;
}

static void worker_loop(bs_server *s) {
    (void)hipSetDevice(0);
    uint32_t tick_id = 0;
    while (true) {
        // ---- admission ----
        std::vector<std::unique_ptr<Job>> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->active_count() == 0 && s->queue.empty() && !s->stop)
                s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop) break;
            while (!s->queue.empty()) {
                const uint64_t need = job_bytes(s->queue.front()->job);
                const bool slot_ok = s->cfg.max_active == 0 || s->active_count() + admitted.size() < s->cfg.max_active;
                if (!slot_ok || s->stats.bytes_in_use + need > s->budget) break;
                s->stats.bytes_in_use += need;
                admitted.push_back(std::move(s->queue.front()));
                s->queue.pop_front();
            }
            // nothing to step and nothing admissible: wait for retirements
            // to return memory, or for new work
            if (s->active_count() == 0 && admitted.empty()) {
                s->cv.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
        }
        for (auto &j : admitted) {
            const uint64_t cells = (uint64_t)j->job.spec.grid[0] * j->job.spec.grid[1] * j->job.spec.grid[2];
            j->group = cells < kSmallJobCells ? 0 : 1;
            const int rc = admit(s, *j);
            j->admit_s = s->now();
            if (rc) {
                hand_over(s, std::move(j), rc, false);
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stats.admitted += 1;
                s->n_active += 1;
            }
            Group &g = s->groups[j->group];
            if (j->group == 0) {
                // a small job's z chunks follow its own size, as on the plain
                // path (a set of a few small boxes is bound by the latency of
                // the march, not by the device), but no shorter than 8 planes,
                // where the prologue would cost more than the latency it
                // saves once the class holds several jobs.  At the 6 Hz sweep
                // stream: p50 0.70 -> 0.56 s, p99 2.31 -> 2.05 s, the same
                // capacity; chunks of 2-4 planes gave p50 0.62 s and lost 6 %
                // of the capacity.
                for (int p = 0; p < 2; ++p)
                    j->entry->zc[p] =
                        std::min(g.eng.plan[p].zc, std::max(8, default_zc(j->job.spec.grid, g.eng.plan[p].tile)));
            }
            g.active.push_back(std::move(j));
            g.set_changed = true;
        }
        if (s->active_count() == 0) continue;

        // ---- advance the active set ----
        const double tb = s->now();
        const int rc = tick(s, ++tick_id);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.ticks += 1;
            s->stats.busy_s += s->now() - tb;
            if (rc) {
                s->worker_error = rc;
                s->worker_msg = bs_last_error();
            } else {
                s->worker_error = 0;  // a tick went through: the device works again
            }
        }
        // ---- retire ----
        for (Group &g : s->groups) {
            std::vector<std::unique_ptr<Job>> keep;
            for (size_t i = 0; i < g.active.size(); ++i) {
                std::unique_ptr<Job> &j = g.active[i];
                int status = BS_OK;
                bool go = false;
                if (rc) {
                    status = rc;  // a failed launch leaves no job's state trustworthy
                    go = true;
                } else if (i < g.bad.size() && g.bad[i]) {
                    status = BS_ERR_NONFINITE;  // evicted early, as the headless run aborts (headless.rs:63-67)
                    go = true;
                } else if (j->entry->steps_done >= j->entry->steps_target) {
                    go = true;
                }
                if (go) {
                    hand_over(s, std::move(j), status, true);
                    g.set_changed = true;
                } else {
                    keep.push_back(std::move(j));
                }
            }
            g.active.swap(keep);
        }
    }
    // shutdown: hand what is left to the retire thread as cancelled
    for (Group &g : s->groups) {
        for (auto &j : g.active) hand_over(s, std::move(j), BS_ERR_NOT_READY, true);
        g.active.clear();
        (void)hipStreamSynchronize(g.stream);
    }
}

extern "C" {

int bs_server_create(const bs_server_config *cfg, bs_server **out) {
    if (!cfg || !out) return set_error(BS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->precision != BS_PRECISION_F32 && cfg->precision != BS_PRECISION_F64)
        return set_error(BS_ERR_INVALID, "invalid precision");
    if (bs_device_count() == 0) return set_error(BS_ERR_DEVICE, "no HIP device available");
    const int math = cfg->precision == BS_PRECISION_F64 ? BS_MATH_EXACT : cfg->math;
    TileShape t0, t1;
    int rc = tile_from_env(TileShape{16, 16}, math, cfg->precision, &t0);
    if (!rc) rc = tile_from_env(TileShape{32, 8}, math, cfg->precision, &t1);
    if (rc) return rc;
    bs_server *s = new bs_server();
    s->cfg = *cfg;
    if (s->cfg.chunk_steps == 0) s->cfg.chunk_steps = 10;
    uint64_t fr = 0, tot = 0;
    if (bs_device_memory(&fr, &tot) != BS_OK) {
        delete s;
        return BS_ERR_DEVICE;
    }
    s->budget = cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (tot > cfg->reserve_bytes ? tot - cfg->reserve_bytes : 0);
    if (s->budget > fr) s->budget = fr > cfg->reserve_bytes ? fr - cfg->reserve_bytes : 0;
    s->stats.budget_bytes = s->budget;
    auto destroy_streams = [&]() {
        for (Group &g : s->groups) {
            g.eng.stream = nullptr;
            if (g.stream) (void)hipStreamDestroy(g.stream);
            g.stream = nullptr;
        }
        if (s->rstream) (void)hipStreamDestroy(s->rstream);
    };
    bool streams_ok = hipStreamCreateWithFlags(&s->rstream, hipStreamNonBlocking) == hipSuccess;
    for (Group &g : s->groups) streams_ok = streams_ok && hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) == hipSuccess;
    if (!streams_ok) {
        destroy_streams();
        delete s;
        return set_error(BS_ERR_DEVICE, "hipStreamCreate failed");
    }
    for (Group &g : s->groups) {
        g.eng.precision = cfg->precision;
        g.eng.math = cfg->math;
        g.eng.stream = g.stream;
        g.eng.use_graphs = (cfg->flags & BS_SERVER_NO_GRAPHS) == 0;
        // the first RK stage, bound by instruction issue, on 16x16 tiles; the
        // other two, which also stream U^n, on 32x8 (as the single box)
        g.eng.plan[0].tile = t0;
        g.eng.plan[1].tile = t1;
        // 24-plane z chunks: the 48-job sweep stream all at once runs 0.5 %
        // faster than with 32, at the same latencies under 6 Hz arrivals
        g.eng.plan[0].zc = g.eng.plan[1].zc = 24;
    }
    rc = s->rs.alloc();
    if (rc) {
        const std::string msg = bs_last_error();
        destroy_streams();
        delete s;
        return set_error(rc, msg);
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(worker_loop, s);
    s->retirer = std::thread(retire_loop, s);
    *out = s;
    return BS_OK;
}

int bs_server_submit(bs_server *s, const bs_job *job, uint64_t *job_id_out) {
    if (!s || !job) return set_error(BS_ERR_INVALID, "null argument");
    bs_spec sp = job->spec;
    sp.precision = s->cfg.precision;
    sp.math = s->cfg.math;
    sp.kernel = BS_KERNEL_FUSED;
    std::string err;
    int rc = spec_validate(sp, err);
    if (rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected += 1;
        return set_error(rc, err);
    }
    if ((job->init_field == nullptr) != (job->init_fluid == nullptr)) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected += 1;
        return set_error(BS_ERR_INVALID, "an initial state needs both the field and the fluid");
    }
    std::unique_ptr<Job> j(new Job());
    j->job = *job;
    j->job.spec = sp;
    if (job->init_field) {
        const size_t cells = (size_t)sp.grid[0] * sp.grid[1] * sp.grid[2];
        const size_t r = real_size(sp.precision);
        try {
            j->init.resize(cells * 6 * r);
        } catch (const std::bad_alloc &) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.rejected += 1;
            return set_error(BS_ERR_OOM, "no host memory for the job's initial state");
        }
        std::memcpy(j->init.data(), job->init_field, cells * 2 * r);
        std::memcpy(j->init.data() + cells * 2 * r, job->init_fluid, cells * 4 * r);
    }
    j->job.init_field = j->job.init_fluid = nullptr;  // the copy above is the job's
    if (job_bytes(j->job) > s->budget) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected += 1;
        return set_error(BS_ERR_OOM, "job lattice exceeds the server's HBM budget");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    j->submit_s = s->now();
    if (job_id_out) *job_id_out = j->id;
    s->queue.push_back(std::move(j));
    s->stats.submitted += 1;
    s->cv.notify_all();
    return BS_OK;
}

int bs_server_poll(bs_server *s, uint64_t id, bs_job_result *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(id);
    if (it == s->done.end()) {
        if (id == 0 || id >= s->next_id) return set_error(BS_ERR_NOT_FOUND, "unknown job id");
        return BS_ERR_NOT_READY;
    }
    *out = it->second.res;
    return BS_OK;
}

int bs_server_wait(bs_server *s, uint64_t id, double timeout_s, bs_job_result *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (id == 0 || id >= s->next_id) return set_error(BS_ERR_NOT_FOUND, "unknown job id");
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (s->done.find(id) == s->done.end()) {
        if (s->cv.wait_until(lk, deadline) == std::cv_status::timeout && s->done.find(id) == s->done.end())
            return BS_ERR_NOT_READY;
    }
    *out = s->done[id].res;
    return BS_OK;
}

int bs_server_fetch_state(bs_server *s, uint64_t id, void *field, void *fluid) {
    if (!s) return set_error(BS_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(id);
    if (it == s->done.end()) return set_error(BS_ERR_NOT_FOUND, "job not completed");
    const Done &d = it->second;
    if (d.state.empty()) return set_error(BS_ERR_NOT_FOUND, "job was not submitted with keep_state");
    if (field) std::memcpy(field, d.state.data(), d.field_bytes);
    if (fluid) std::memcpy(fluid, d.state.data() + d.field_bytes, d.state.size() - d.field_bytes);
    return BS_OK;
}

int bs_server_stats_get(bs_server *s, bs_server_stats *out) {
    if (!s || !out) return set_error(BS_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->active = s->n_active;
    out->queued = (uint32_t)s->queue.size();
    out->elapsed_s = s->now();
    if (s->worker_error) {
        // reported once; a later tick that goes through clears it as well
        const int rc = s->worker_error;
        s->worker_error = 0;
        return set_error(rc, s->worker_msg);
    }
    return BS_OK;
}

void bs_server_destroy(bs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->retire_stop = true;
        s->rcv.notify_all();
    }
    if (s->retirer.joinable()) s->retirer.join();
    s->rs.release();
    for (Group &g : s->groups) {
        g.eng.entries.clear();
        if (g.h_flags) (void)hipHostFree(g.h_flags);
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        g.eng.drop_graph();
        g.eng.reap_graphs(true);
        g.eng.stream = nullptr;  // destroyed below, before the engine
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    if (s->rstream) (void)hipStreamDestroy(s->rstream);
    delete s;
}

}  // extern "C"
