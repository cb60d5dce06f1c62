// server.cpp -- bs_server: continuous batching of sweep jobs.
//
// A worker thread owns the device.  Each scheduling tick it
//   1. admits queued jobs in arrival order while their lattice fits the HBM
//      budget (and the active-slot cap), allocating their state
//      stream-ordered and initialising it on the device;
//   2. advances every active job by `chunk_steps` steps with one fused launch
//      per RK stage across all of them (a captured hipGraph when the active
//      set did not change since the previous tick);
//   3. retires jobs that reached their step count: final diagnostics, state
//      checksum (and optionally the state itself) are kept as the result and
//      the device memory is returned to the pool.
// A job's arithmetic depends only on its own spec: tasks never span boxes and
// the tile shape and z chunk are fixed for the server, so a job gives the
// same bits alone and in any mix.
#include <chrono>
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
    double submit_s = 0, admit_s = 0;
    std::unique_ptr<Entry> entry;
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
    hipStream_t stream = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    std::deque<std::unique_ptr<Job>> queue;
    std::vector<std::unique_ptr<Job>> active;
    std::map<uint64_t, Done> done;
    uint64_t next_id = 1;
    bs_server_stats stats{};
    std::chrono::steady_clock::time_point t0;
    Engine eng;
    ReduceScratch rs;
    bool set_changed = true;
    int worker_error = 0;
    std::string worker_msg;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static uint64_t job_bytes(const bs_job &j) {
    return bytes_per_cell(j.spec.precision, BS_KERNEL_FUSED) * (uint64_t)j.spec.grid[0] * j.spec.grid[1] *
           j.spec.grid[2];
}

static void finish_job(bs_server *s, Job &j, int status) {
    Done d;
    d.res.job_id = j.id;
    d.res.status = status;
    d.res.submit_s = j.submit_s;
    d.res.admit_s = j.admit_s;
    if (j.entry && status == BS_OK) {
        Entry *e = j.entry.get();
        d.res.steps = e->steps_done;
        if (launch_diagnostics(e->bx, s->rs, s->stream) == BS_OK && hipStreamSynchronize(s->stream) == hipSuccess)
            diagnostics_from_result(*s->rs.hres, e->bx.n_cells, &d.res.final_diag);
        const size_t r = real_size(e->spec.precision);
        std::vector<unsigned char> buf;
        if (j.job.keep_state || j.job.checksum) buf.resize(e->bx.n_cells * 6 * r);
        if (!buf.empty() &&
            download_state(e->bx, s->stream, buf.data(), buf.data() + e->bx.n_cells * 2 * r) == BS_OK) {
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
    if (j.entry) j.entry->release(s->stream);
    d.res.finish_s = s->now();
    d.res.latency_s = d.res.finish_s - d.res.submit_s;
    std::lock_guard<std::mutex> lk(s->mu);
    if (status == BS_OK) {
        s->stats.completed += 1;
        if (j.entry) s->stats.site_updates += j.entry->bx.n_cells * j.entry->steps_done;
    }
    s->stats.bytes_in_use -= std::min<uint64_t>(s->stats.bytes_in_use, job_bytes(j.job));
    s->done[j.id] = std::move(d);
    s->cv.notify_all();
}

static void worker_loop(bs_server *s) {
    (void)hipSetDevice(0);
    while (true) {
        // ---- admission ----
        std::vector<std::unique_ptr<Job>> admitted;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->active.empty() && s->queue.empty() && !s->stop) s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop) break;
            while (!s->queue.empty()) {
                const uint64_t need = job_bytes(s->queue.front()->job);
                const bool slot_ok = s->cfg.max_active == 0 || s->active.size() + admitted.size() < s->cfg.max_active;
                if (!slot_ok || s->stats.bytes_in_use + need > s->budget) break;
                s->stats.bytes_in_use += need;
                admitted.push_back(std::move(s->queue.front()));
                s->queue.pop_front();
            }
        }
        for (auto &j : admitted) {
            j->entry.reset(new Entry());
            int rc = j->entry->create(j->job.spec, s->stream, true);
            j->admit_s = s->now();
            if (rc) {
                finish_job(s, *j, rc);
                continue;
            }
            j->entry->steps_target = j->job.steps;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stats.admitted += 1;
                s->active.push_back(std::move(j));
            }
            s->set_changed = true;
        }
        if (s->active.empty()) continue;

        // ---- advance the active set ----
        const double tb = s->now();
        int rc = BS_OK;
        if (s->set_changed) {
            s->eng.entries.clear();
            for (auto &j : s->active) s->eng.entries.push_back(j->entry.get());
            rc = s->eng.upload();
        }
        const bool reuse = !s->set_changed;
        s->set_changed = false;
        if (!rc) rc = s->eng.run(s->cfg.chunk_steps, reuse, (int)s->cfg.chunk_steps);
        if (!rc) rc = s->eng.download_progress();
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->worker_error = rc;
            s->worker_msg = bs_last_error();
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.ticks += 1;
            s->stats.busy_s += s->now() - tb;
        }
        // ---- retire ----
        for (size_t i = 0; i < s->active.size();) {
            Job &j = *s->active[i];
            if (rc || j.entry->steps_done >= j.entry->steps_target) {
                finish_job(s, j, rc ? rc : BS_OK);
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->active.erase(s->active.begin() + (long)i);
                }
                s->set_changed = true;
            } else {
                ++i;
            }
        }
    }
    // shutdown: drop whatever is left
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &j : s->active)
        if (j->entry) j->entry->release(s->stream);
    s->active.clear();
    (void)hipStreamSynchronize(s->stream);
}

extern "C" {

int bs_server_create(const bs_server_config *cfg, bs_server **out) {
    if (!cfg || !out) return set_error(BS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->precision != BS_PRECISION_F32 && cfg->precision != BS_PRECISION_F64)
        return set_error(BS_ERR_INVALID, "invalid precision");
    if (bs_device_count() == 0) return set_error(BS_ERR_DEVICE, "no HIP device available");
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
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return set_error(BS_ERR_DEVICE, "hipStreamCreate failed");
    }
    s->eng.precision = cfg->precision;
    s->eng.math = cfg->math;
    s->eng.stream = s->stream;
    s->eng.tile = tile_from_env(TileShape{16, 16});
    s->eng.zc = 32;
    int rc = s->rs.alloc();
    if (rc) {
        const std::string msg = bs_last_error();
        (void)hipStreamDestroy(s->stream);
        delete s;
        return set_error(rc, msg);
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(worker_loop, s);
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
    if (!rc && sp.eos != BS_EOS_BAG) {
        err = "the server runs the bag equation of state only";
        rc = BS_ERR_INVALID;
    }
    if (rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected += 1;
        return set_error(rc, err);
    }
    std::unique_ptr<Job> j(new Job());
    j->job = *job;
    j->job.spec = sp;
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
    out->active = (uint32_t)s->active.size();
    out->queued = (uint32_t)s->queue.size();
    out->elapsed_s = s->now();
    if (s->worker_error) return set_error(s->worker_error, s->worker_msg);
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
    s->eng.entries.clear();
    s->rs.release();
    s->eng.drop_graph();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
