// multibox.cpp -- Entry / Engine of the batched and served paths.
#include "multibox.h"

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

namespace bs {

int Entry::create(const bs_spec &s, hipStream_t stream, bool async) {
    spec = s;
    std::string err;
    int rc = spec_validate(s, err);
    if (rc) return set_error(rc, err);
    if (s.eos != BS_EOS_BAG) return set_error(BS_ERR_INVALID, "the batched engine runs the bag equation of state only");
    rc = spec_schedule(s, schedule, err);
    if (rc) return set_error(rc, err);
    rc = bx.alloc(s.precision, s.grid, stream, async);
    if (rc) return rc;
    sim_params_fill(s.model, s.grid, 0.0f, &bx.params);
    bx.cur = 0;
    rc = bx.init(stream);
    if (rc) return rc;
    const size_t ne = schedule.size();
    std::vector<uint64_t> steps(ne);
    std::vector<float> bub(4 * ne);
    const float dt = model_dt(s.model);
    for (size_t k = 0; k < ne; ++k) {
        steps[k] = event_step(schedule[k].time, dt);
        bub[4 * k + 0] = schedule[k].pos[0];
        bub[4 * k + 1] = schedule[k].pos[1];
        bub[4 * k + 2] = schedule[k].pos[2];
        bub[4 * k + 3] = schedule[k].radius;
    }
    if (ne) {
        if (async) {
            BS_HIP_TRY(hipMallocAsync((void **)&d_steps, ne * sizeof(uint64_t), stream));
            BS_HIP_TRY(hipMallocAsync((void **)&d_bubbles, ne * 4 * sizeof(float), stream));
        } else {
            BS_HIP_TRY(hipMalloc((void **)&d_steps, ne * sizeof(uint64_t)));
            BS_HIP_TRY(hipMalloc((void **)&d_bubbles, ne * 4 * sizeof(float)));
        }
        BS_HIP_TRY(hipMemcpyAsync(d_steps, steps.data(), ne * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
        BS_HIP_TRY(hipMemcpyAsync(d_bubbles, bub.data(), ne * 4 * sizeof(float), hipMemcpyHostToDevice, stream));
        // the host vectors must outlive the copies
        BS_HIP_TRY(hipStreamSynchronize(stream));
    }
    steps_done = 0;
    next_event = 0;
    return BS_OK;
}

void Entry::release(hipStream_t stream) {
    const bool async = bx.async_alloc;
    bx.release(stream);
    if (d_steps) (void)(async ? hipFreeAsync(d_steps, stream) : hipFree(d_steps));
    if (d_bubbles) (void)(async ? hipFreeAsync(d_bubbles, stream) : hipFree(d_bubbles));
    d_steps = nullptr;
    d_bubbles = nullptr;
}

uint64_t Entry::device_bytes() const {
    return bytes_per_cell(spec.precision, BS_KERNEL_FUSED) * (uint64_t)spec.grid[0] * spec.grid[1] * spec.grid[2] +
           schedule.size() * 20;
}

Engine::~Engine() {
    drop_graph();
    if (d_boxes) (void)hipFree(d_boxes);
    if (d_events) (void)hipFree(d_events);
    if (d_tasks) (void)hipFree(d_tasks);
}

void Engine::drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_steps = 0;
}

template <class R> static void fill_boxes(const std::vector<Entry *> &es, std::vector<unsigned char> &buf) {
    buf.resize(es.size() * sizeof(Box<R>));
    Box<R> *b = (Box<R> *)buf.data();
    for (size_t i = 0; i < es.size(); ++i) b[i] = es[i]->bx.box<R>();
}

int Engine::upload() {
    drop_graph();
    const size_t n = entries.size();
    const size_t box_bytes = precision == BS_PRECISION_F64 ? sizeof(Box<double>) : sizeof(Box<float>);
    // tasks
    std::vector<Task> tasks;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t *g = entries[i]->spec.grid;
        for (uint32_t zb = 0; zb < g[2]; zb += (uint32_t)zc)
            for (uint32_t y0 = 0; y0 < g[1]; y0 += (uint32_t)tile.ty)
                for (uint32_t x0 = 0; x0 < g[0]; x0 += (uint32_t)tile.tx) {
                    Task t;
                    std::memset(&t, 0, sizeof t);
                    t.box = (int)i;
                    t.x0 = (int)x0;
                    t.y0 = (int)y0;
                    t.zb = (int)zb;
                    t.ze = (int)std::min<uint32_t>(zb + (uint32_t)zc, g[2]);
                    tasks.push_back(t);
                }
    }
    // pad to a multiple of 8 with empty tasks, so that the kernel can hand
    // each XCD a contiguous run of tiles (blocks go round-robin over the 8
    // XCDs; neighbouring tiles then share that XCD's L2)
    while (!tasks.empty() && tasks.size() % 8 != 0) {
        Task t;
        std::memset(&t, 0, sizeof t);
        tasks.push_back(t);
    }
    ntasks = (int)tasks.size();
    if (n > cap_boxes) {
        if (d_boxes) (void)hipFree(d_boxes);
        if (d_events) (void)hipFree(d_events);
        d_boxes = nullptr;
        d_events = nullptr;
        const size_t cap = std::max<size_t>(n, 16);
        BS_HIP_TRY(hipMalloc(&d_boxes, cap * sizeof(Box<double>)));
        BS_HIP_TRY(hipMalloc((void **)&d_events, cap * sizeof(BoxEvents)));
        cap_boxes = cap;
    }
    if ((size_t)ntasks > cap_tasks) {
        if (d_tasks) (void)hipFree(d_tasks);
        d_tasks = nullptr;
        const size_t cap = std::max<size_t>((size_t)ntasks, 256);
        BS_HIP_TRY(hipMalloc((void **)&d_tasks, cap * sizeof(Task)));
        cap_tasks = cap;
    }
    std::vector<unsigned char> boxes;
    if (precision == BS_PRECISION_F64) fill_boxes<double>(entries, boxes);
    else fill_boxes<float>(entries, boxes);
    std::vector<BoxEvents> ev(n);
    for (size_t i = 0; i < n; ++i) {
        Entry *e = entries[i];
        std::memset(&ev[i], 0, sizeof(BoxEvents));
        ev[i].step = e->d_steps;
        ev[i].bubble = e->d_bubbles;
        ev[i].count = (int)e->schedule.size();
        ev[i].next = e->next_event;
        ev[i].steps_done = e->steps_done;
        ev[i].steps_target = e->steps_target;
        ev[i].active = 1;
    }
    if (n) {
        BS_HIP_TRY(hipMemcpyAsync(d_boxes, boxes.data(), n * box_bytes, hipMemcpyHostToDevice, stream));
        BS_HIP_TRY(hipMemcpyAsync(d_events, ev.data(), n * sizeof(BoxEvents), hipMemcpyHostToDevice, stream));
    }
    if (ntasks)
        BS_HIP_TRY(hipMemcpyAsync(d_tasks, tasks.data(), (size_t)ntasks * sizeof(Task), hipMemcpyHostToDevice, stream));
    BS_HIP_TRY(hipStreamSynchronize(stream));
    return BS_OK;
}

int Engine::download_progress() {
    const size_t n = entries.size();
    if (!n) return BS_OK;
    const size_t box_bytes = precision == BS_PRECISION_F64 ? sizeof(Box<double>) : sizeof(Box<float>);
    std::vector<unsigned char> boxes(n * box_bytes);
    std::vector<BoxEvents> ev(n);
    BS_HIP_TRY(hipMemcpyAsync(boxes.data(), d_boxes, n * box_bytes, hipMemcpyDeviceToHost, stream));
    BS_HIP_TRY(hipMemcpyAsync(ev.data(), d_events, n * sizeof(BoxEvents), hipMemcpyDeviceToHost, stream));
    BS_HIP_TRY(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; ++i) {
        int cur;
        if (precision == BS_PRECISION_F64) cur = ((Box<double> *)boxes.data())[i].cur;
        else cur = ((Box<float> *)boxes.data())[i].cur;
        entries[i]->bx.cur = cur;
        entries[i]->steps_done = ev[i].steps_done;
        entries[i]->next_event = ev[i].next;
    }
    return BS_OK;
}

int Engine::step_eager() {
    const int n = (int)entries.size();
    if (!n) return BS_OK;
    if (precision == BS_PRECISION_F64) {
        Box<double> *db = (Box<double> *)d_boxes;
        BS_HIP_TRY(launch_stamp_multi<double>(db, d_events, n, stream));
        for (int s = 0; s < 3; ++s)
            BS_HIP_TRY(launch_stage_fused_multi<double>(math, db, d_events, d_tasks, ntasks, s, tile, stream));
        BS_HIP_TRY(launch_advance_multi<double>(db, d_events, n, stream));
    } else {
        Box<float> *db = (Box<float> *)d_boxes;
        BS_HIP_TRY(launch_stamp_multi<float>(db, d_events, n, stream));
        for (int s = 0; s < 3; ++s)
            BS_HIP_TRY(launch_stage_fused_multi<float>(math, db, d_events, d_tasks, ntasks, s, tile, stream));
        BS_HIP_TRY(launch_advance_multi<float>(db, d_events, n, stream));
    }
    return BS_OK;
}

int Engine::run(uint64_t n, bool use_graph, int graph_chunk) {
    if (entries.empty() || n == 0) return BS_OK;
    if (!use_graph || graph_chunk <= 1 || n < (uint64_t)graph_chunk) {
        for (uint64_t i = 0; i < n; ++i) {
            int rc = step_eager();
            if (rc) return rc;
        }
        return BS_OK;
    }
    if (!gexec || graph_steps != graph_chunk) {
        drop_graph();
        BS_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = BS_OK;
        for (int i = 0; i < graph_chunk && !rc; ++i) rc = step_eager();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(BS_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        graph = g;
        BS_HIP_TRY(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_steps = graph_chunk;
    }
    uint64_t done = 0;
    while (n - done >= (uint64_t)graph_steps) {
        BS_HIP_TRY(hipGraphLaunch(gexec, stream));
        done += (uint64_t)graph_steps;
    }
    for (; done < n; ++done) {
        int rc = step_eager();
        if (rc) return rc;
    }
    return BS_OK;
}

int Engine::time_stage0(float *ms) {
    hipEvent_t a, b;
    BS_HIP_TRY(hipEventCreate(&a));
    BS_HIP_TRY(hipEventCreate(&b));
    int rc = BS_OK;
    hipError_t e = hipSuccess;
    // stage 0 reads slot cur and writes the next slot: repeating it leaves
    // U^n untouched
    auto launch = [&]() -> hipError_t {
        if (precision == BS_PRECISION_F64)
            return launch_stage_fused_multi<double>(math, (Box<double> *)d_boxes, d_events, d_tasks, ntasks, 0, tile, stream);
        return launch_stage_fused_multi<float>(math, (Box<float> *)d_boxes, d_events, d_tasks, ntasks, 0, tile, stream);
    };
    e = launch();  // warm
    if (e == hipSuccess) e = hipEventRecord(a, stream);
    for (int r = 0; r < 3 && e == hipSuccess; ++r) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (e != hipSuccess) rc = set_error(BS_ERR_DEVICE, std::string("autotune timing: ") + hipGetErrorString(e));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / 3.0f;
    return rc;
}

int autotune(Engine &eng) {
    if (const char *e = std::getenv("BS_TILE")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%dx%d", &a, &b) == 2 && tile_index(TileShape{a, b}) >= 0) {
            eng.tile = TileShape{a, b};
            if (const char *z = std::getenv("BS_ZC")) eng.zc = std::max(1, std::atoi(z));
            return eng.upload();
        }
    }
    static std::mutex mu;
    static std::map<std::string, std::pair<int, int>> cache;
    std::string key = std::to_string(eng.precision) + ":" + std::to_string(eng.math);
    for (Entry *e : eng.entries)
        key += ":" + std::to_string(e->spec.grid[0]) + "x" + std::to_string(e->spec.grid[1]) + "x" +
               std::to_string(e->spec.grid[2]);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            eng.tile = kTiles[it->second.first];
            eng.zc = it->second.second;
            return eng.upload();
        }
    }
    float best = 1e30f;
    int best_t = 0, best_zc = 32;
    const int zcs[3] = {16, 32, 64};
    for (int ti = 0; ti < kNumTiles; ++ti) {
        if (eng.precision == BS_PRECISION_F64 && kTiles[ti].tx * kTiles[ti].ty > 256) continue;
        for (int zc : zcs) {
            eng.tile = kTiles[ti];
            eng.zc = zc;
            int rc = eng.upload();
            if (rc) return rc;
            float ms = 0.0f;
            rc = eng.time_stage0(&ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                best_t = ti;
                best_zc = zc;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = {best_t, best_zc};
    }
    eng.tile = kTiles[best_t];
    eng.zc = best_zc;
    return eng.upload();
}

}  // namespace bs
