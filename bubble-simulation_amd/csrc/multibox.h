// multibox.h -- many independent boxes advanced together: one fused launch
// per RK stage over a task table that spans every active box, with per-box
// nucleation and progress kept on the device so that a captured hipGraph of
// steps replays without the host.
#pragma once

#include <memory>

#include "device.h"

namespace bs {

// One box of a batch or a served job.
struct Entry {
    bs_spec spec{};
    BoxState bx;
    std::vector<bs_event> schedule;
    uint64_t *d_steps = nullptr;   // event stamping steps
    float *d_bubbles = nullptr;    // 4 floats per event
    uint64_t steps_target = UINT64_MAX;
    // host mirrors of the device progress
    uint64_t steps_done = 0;
    int next_event = 0;

    // allocate slots + event tables, init (async on `stream`)
    int create(const bs_spec &s, hipStream_t stream, bool async);
    void release(hipStream_t stream);
    uint64_t device_bytes() const;
};

// The launch machinery over a set of entries.
struct Engine {
    int precision = BS_PRECISION_F32;
    int math = BS_MATH_FAST;
    TileShape tile{32, 16};
    int zc = 32;
    hipStream_t stream = nullptr;

    std::vector<Entry *> entries;   // the current set (order = device index)
    void *d_boxes = nullptr;        // Box<R>[n]
    BoxEvents *d_events = nullptr;  // BoxEvents[n]
    Task *d_tasks = nullptr;
    int ntasks = 0;
    size_t cap_boxes = 0, cap_tasks = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;

    ~Engine();
    // upload the current host mirrors of `entries` (progress, cur) and the
    // task table; invalidates a captured graph
    int upload();
    // read device progress (steps_done, next, cur) back into the entries
    int download_progress();
    // one step of every entry (eager)
    int step_eager();
    // n steps, through a captured graph of `graph_chunk` steps when use_graph
    int run(uint64_t n, bool use_graph, int graph_chunk);
    // time one stage-0 launch of the current set (idempotent) in ms
    int time_stage0(float *ms);
    void drop_graph();
};

// pick the tile shape and z chunk minimising the stage time of the set
int autotune(Engine &eng);

}  // namespace bs
