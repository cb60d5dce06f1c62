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
    std::vector<uint64_t> h_steps;  // stamping step of each event (host copy)
    std::vector<float> h_bubbles;   // 4 floats per event (host copy, alive until the upload has run)
    uint64_t *d_steps = nullptr;    // event stamping steps
    float *d_bubbles = nullptr;     // 4 floats per event
    uint64_t steps_target = UINT64_MAX;
    // z chunk of the box's tasks per stage class ([0] RK stage 1, [1] stages
    // 2 and 3); 0 takes the plan's.  A chunk fixed by the box alone keeps
    // its bits the same in any set.
    int zc[2] = {0, 0};
    // host mirrors of the device progress
    uint64_t steps_done = 0;
    int next_event = 0;

    // allocate slots + event tables, init; stream-ordered when `async`, and
    // then nothing here waits for the stream
    int create(const bs_spec &s, hipStream_t stream, bool async);
    void release(hipStream_t stream);
    uint64_t device_bytes() const;
    // the step before which the last event is stamped (0 without events)
    uint64_t last_event_step() const { return h_steps.empty() ? 0 : h_steps.back(); }
    // mirror of n device steps (k_advance_multi): events due at a step are
    // passed, then the step is counted and the slot ring turns
    void host_advance(uint64_t n);
};

// Equation-of-state classes of a set (BS_EOS_BAG, BS_EOS_GAMMA_LAW): each
// stage launch runs the tasks of one class, so a set that holds both takes
// two launches per stage and a bag-only set runs the bag kernels alone.
constexpr int kEosClasses = 2;

// Tile shape and z chunk of one class of stage launches, with its task
// tables (one per equation of state): the first RK stage (reads U_in only)
// and the other two (stream U^n as well) are bound differently and may take
// different tiles.  The gamma-law tasks are cut with eos_tile(kEosGammaLaw,
// tile).
struct Plan {
    TileShape tile{16, 16};
    int zc = 32;
    Task *d_tasks[kEosClasses] = {nullptr, nullptr};
    int ntasks[kEosClasses] = {0, 0};
    size_t cap[kEosClasses] = {0, 0};
};

// The launch machinery over a set of entries.
struct Engine {
    int precision = BS_PRECISION_F32;
    int math = BS_MATH_FAST;
    Plan plan[2];  // [0]: RK stage 1, [1]: RK stages 2 and 3
    hipStream_t stream = nullptr;

    std::vector<Entry *> entries;   // the current set (order = device index)
    void *d_boxes = nullptr;        // Box<R>[n]
    BoxEvents *d_events = nullptr;  // BoxEvents[n]
    uint32_t *d_flags = nullptr;    // per box: tick of the last non-finite check that failed
    size_t cap_boxes = 0;
    // pinned staging of the uploads, two buffers used in turn; an upload
    // waits only for the copies out of the buffer it reuses
    void *stage_buf[2] = {nullptr, nullptr};
    size_t stage_cap[2] = {0, 0};
    hipEvent_t stage_done[2] = {nullptr, nullptr};
    int stage_next = 0;
    // captured chunks of steps, [0] without and [1] with the nucleation
    // stamp ahead of each step
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    uint64_t graph_replayed = 0;  // steps run through graph replay
    uint64_t stamps_skipped = 0;  // steps run without the stamp launch
    bool use_graphs = true;
    // dropped graphs, destroyed once the stream has passed their last launch
    struct Dead {
        hipGraph_t g[2];
        hipGraphExec_t x[2];
        hipEvent_t after;
    };
    std::vector<Dead> graveyard;

    ~Engine();
    // upload the current host mirrors of `entries` (progress, cur) and the
    // task tables (stream-ordered; the host does not wait for the stream);
    // invalidates a captured graph
    int upload();
    // read device progress (steps_done, next, cur) back into the entries
    int download_progress();
    // whether any entry passes an event in its next k steps (from the host
    // mirrors, which follow the device exactly)
    bool event_due(uint64_t k) const;
    // one step of every entry (eager), with or without the stamp launch
    int step_eager(bool stamp);
    // n steps, through a captured graph of `graph_chunk` steps when
    // use_graph; the stamp launch is left out of every step (or chunk) in
    // which no entry has an event due.  Keeps the host mirrors of every
    // entry in step with the launches enqueued.
    int run(uint64_t n, bool use_graph, int graph_chunk);
    // mean time of one launch of each stage class over the current set, in
    // ms (stage 1 alone; stages 2 and 3 as a pair, per launch)
    int time_stages(float ms[2]);
    // flag every box of the set whose slot cur holds a non-finite value:
    // d_flags[box] = tick (never cleared: a flag is current when it equals
    // the tick of the check)
    int check_finite(uint32_t tick);
    void drop_graph();
    void reap_graphs(bool wait);
};

// pick the tile shape and z chunk of each stage class, minimising the stage
// times of the set; the set's entries must be freshly initialised (slot cur
// is stamped with every event of its box for the timing and initialised
// again afterwards)
int autotune(Engine &eng);

}  // namespace bs
