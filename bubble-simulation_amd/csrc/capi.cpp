// capi.cpp -- error surface, version, ABI struct sizes and device queries.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "bubblesim.h"
#include "internal.h"

static_assert(sizeof(bs_sim_params) == 96, "SimParams is 96 bytes (sim.rs:1072)");
static_assert(sizeof(bs_stage_params) == 16, "StageParams is 16 bytes (sim.rs:1073)");
static_assert(sizeof(bs_bubble) == 16, "Bubble is 16 bytes (sim.rs:1074)");
static_assert(sizeof(bs_nucleation_batch) == 16, "NucleationBatch is 16 bytes");
static_assert(sizeof(bs_diagnostics_gpu) == 32, "Diagnostics block is 32 bytes (sim.rs:1075)");

namespace bs {

static thread_local std::string g_last_error;

int set_error(int rc, const std::string &msg) {
    g_last_error = msg;
    return rc;
}
void clear_error() { g_last_error.clear(); }

}  // namespace bs

extern "C" {

const char *bs_last_error(void) { return bs::g_last_error.c_str(); }

const char *bs_version(void) { return "bubblesim-mi355x 0.1.0 (gfx950)"; }

size_t bs_struct_size(int which) {
    switch (which) {
    case 0: return sizeof(bs_sim_params);
    case 1: return sizeof(bs_stage_params);
    case 2: return sizeof(bs_bubble);
    case 3: return sizeof(bs_nucleation_batch);
    case 4: return sizeof(bs_diagnostics_gpu);
    case 5: return sizeof(bs_diagnostics);
    case 6: return sizeof(bs_spec);
    case 7: return sizeof(bs_model);
    case 8: return sizeof(bs_event);
    case 9: return sizeof(bs_derived);
    case 10: return sizeof(bs_server_config);
    case 11: return sizeof(bs_job);
    case 12: return sizeof(bs_job_result);
    case 13: return sizeof(bs_server_stats);
    case 14: return sizeof(bs_render_settings);
    case 15: return sizeof(bs_render_view);
    default: return 0;
    }
}

int bs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int bs_device_memory(uint64_t *free_out, uint64_t *total_out) {
    size_t f = 0, t = 0;
    hipError_t e = hipMemGetInfo(&f, &t);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return bs::set_error(BS_ERR_DEVICE, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
    }
    if (free_out) *free_out = f;
    if (total_out) *total_out = t;
    return BS_OK;
}

int bs_device_name(char *buf, size_t len) {
    int dev = 0;
    hipDeviceProp_t p;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return bs::set_error(BS_ERR_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    if (buf && len) {
        // some drivers leave the marketing name empty; the architecture and
        // CU count identify the part either way
        const char *name = p.name[0] ? p.name : "AMD Instinct GPU";
        std::snprintf(buf, len, "%s (%s, %d CUs)", name, p.gcnArchName, p.multiProcessorCount);
    }
    return BS_OK;
}

}  // extern "C"
