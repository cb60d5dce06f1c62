// noalloc.hip -- bs_reduce_noalloc: the diagnostics reduction of
// bs_sim_diagnostics over a lattice the caller holds in device memory, for
// callers that cannot allocate (code under hipGraph capture, tight serving
// loops).  The reduction's first pass leaves one partial result per block
// (kReduceBlocks of them, whatever the lattice's size); here they go to a
// fixed device array instead of memory allocated per handle.  The kernels
// and the grid are bs_sim_diagnostics' own, so the sums are the same bits.
#include <cstring>
#include <mutex>

#include "device.h"

namespace bs {

// The partials of one reduction, 40 KB, on every device the library runs on
// (held until the process ends).  In host code this name is only a stand-in
// for the symbol: its device address comes from hipGetSymbolAddress, per
// device, below.
static __device__ __align__(256) ReducePartial scratch[kReduceBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
ReducePartial *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// `scratch` on the current device, looked up on the first call there
int scratch_ptr(ReducePartial **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(BS_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(BS_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(BS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (ReducePartial *)p;
    }
    *out = g_scratch[dev];
    return BS_OK;
}
}  // namespace

}  // namespace bs

using namespace bs;

extern "C" {

static_assert(sizeof(bs_reduce_result) == sizeof(ReduceResult), "bs_reduce_result mirrors the reduction's result");

int bs_reduce_noalloc(const bs_sim_params *params, int32_t precision, const void *field, const void *fluid,
                      void *out_device, void *stream) {
    if (!params || !field || !fluid || !out_device) return set_error(BS_ERR_INVALID, "null argument");
    if (precision != BS_PRECISION_F32 && precision != BS_PRECISION_F64)
        return set_error(BS_ERR_INVALID, "precision must be BS_PRECISION_F32 or BS_PRECISION_F64");
    // the lattices a spec may describe (spec_validate): every axis >= 8, and
    // an x-y plane below 2^28 cells (f32) / 2^27 (f64)
    const uint32_t g[3] = {params->nx, params->ny, params->nz};
    for (int a = 0; a < 3; ++a)
        if (g[a] < 8) return set_error(BS_ERR_INVALID, "grid axis " + std::to_string(a) + " is below 8 cells");
    const uint64_t plane = (uint64_t)g[0] * g[1];
    if (plane >= (precision == BS_PRECISION_F64 ? (1ull << 27) : (1ull << 28)))
        return set_error(BS_ERR_INVALID, "an x-y plane of " + std::to_string(plane) + " cells is too large");
    ReducePartial *partials = nullptr;
    int rc = scratch_ptr(&partials);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (precision == BS_PRECISION_F64) {
        const Params<double> P = widen_params<double>(*params);
        e = launch_reduce<double>((const V2<double> *)field, (const V4<double> *)fluid, P, partials,
                                  (ReduceResult *)out_device, st);
    } else {
        const Params<float> P = widen_params<float>(*params);
        e = launch_reduce<float>((const V2<float> *)field, (const V4<float> *)fluid, P, partials,
                                 (ReduceResult *)out_device, st);
    }
    if (e != hipSuccess) return set_error(BS_ERR_DEVICE, std::string("reduction launch: ") + hipGetErrorString(e));
    return BS_OK;
}

}  // extern "C"
