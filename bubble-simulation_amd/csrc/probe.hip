// probe.hip -- achievable HBM bandwidth of this device (bs_hbm_probe).
//
// Three streaming kernels over buffers far larger than the last-level caches:
// read-only (non-temporal loads, a reduction that the compiler must keep),
// write-only (non-temporal stores) and copy.  Each runs at a few grid sizes
// (8 .. 64 workgroups of 256 threads per CU, grid-stride loops) and the best
// rate is kept: the practical ceilings a bandwidth-bound kernel's measured
// traffic is compared against, next to the published peak.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "internal.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_probe_read(const f4 *__restrict__ a, size_t n, float *__restrict__ sink) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    f4 s = {0.0f, 0.0f, 0.0f, 0.0f};
    for (; i < n; i += stride) s += __builtin_nontemporal_load(a + i);
    // the buffers hold zeros: the store never happens, but the sum is live
    if (s.x + s.y + s.z + s.w != 0.0f) sink[0] = s.x + s.y + s.z + s.w;
}

__global__ void __launch_bounds__(256) k_probe_write(f4 *__restrict__ b, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    for (; i < n; i += stride) __builtin_nontemporal_store(z, b + i);
}

__global__ void __launch_bounds__(256) k_probe_copy(const f4 *__restrict__ a, f4 *__restrict__ b, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

}  // namespace

extern "C" int bs_hbm_probe(uint64_t bytes, uint32_t repeats, bs_hbm_rates *out) {
    if (!out || bytes < (1u << 20) || repeats == 0) return bs::set_error(BS_ERR_INVALID, "bs_hbm_probe: bad argument");
    const size_t n = bytes / sizeof(f4);
    int dev = 0, cus = 0;
    f4 *a = nullptr, *b = nullptr;
    float *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best[3] = {0.0, 0.0, 0.0};
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipMalloc((void **)&a, n * sizeof(f4));
    if (e == hipSuccess) e = hipMalloc((void **)&b, n * sizeof(f4));
    if (e == hipSuccess) e = hipMalloc((void **)&sink, sizeof(float));
    if (e == hipSuccess) e = hipMemset(a, 0, n * sizeof(f4));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int per_cu[4] = {8, 16, 32, 64};
    for (int kind = 0; kind < 3 && e == hipSuccess; ++kind) {
        for (int g = 0; g < 4 && e == hipSuccess; ++g) {
            const unsigned grid = (unsigned)(per_cu[g] * std::max(cus, 1));
            auto launch = [&]() {
                if (kind == 0) hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, 0, a, n, sink);
                else if (kind == 1) hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(256), 0, 0, b, n);
                else hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
                return hipGetLastError();
            };
            e = launch();  // warm
            if (e == hipSuccess) e = hipEventRecord(e0, 0);
            for (uint32_t r = 0; r < repeats && e == hipSuccess; ++r) e = launch();
            if (e == hipSuccess) e = hipEventRecord(e1, 0);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && ms > 0.0f) {
                const double moved = (kind == 2 ? 2.0 : 1.0) * (double)(n * sizeof(f4)) * repeats;
                best[kind] = std::max(best[kind], moved / (ms * 1e-3) / 1e9);
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (b) (void)hipFree(b);
    if (a) (void)hipFree(a);
    if (e != hipSuccess)
        return bs::set_error(e == hipErrorOutOfMemory ? BS_ERR_OOM : BS_ERR_DEVICE,
                             std::string("bs_hbm_probe: ") + hipGetErrorString(e));
    out->read_gbps = best[0];
    out->write_gbps = best[1];
    out->copy_gbps = best[2];
    return BS_OK;
}
