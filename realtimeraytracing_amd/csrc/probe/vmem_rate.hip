// Vector-memory rate probe (measurement only; bench.py loads it beside the
// library, nothing in librtq uses it).
//
// Every lane follows its own chain through an array of 128-B node slots: each
// step loads `loads` 16-B words of the current slot (global_load_dwordx4, as
// a traversal node visit does), mixes them and picks the next slot from the
// result.  The rate of these per-lane loads, with thousands of chains in
// flight, is the ceiling of a kernel whose lanes fetch scattered 16-B words:
// an array far larger than L1 gives the rate of loads that miss L1, a few
// slots shared by every lane the rate of loads that hit it.  With fewer than
// 64 active lanes per wave (the others skip the chain) the same launch shows
// what a partially active wave pays per load instruction.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr int kSteps = 128;

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_fill(uint4* __restrict__ slots, uint64_t words) {
    const uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < words) {
        const uint32_t s = (uint32_t)i * 4u;
        slots[i] = make_uint4(mix32(s), mix32(s + 1), mix32(s + 2), mix32(s + 3));
    }
}

template <int L>
__global__ __launch_bounds__(kThreads) void k_chain(const uint4* __restrict__ slots, uint32_t n, uint32_t active,
                                                    uint32_t* __restrict__ out) {
    uint32_t acc = 0;
    if ((threadIdx.x & 63u) < active) {   // lanes 0 .. active-1 of each wave
        uint32_t cur = mix32(blockIdx.x * kThreads + threadIdx.x) % n;
        for (int s = 0; s < kSteps; ++s) {
            const uint4* p = slots + 8ull * cur;
            uint32_t h = 0;
#pragma unroll
            for (int k = 0; k < L; ++k) {
                const uint4 v = p[k];
                h += (v.x ^ v.y) + (v.z ^ (v.w + (uint32_t)k));
            }
            acc += h;
            cur = mix32(h + (uint32_t)s) % n;
        }
    }
    out[blockIdx.x * kThreads + threadIdx.x] = acc;   // keeps the loads live
}

using ChainKernel = void (*)(const uint4*, uint32_t, uint32_t, uint32_t*);
constexpr ChainKernel kChains[8] = {k_chain<1>, k_chain<2>, k_chain<3>, k_chain<4>,
                                    k_chain<5>, k_chain<6>, k_chain<7>, k_chain<8>};

}  // namespace

// Per-lane 16-B loads per second over n_slots 128-B slots, `loads` (1..8) per
// step, blocks_per_cu 256-thread workgroups per CU, active_lanes (1..64) lanes
// of each wave loading (the rate counts their loads only); best of 3 launches.
// Returns 0, or a HIP error code.
extern "C" __attribute__((visibility("default"))) int rtqprobe_vmem_rate_lanes(int device, uint32_t loads,
                                                                               uint32_t n_slots, uint32_t blocks_per_cu,
                                                                               uint32_t active_lanes, double* rate) {
    if (!rate || loads < 1 || loads > 8 || n_slots < 1 || blocks_per_cu < 1 || blocks_per_cu > 32 ||
        active_lanes < 1 || active_lanes > 64)
        return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return (int)e;
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * blocks_per_cu;
    uint4* slots = nullptr;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const uint64_t words = 8ull * n_slots;
    e = hipMalloc(&slots, words * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&out, (size_t)blocks * kThreads * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill, dim3((uint32_t)((words + kThreads - 1) / kThreads)), dim3(kThreads), 0, 0, slots,
                           words);
        e = hipGetLastError();
    }
    float best = 0.0f;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {   // the first launch warms the caches
        e = hipEventRecord(e0, 0);
        hipLaunchKernelGGL(kChains[loads - 1], dim3(blocks), dim3(kThreads), 0, 0, slots, n_slots, active_lanes, out);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && rep > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    if (e == hipSuccess) *rate = (double)blocks * (kThreads / 64) * active_lanes * kSteps * loads / (best * 1e-3);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if (slots) (void)hipFree(slots);
    return (int)e;
}

// The same with every lane of each wave loading.
extern "C" __attribute__((visibility("default"))) int rtqprobe_vmem_rate(int device, uint32_t loads, uint32_t n_slots,
                                                                         uint32_t blocks_per_cu, double* rate) {
    return rtqprobe_vmem_rate_lanes(device, loads, n_slots, blocks_per_cu, 64, rate);
}
