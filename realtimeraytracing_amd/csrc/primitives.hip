// Device primitives for the build: exclusive scan and stable LSD radix sort.
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8 items,
// blocked), tile sums scanned recursively, then a block scan per tile.
// Radix sort: 8-bit digits, one histogram + one scatter launch per digit.  The
// scatter ranks keys stably inside a tile with wave-wide digit match masks
// (8 ballots per key) and per-wave digit counts in LDS, so equal keys keep
// their input order: with iota values this is std::sort on (code, index)
// pairs (bvh.cpp:223-231).  The reference's own GPU sort never got past a
// set-bit histogram and a per-group scan (histogramOfGlobalDigitCounts.glsl,
// prefixSumOfGlobalDigitCounts.glsl, chainedScanDigitBinning.glsl:31-32 empty).
#include "rtq_internal.hpp"

namespace rtq {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline uint32_t shfl_up_u32(uint32_t x, int o) { return __shfl_up(x, o, 64); }

struct U1 {
    uint32_t x;
    __device__ U1 operator+(const U1& b) const { return {x + b.x}; }
    __device__ U1 shfl_up(int o) const { return {shfl_up_u32(x, o)}; }
};
struct U2 {
    uint32_t x, y;
    __device__ U2 operator+(const U2& b) const { return {x + b.x, y + b.y}; }
    __device__ U2 shfl_up(int o) const { return {shfl_up_u32(x, o), shfl_up_u32(y, o)}; }
};

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_tile_reduce(const T* __restrict__ in, uint64_t n,
                                                              T* __restrict__ sums) {
    __shared__ T s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    T acc{};
    for (int r = 0; r < kScanItems; ++r) {
        uint64_t i = base + (uint64_t)r * kScanThreads + threadIdx.x;
        if (i < n) acc = acc + in[i];
    }
    for (int o = 1; o < 64; o <<= 1) {
        T y = acc.shfl_up(o);
        if ((threadIdx.x & 63) >= (unsigned)o) acc = acc + y;
    }
    if ((threadIdx.x & 63) == 63) s_wave[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t{};
        for (int w = 0; w < kScanThreads / 64; ++w) t = t + s_wave[w];
        sums[blockIdx.x] = t;
    }
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const T* __restrict__ in, T* __restrict__ out,
                                                            uint64_t n, const T* __restrict__ offsets,
                                                            T* __restrict__ total) {
    __shared__ T s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T tsum{};
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        uint64_t i = base + r;
        v[r] = (i < n) ? in[i] : T{};
        tsum = tsum + v[r];
    }
    T x = tsum;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        T y = x.shfl_up(o);
        if (lane >= o) x = x + y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    T off = offsets ? offsets[blockIdx.x] : T{};
    T block_total{};
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) off = off + s_wave[w];
        block_total = block_total + s_wave[w];
    }
    // exclusive prefix of this thread = off + (x - tsum); computed additively
    T run = off;
    {
        // x - tsum == inclusive-of-previous-lanes; rebuild it with a shifted scan value
        T prev = x.shfl_up(1);
        if (lane > 0) run = run + prev;
    }
#pragma unroll
    for (int r = 0; r < kScanItems; ++r) {
        uint64_t i = base + r;
        if (i < n) out[i] = run;
        run = run + v[r];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        T o = offsets ? offsets[blockIdx.x] : T{};
        *total = o + block_total;
    }
}

template <typename T>
rtq_status scan_impl(rtq_ctx* ctx, const T* in, T* out, uint64_t n, T* total, hipStream_t st) {
    if (n == 0) {
        if (total) RTQ_HIP(ctx, hipMemsetAsync(total, 0, sizeof(T), st));
        return RTQ_OK;
    }
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kScanThreads), 0, st, in, out, n,
                           (const T*)nullptr, total);
        RTQ_HIP(ctx, hipGetLastError());
        return RTQ_OK;
    }
    Temps tmp(ctx);
    T* sums = nullptr;
    RTQ_HIP(ctx, tmp.get(&sums, 2 * nb));
    T* offs = sums + nb;
    hipLaunchKernelGGL(k_tile_reduce<T>, dim3((uint32_t)nb), dim3(kScanThreads), 0, st, in, n, sums);
    RTQ_HIP(ctx, hipGetLastError());
    const rtq_status s = scan_impl<T>(ctx, sums, offs, nb, nullptr, st);
    if (s != RTQ_OK) return s;
    hipLaunchKernelGGL(k_tile_scan<T>, dim3((uint32_t)nb), dim3(kScanThreads), 0, st, in, out, n,
                       (const T*)offs, total);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

// Elements of workspace scan_impl_ws needs for n elements (the tile sums and
// their offsets at every level).
template <typename T>
uint64_t scan_workspace(uint64_t n) {
    uint64_t total = 0;
    while (n > (uint64_t)kScanTile) {
        const uint64_t nb = (n + kScanTile - 1) / kScanTile;
        total += 2 * nb;
        n = nb;
    }
    return total;
}

// The same scan with the caller's workspace: no allocation and no
// synchronisation, so it can be queued back to back.
template <typename T>
rtq_status scan_impl_ws(rtq_ctx* ctx, const T* in, T* out, uint64_t n, T* total, T* ws, hipStream_t st) {
    if (n == 0) return fail(ctx, RTQ_EINTERNAL, "scan_ws: n must be > 0");
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kScanThreads), 0, st, in, out, n,
                           (const T*)nullptr, total);
        RTQ_HIP(ctx, hipGetLastError());
        return RTQ_OK;
    }
    T* sums = ws;
    T* offs = ws + nb;
    hipLaunchKernelGGL(k_tile_reduce<T>, dim3((uint32_t)nb), dim3(kScanThreads), 0, st, in, n, sums);
    RTQ_HIP(ctx, hipGetLastError());
    rtq_status s = scan_impl_ws<T>(ctx, sums, offs, nb, nullptr, ws + 2 * nb, st);
    if (s != RTQ_OK) return s;
    hipLaunchKernelGGL(k_tile_scan<T>, dim3((uint32_t)nb), dim3(kScanThreads), 0, st, in, out, n,
                       (const T*)offs, total);
    RTQ_HIP(ctx, hipGetLastError());
    return RTQ_OK;
}

// ---------------------------------------------------------------- radix sort

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;

__global__ __launch_bounds__(kSortThreads) void k_digit_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t shift, uint32_t n_tiles,
                                                             uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * n_tiles + blockIdx.x] = s_h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_digit_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n, uint32_t shift,
    uint32_t n_tiles, const uint32_t* __restrict__ offsets, uint32_t* __restrict__ out_k,
    uint32_t* __restrict__ out_v) {
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_wcnt[kSortThreads / 64][256];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    s_off[threadIdx.x] = offsets[threadIdx.x * n_tiles + blockIdx.x];
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) s_wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        uint32_t key = 0, val = 0, d = 0;
        if (valid) {
            key = keys[i];
            val = vals[i];
            d = (key >> shift) & 0xFFu;
        }
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        if (valid && rank == 0) s_wcnt[wid][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_off[d] + rank;
            for (int w = 0; w < wid; ++w) pos += s_wcnt[w][d];
            out_k[pos] = key;
            out_v[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) add += s_wcnt[w][threadIdx.x];
        s_off[threadIdx.x] += add;
        __syncthreads();
    }
    (void)lane;
}

}  // namespace

rtq_status scan_u32(rtq_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total,
                    hipStream_t st) {
    return scan_impl<U1>(ctx, reinterpret_cast<const U1*>(in), reinterpret_cast<U1*>(out), n,
                         reinterpret_cast<U1*>(total), st);
}

uint64_t scan_u32x2_workspace(uint64_t n) { return scan_workspace<U2>(n); }

rtq_status scan_u32x2_ws(rtq_ctx* ctx, const uint2* in, uint2* out, uint64_t n, uint2* total, uint2* ws,
                         hipStream_t st) {
    return scan_impl_ws<U2>(ctx, reinterpret_cast<const U2*>(in), reinterpret_cast<U2*>(out), n,
                            reinterpret_cast<U2*>(total), reinterpret_cast<U2*>(ws), st);
}

rtq_status radix_sort_pairs(rtq_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tmp_k,
                            uint32_t* tmp_v, uint64_t n, uint32_t key_bits, hipStream_t st) {
    if (n == 0 || key_bits == 0) return RTQ_OK;
    if (n >= (1ull << 31)) return fail(ctx, RTQ_ECAPACITY, "sort: n must be < 2^31");
    const uint32_t nn = (uint32_t)n;
    const uint32_t n_tiles = (nn + kSortTile - 1) / kSortTile;
    // digit counts, their offsets and the offsets scan's workspace in one
    // block, so the passes queue without synchronising
    const uint64_t n_hist = 256ull * n_tiles;
    const uint64_t ws = scan_workspace<U1>(n_hist);
    Temps tmp(ctx);
    uint32_t* hist = nullptr;
    RTQ_HIP(ctx, tmp.get(&hist, 2ull * n_hist + ws));
    uint32_t* offs = hist + n_hist;
    U1* scan_ws = reinterpret_cast<U1*>(offs + n_hist);
    uint32_t *src_k = keys, *src_v = vals, *dst_k = tmp_k, *dst_v = tmp_v;
    rtq_status s = RTQ_OK;
    int passes = 0;
    for (uint32_t shift = 0; shift < key_bits && s == RTQ_OK; shift += 8, ++passes) {
        hipLaunchKernelGGL(k_digit_hist, dim3(n_tiles), dim3(kSortThreads), 0, st, src_k, nn, shift,
                           n_tiles, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { s = hip_fail(ctx, e, "k_digit_hist"); break; }
        s = scan_impl_ws<U1>(ctx, reinterpret_cast<const U1*>(hist), reinterpret_cast<U1*>(offs), n_hist,
                             nullptr, scan_ws, st);
        if (s != RTQ_OK) break;
        hipLaunchKernelGGL(k_digit_scatter, dim3(n_tiles), dim3(kSortThreads), 0, st, src_k, src_v, nn,
                           shift, n_tiles, offs, dst_k, dst_v);
        e = hipGetLastError();
        if (e != hipSuccess) { s = hip_fail(ctx, e, "k_digit_scatter"); break; }
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (s == RTQ_OK && (passes & 1)) {
        hipError_t e = hipMemcpyAsync(keys, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) s = hip_fail(ctx, e, "sort copy-back");
    }
    return s;
}

}  // namespace rtq
