// Device primitives for the build: exclusive scan and stable LSD radix sort.
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8 items,
// blocked), tile sums scanned recursively, then a block scan per tile.
// Radix sort: 8-bit digits; one launch counts every digit position, then one
// Onesweep launch per digit (decoupled look-back over tiles for the digit
// offsets, then the scatter).  The scatter ranks keys stably inside a tile
// with wave-wide digit match masks (8 ballots per key) and per-wave digit
// counts in LDS, so equal keys keep their input order: with iota values this
// is std::sort on (code, index) pairs (bvh.cpp:223-231).  The reference's own GPU sort never got past a
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
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;   // keys per tile
constexpr int kSortHistBlocks = 512;
// Look-back status word of (tile, digit): flags in the top two bits, a count
// (below 2^30) in the others.  Aggregate = the tile's own count of the digit,
// prefix = its count plus every earlier tile's.
constexpr uint32_t kDigitAgg = 1u << 30;
constexpr uint32_t kDigitPre = 2u << 30;
constexpr uint32_t kDigitFlags = 3u << 30;
constexpr uint32_t kSortSpinLimit = 1u << 24;   // ~0.5 s of s_sleep(1) polls

// Counts of every digit at every digit position in one read of the keys:
// hist[p * 256 + d] = keys whose digit p (bits 8p .. 8p+7) is d, p < n_digits.
// Each workgroup counts its tiles in LDS and adds its non-zero bins.
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                            uint32_t n_digits, uint32_t n_tiles,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[4][256];
#pragma unroll
    for (int p = 0; p < 4; ++p) s_h[p][threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t base = t * kSortTile;
        for (int r = 0; r < kSortRounds; ++r) {
            const uint32_t i = base + r * kSortThreads + threadIdx.x;
            if (i < n) {
                const uint32_t k = keys[i];
                for (uint32_t p = 0; p < n_digits; ++p) atomicAdd(&s_h[p][(k >> (8 * p)) & 0xFFu], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < n_digits; ++p)
        if (s_h[p][threadIdx.x]) atomicAdd(&hist[p * 256 + threadIdx.x], s_h[p][threadIdx.x]);
}

// Exclusive scan of each digit position's 256 counts: base[p * 256 + d] =
// keys whose digit p is below d.  One workgroup.
__global__ __launch_bounds__(kSortThreads) void k_sort_digit_base(const uint32_t* __restrict__ hist,
                                                                  uint32_t n_digits, uint32_t* __restrict__ base) {
    __shared__ uint32_t s_w[kSortThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    for (uint32_t p = 0; p < n_digits; ++p) {
        const uint32_t c = hist[p * 256 + threadIdx.x];
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) s_w[wid] = x;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t w = 0; w < wid; ++w) before += s_w[w];
        base[p * 256 + threadIdx.x] = before + x - c;
        __syncthreads();
    }
}

// One digit pass in one launch (Adinets & Merrill, "Onesweep: A Faster Least
// Significant Digit Radix Sort for GPUs", 2022): a workgroup takes the next
// tile (atomic counter, so every tile it waits on is already running),
// counts the tile's digits in LDS, publishes them and sums the counts of the
// earlier tiles by decoupled look-back per digit (Merrill & Garland 2016),
// then scatters each key to digit_base[d] + earlier tiles' count + its rank
// in the tile.  Ranks are stable: rounds in order, waves in order, lanes in
// order (wave-wide digit match masks, 8 ballots per key).  A look-back that
// waits longer than kSortSpinLimit polls sets *err and goes on.
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n, uint32_t shift,
    const uint32_t* __restrict__ digit_base, uint32_t* __restrict__ tile_ctr, uint32_t* __restrict__ status,
    uint32_t* __restrict__ err, uint32_t* __restrict__ out_k, uint32_t* __restrict__ out_v) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[256];
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_wcnt[kSortThreads / 64][256];
    const uint32_t tid = threadIdx.x, wid = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    s_cnt[tid] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kSortTile;
    uint32_t key[kSortRounds], val[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t i = base + r * kSortThreads + tid;
        key[r] = 0;
        val[r] = 0;
        if (i < n) {
            key[r] = keys[i];
            val[r] = vals[i];
            atomicAdd(&s_cnt[(key[r] >> shift) & 0xFFu], 1u);
        }
    }
    __syncthreads();
    // thread d: digit d's count through this tile
    const uint32_t c = s_cnt[tid];
    uint32_t* const mine = status + (size_t)tile * 256 + tid;
    uint32_t excl = 0;
    if (tile == 0) {
        __hip_atomic_store(mine, kDigitPre | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(mine, kDigitAgg | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t k = tile - 1, spins = 0;
        for (;;) {
            const uint32_t w = __hip_atomic_load(status + (size_t)k * 256 + tid, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t f = w & kDigitFlags;
            if (f == 0) {
                if (++spins > kSortSpinLimit) {
                    atomicOr(err, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += w & ~kDigitFlags;
            if (f == kDigitPre) break;
            if (k == 0) {   // tile 0 publishes a prefix: a broken invariant
                atomicOr(err, 1u);
                break;
            }
            --k;
        }
        __hip_atomic_store(mine, kDigitPre | (excl + c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_off[tid] = digit_base[tid] + excl;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) s_wcnt[w][tid] = 0;
        __syncthreads();
        const uint32_t i = base + r * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t d = (key[r] >> shift) & 0xFFu;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (valid && rank == 0) s_wcnt[wid][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_off[d] + rank;
            for (uint32_t w = 0; w < wid; ++w) pos += s_wcnt[w][d];
            out_k[pos] = key[r];
            out_v[pos] = val[r];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; ++w) add += s_wcnt[w][tid];
        s_off[tid] += add;
        __syncthreads();
    }
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
                            uint32_t* tmp_v, uint64_t n, uint32_t key_bits, uint32_t* err, hipStream_t st) {
    if (n == 0 || key_bits == 0) return RTQ_OK;
    if (n >= (1ull << 30)) return fail(ctx, RTQ_ECAPACITY, "sort: n must be < 2^30");
    const uint32_t nn = (uint32_t)n;
    const uint32_t n_tiles = (nn + kSortTile - 1) / kSortTile;
    const uint32_t passes = (key_bits + 7) / 8;
    // digit counts and bases of every pass, the passes' tile counters and one
    // pass's look-back status words (cleared before each pass): the passes
    // queue without synchronising
    Temps tmp(ctx);
    uint32_t *hist = nullptr, *dbase = nullptr, *ctr = nullptr, *status = nullptr;
    RTQ_HIP(ctx, tmp.get(&status, (size_t)n_tiles * 256));
    RTQ_HIP(ctx, tmp.get(&hist, 3 * 1024 + 64));
    dbase = hist + 1024;
    ctr = hist + 2048;
    RTQ_HIP(ctx, hipMemsetAsync(hist, 0, sizeof(uint32_t) * (3 * 1024 + 64), st));
    hipLaunchKernelGGL(k_sort_hist, dim3(std::min<uint32_t>(n_tiles, kSortHistBlocks)), dim3(kSortThreads), 0, st,
                       keys, nn, passes, n_tiles, hist);
    RTQ_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_sort_digit_base, dim3(1), dim3(kSortThreads), 0, st, hist, passes, dbase);
    RTQ_HIP(ctx, hipGetLastError());
    uint32_t *src_k = keys, *src_v = vals, *dst_k = tmp_k, *dst_v = tmp_v;
    for (uint32_t p = 0; p < passes; ++p) {
        RTQ_HIP(ctx, hipMemsetAsync(status, 0, sizeof(uint32_t) * n_tiles * 256, st));
        hipLaunchKernelGGL(k_onesweep, dim3(n_tiles), dim3(kSortThreads), 0, st, src_k, src_v, nn, 8 * p,
                           dbase + 256 * p, ctr + p, status, err, dst_k, dst_v);
        RTQ_HIP(ctx, hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (passes & 1) {
        RTQ_HIP(ctx, hipMemcpyAsync(keys, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        RTQ_HIP(ctx, hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return RTQ_OK;
}

}  // namespace rtq
