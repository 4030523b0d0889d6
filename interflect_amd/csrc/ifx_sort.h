// Stable LSD radix sort (key, u32 value) and exclusive u32 scan built only
// from kernel launches: no memset, no host synchronisation, fixed scratch
// sizes.  That keeps them legal inside a captured HIP graph (the library
// sorts clear their counters with memset nodes).  8-bit digits; a block
// ranks its 1024-item tile in item order (wavefront ballots per digit bit,
// then wavefront prefix counts in LDS), so equal keys keep their input order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ifx {
namespace gsort {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kItems = 4;                     // per thread per tile
constexpr uint32_t kTile = kThreads * kItems;      // items per block
constexpr uint32_t kDigits = 256;

inline uint32_t tiles(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }

// Scratch words needed by sort_pairs / exclusive_scan for n items.
inline uint64_t sort_scratch_words(uint64_t n) {
    uint64_t t = tiles(n);
    uint64_t h = t * kDigits;
    return h + (h + kTile - 1) / kTile + 64;
}
inline uint64_t scan_scratch_words(uint64_t n) { return tiles(n) + 64; }

namespace detail {
// Internal linkage: every translation unit that includes this header keeps
// its own copy of the kernels.
namespace {

__device__ __forceinline__ uint32_t lane() { return threadIdx.x & 63u; }

// Lanes of this wavefront holding the same 8-bit digit (match by ballots).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
        uint64_t bb = __ballot(valid && ((d >> b) & 1u));
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    return valid ? m : 0ull;
}

// Block-wide exclusive scan of one u32 per thread (fixed order); returns the
// block total through `total`.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    const uint32_t t = threadIdx.x;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane() >= (uint32_t)o) x += y;
    }
    if (lane() == 63) lds[t >> 6] = x;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < (t >> 6); ++w) base += lds[w];
    total = 0;
    for (uint32_t w = 0; w < kThreads / 64; ++w) total += lds[w];
    __syncthreads();
    return base + x - v;
}

template <class K>
__global__ __launch_bounds__(kThreads) void k_hist(const K* __restrict__ keys, uint64_t n, uint32_t shift,
                                                   uint32_t nt, uint32_t* hist) {
    __shared__ uint32_t h[kDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * kTile;
    for (uint32_t j = 0; j < kItems; ++j) {
        uint64_t i = b0 + j * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nt + blockIdx.x] = h[threadIdx.x];  // digit-major
}

// Exclusive scan, pass 1: per-tile sums.
__global__ __launch_bounds__(kThreads) void k_tile_sums(const uint32_t* __restrict__ in, uint64_t n, uint32_t* sums) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t s = 0;
    for (uint32_t j = 0; j < kItems; ++j)
        if (b0 + j < n) s += in[b0 + j];
    uint32_t total;
    (void)block_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Pass 2: one block scans the tile sums in place (chunks of kThreads).
__global__ __launch_bounds__(kThreads) void k_scan_sums(uint32_t* sums, uint32_t nt) {
    __shared__ uint32_t lds[kThreads / 64];
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nt; c += kThreads) {
        uint32_t i = c + threadIdx.x;
        uint32_t v = i < nt ? sums[i] : 0u;
        uint32_t total;
        uint32_t ex = block_scan(v, lds, total);
        if (i < nt) sums[i] = carry + ex;
        carry += total;
    }
}

// Pass 3: scan within each tile, offset by the tile's scanned sum.
__global__ __launch_bounds__(kThreads) void k_tile_scan(const uint32_t* in, uint64_t n,
                                                        const uint32_t* __restrict__ sums, uint32_t* out) {
    __shared__ uint32_t lds[kThreads / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t s = 0;
    for (uint32_t j = 0; j < kItems; ++j) {
        v[j] = b0 + j < n ? in[b0 + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t ex = block_scan(s, lds, total) + sums[blockIdx.x];
    for (uint32_t j = 0; j < kItems; ++j) {
        if (b0 + j < n) out[b0 + j] = ex;
        ex += v[j];
    }
}

// Stable scatter of one tile: round r ranks items r*256 + t of the tile.
template <class K>
__global__ __launch_bounds__(kThreads) void k_scatter(const K* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                      uint64_t n, uint32_t shift, uint32_t nt,
                                                      const uint32_t* __restrict__ offs, K* keys_out,
                                                      uint32_t* vals_out) {
    __shared__ uint32_t base[kDigits];
    __shared__ uint32_t wc[kThreads / 64][kDigits];
    const uint32_t t = threadIdx.x, w = t >> 6;
    base[t] = offs[(uint64_t)t * nt + blockIdx.x];
    for (uint32_t k = 0; k < kThreads / 64; ++k) wc[k][t] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * kTile;
    for (uint32_t r = 0; r < kItems; ++r) {
        const uint64_t i = b0 + (uint64_t)r * kThreads + t;
        const bool valid = i < n;
        K key = valid ? keys[i] : K(0);
        uint32_t val = valid ? vals[i] : 0u;
        uint32_t d = (uint32_t)(key >> shift) & 0xffu;
        uint64_t m = match_digit(d, valid);
        uint64_t below = lane() == 0 ? 0ull : (m & ((~0ull) >> (64 - lane())));
        uint32_t rank = (uint32_t)__popcll(below);
        if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (uint32_t k = 0; k < w; ++k) pos += wc[k][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t k = 0; k < kThreads / 64; ++k) {
            add += wc[k][t];
            wc[k][t] = 0;
        }
        base[t] += add;
        __syncthreads();
    }
}

}  // namespace
}  // namespace detail

// out[i] = sum of in[0..i) for i < n.  `scratch` holds scan_scratch_words(n).
inline hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* scratch,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t nt = tiles(n);
    detail::k_tile_sums<<<nt, kThreads, 0, s>>>(in, n, scratch);
    detail::k_scan_sums<<<1, kThreads, 0, s>>>(scratch, nt);
    detail::k_tile_scan<<<nt, kThreads, 0, s>>>(in, n, scratch, out);
    return hipGetLastError();
}

// Sorts (keys, vals) by the low `bits` bits of the keys, stably.  The
// passes ping-pong between (keys, vals) and (keys_alt, vals_alt); the result
// ends in the pair named by the return flag `in_alt`.  `scratch` holds
// sort_scratch_words(n).
template <class K>
inline hipError_t sort_pairs(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, uint64_t n, uint32_t bits,
                             uint32_t* scratch, hipStream_t s, bool& in_alt) {
    in_alt = false;
    if (n == 0) return hipSuccess;
    const uint32_t nt = tiles(n);
    uint32_t* hist = scratch;
    uint32_t* part = scratch + (uint64_t)nt * kDigits;
    K *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        detail::k_hist<K><<<nt, kThreads, 0, s>>>(ki, n, shift, nt, hist);
        hipError_t e = exclusive_scan(hist, hist, (uint64_t)nt * kDigits, part, s);
        if (e != hipSuccess) return e;
        detail::k_scatter<K><<<nt, kThreads, 0, s>>>(ki, vi, n, shift, nt, hist, ko, vo);
        K* tk = ki;
        ki = ko;
        ko = tk;
        uint32_t* tv = vi;
        vi = vo;
        vo = tv;
        in_alt = !in_alt;
    }
    return hipGetLastError();
}

// Sorts (keys, vals) by the low `bits` bits of the keys, stably, into
// (keys_out, vals_out), leaving the inputs untouched: the passes alternate
// between the output pair and (keys_tmp, vals_tmp) so that the last one
// lands in the output.  `scratch` holds sort_scratch_words(n).
template <class K>
inline hipError_t sort_pairs_to(const K* keys, const uint32_t* vals, K* keys_out, uint32_t* vals_out, K* keys_tmp,
                                uint32_t* vals_tmp, uint64_t n, uint32_t bits, uint32_t* scratch, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t passes = (bits + 7u) / 8u;
    if (passes == 0) {
        hipError_t e = hipMemcpyAsync(keys_out, keys, sizeof(K) * n, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(vals_out, vals, 4ull * n, hipMemcpyDeviceToDevice, s);
    }
    const uint32_t nt = tiles(n);
    uint32_t* hist = scratch;
    uint32_t* part = scratch + (uint64_t)nt * kDigits;
    const K* ki = keys;
    const uint32_t* vi = vals;
    for (uint32_t p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1u - p) & 1u) == 0u;
        K* ko = to_out ? keys_out : keys_tmp;
        uint32_t* vo = to_out ? vals_out : vals_tmp;
        detail::k_hist<K><<<nt, kThreads, 0, s>>>(ki, n, 8u * p, nt, hist);
        hipError_t e = exclusive_scan(hist, hist, (uint64_t)nt * kDigits, part, s);
        if (e != hipSuccess) return e;
        detail::k_scatter<K><<<nt, kThreads, 0, s>>>(ki, vi, n, 8u * p, nt, hist, ko, vo);
        ki = ko;
        vi = vo;
    }
    return hipGetLastError();
}

}  // namespace gsort
}  // namespace ifx
