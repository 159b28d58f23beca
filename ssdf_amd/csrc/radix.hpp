// radix.hpp — stable LSD radix sort of (key, uint32 value) pairs on gfx950.
//
// One pass per 8-bit digit, three kernels per pass:
//   hist    : per-tile digit histogram in LDS, written digit-major (d * ntiles + tile);
//   scan    : one-block exclusive scan of that digit-major table -> global scatter offsets;
//   scatter : each wave ranks its 16 chunks of 64 keys in order — the lanes holding the same digit
//             are found with 8 ballots (one per digit bit), a lane's rank is the popcount of its
//             peers below it — so the pass is stable without any LDS sort.
// Tiles are 4096 keys (256 threads x 16).  All launches go to the caller's stream and nothing is
// read back, so a sort can sit inside a captured graph.
#pragma once

#include "common.hpp"

namespace ssdf_amd {

constexpr int kRadixBlock = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixBlock * kRadixItems;

inline unsigned radix_tiles(size_t n) { return (unsigned)((n + kRadixTile - 1) / kRadixTile); }

// scratch: 256 * radix_tiles(n) uint32
inline size_t radix_scratch_words(size_t n) { return (size_t)256 * radix_tiles(n) + 64; }

template <typename K>
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const K *__restrict__ keys, size_t n, int shift,
                                                            uint32_t *__restrict__ hist, unsigned ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int i = 0; i < kRadixItems; i++) {
        const size_t idx = base + (size_t)i * kRadixBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m words in place (one block of 1024 threads; lbvh.hip).
__global__ __launch_bounds__(1024) void k_scan_exclusive_u32(uint32_t *__restrict__ data, size_t m);

template <typename K>
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const K *__restrict__ kin,
                                                               const uint32_t *__restrict__ vin,
                                                               K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                               size_t n, int shift, const uint32_t *__restrict__ offs,
                                                               unsigned ntiles) {
    __shared__ uint32_t cnt[4][256];
    __shared__ uint32_t base[4][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < 1024; i += kRadixBlock) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const size_t wbase = (size_t)blockIdx.x * kRadixTile + (size_t)w * (64 * kRadixItems);
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    K key[kRadixItems];
    uint32_t val[kRadixItems], dig[kRadixItems];
    uint64_t peers[kRadixItems];
#pragma unroll
    for (int c = 0; c < kRadixItems; c++) {
        const size_t idx = wbase + (size_t)c * 64 + lane;
        const bool ok = idx < n;
        key[c] = ok ? kin[idx] : K(0);
        val[c] = ok ? vin[idx] : 0u;
        const uint32_t d = (uint32_t)(key[c] >> shift) & 255u;
        dig[c] = d;
        uint64_t pm = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            pm &= bit ? m : ~m;
        }
        peers[c] = ok ? pm : 0ull;
        if (ok && (pm & below) == 0) cnt[w][d] += (uint32_t)__popcll(pm);
    }
    __syncthreads();
    {
        const uint32_t d = threadIdx.x;
        uint32_t run = offs[(size_t)d * ntiles + blockIdx.x];
#pragma unroll
        for (int ww = 0; ww < 4; ww++) {
            base[ww][d] = run;
            run += cnt[ww][d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kRadixItems; c++) {
        const size_t idx = wbase + (size_t)c * 64 + lane;
        const bool ok = idx < n;
        const uint32_t d = dig[c];
        const uint64_t pm = peers[c];
        const uint32_t b = base[w][d];
        const uint32_t rank = (uint32_t)__popcll(pm & below);
        if (ok) {
            kout[b + rank] = key[c];
            vout[b + rank] = val[c];
        }
        if (ok && (pm & below) == 0) base[w][d] = b + (uint32_t)__popcll(pm);
    }
}

// Sorts (keys, vals) by the low `key_bits` bits of the keys.  The buffers ping-pong; on return
// *out_keys / *out_vals point at whichever pair holds the sorted data.
template <typename K>
hipError_t radix_sort_pairs(K *keys, K *keys_alt, uint32_t *vals, uint32_t *vals_alt, size_t n, int key_bits,
                            uint32_t *scratch, hipStream_t stream, K **out_keys, uint32_t **out_vals) {
    K *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    if (n > 1) {
        const unsigned nt = radix_tiles(n);
        for (int shift = 0; shift < key_bits; shift += 8) {
            hipLaunchKernelGGL(k_radix_hist<K>, dim3(nt), dim3(kRadixBlock), 0, stream, ki, n, shift, scratch, nt);
            hipLaunchKernelGGL(k_scan_exclusive_u32, dim3(1), dim3(1024), 0, stream, scratch, (size_t)256 * nt);
            hipLaunchKernelGGL(k_radix_scatter<K>, dim3(nt), dim3(kRadixBlock), 0, stream, ki, vi, ko, vo, n, shift,
                               scratch, nt);
            K *tk = ki;
            ki = ko, ko = tk;
            uint32_t *tv = vi;
            vi = vo, vo = tv;
        }
    }
    *out_keys = ki;
    *out_vals = vi;
    return hipGetLastError();
}

}  // namespace ssdf_amd
