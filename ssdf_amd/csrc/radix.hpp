// radix.hpp — stable LSD radix sort of (key, uint32 value) pairs on gfx950.
//
// One pass per 8-bit digit:
//   hist    : per-tile digit histogram in LDS;
//   scan    : exclusive scan of the (digit, tile) table -> global scatter offsets;
//   scatter : each wave ranks its 4 chunks of 64 keys in order — the lanes holding the same digit
//             are found with 8 ballots (one per digit bit), a lane's rank is the popcount of its
//             peers below it — so the pass is stable without any LDS sort.
// Up to kFuseTiles tiles the scan lives inside the other two kernels: hist writes its counts
// tile-major and adds them into the pass's digit totals (zeroed once per sort), and each scatter
// block scans the 256 totals and sums the earlier tiles' counts itself (coalesced, L2-resident,
// four threads per digit).  Larger sorts take a third kernel between them over a digit-major
// table: one block per digit scans that digit's row of tile counts, offset by the totals of the
// lower digits.
// Tiles are 4096 keys, one 1024-thread block (16 waves x 4 chunks) each: the sorts of a mesh build
// are 0.2-3 M keys, i.e. 50-750 tiles, and with 256-thread blocks each wave ranked 16 chunks in
// series on an otherwise idle SIMD.  All launches go to the caller's stream and nothing is read
// back, so a sort can sit inside a captured graph.
#pragma once

#include "common.hpp"

namespace ssdf_amd {

constexpr int kRadixBlock = 1024;  // hist and scatter blocks
constexpr int kRadixWaves = kRadixBlock / 64;
constexpr int kRadixItems = 4;     // 64-key chunks per wave
constexpr int kRadixTile = kRadixBlock * kRadixItems;
constexpr int kScanBlock = 256;    // zero and row-scan blocks: one thread per digit

// Sorts of at most this many tiles fold the scan into hist and scatter.
constexpr unsigned kFuseTiles = 192;
constexpr int kMaxPasses = 8;

constexpr unsigned radix_tiles(size_t n) { return (unsigned)((n + kRadixTile - 1) / kRadixTile); }

// scratch: the (digit, tile) table, then 256 digit totals per pass
constexpr size_t radix_scratch_words(size_t n) { return (size_t)256 * radix_tiles(n) + 256 * kMaxPasses + 64; }

template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_radix_zero(T *__restrict__ p, unsigned n) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

// FUSED: counts tile-major (tile * 256 + d) and added into totals; else digit-major (d * ntiles
// + tile) for the scan kernel.
template <typename K, bool FUSED>
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const K *__restrict__ keys, size_t n, int shift,
                                                            uint32_t *__restrict__ hist, unsigned ntiles,
                                                            uint32_t *__restrict__ totals) {
    __shared__ uint32_t h[256];
    if (threadIdx.x < 256) h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int i = 0; i < kRadixItems; i++) {
        const size_t idx = base + (size_t)i * kRadixBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const uint32_t c = h[threadIdx.x];
        if (FUSED)
            hist[(size_t)blockIdx.x * 256 + threadIdx.x] = c;
        else
            hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = c;
        if (c) atomicAdd(&totals[threadIdx.x], c);
    }
}

// Inclusive scan of one value per thread over a 256-thread block (x: this thread's value).
__device__ __forceinline__ uint32_t radix_block_scan256(uint32_t x, uint32_t *wsum) {
    const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    for (unsigned j = 0; j < w; j++) x += wsum[j];
    __syncthreads();
    return x;
}

// Digit-major table -> scatter offsets, block d scanning row d (ntiles counts) in place.
template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_radix_scan_rows(T *__restrict__ hist, unsigned ntiles,
                                                                const T *__restrict__ totals) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    const unsigned d = blockIdx.x, tid = threadIdx.x;
    // offset of digit d: the totals of the digits below it (thread 255 ends with their sum)
    const uint32_t lower = radix_block_scan256(tid < d ? totals[tid] : 0u, wsum);
    if (tid == 255) carry = lower;
    __syncthreads();
    uint32_t *row = hist + (size_t)d * ntiles;
    for (unsigned b = 0; b < ntiles; b += kScanBlock) {
        const unsigned i = b + tid;
        const uint32_t v = i < ntiles ? row[i] : 0u;
        const uint32_t inc = radix_block_scan256(v, wsum);
        const uint32_t c = carry;
        if (i < ntiles) row[i] = c + inc - v;
        __syncthreads();
        if (tid == 255) carry = c + inc;
        __syncthreads();
    }
}

// Scatter offset of digit d = threadIdx.x (threads below 256; the others return 0) for tile
// blockIdx.x from the tile-major counts and the digit totals: (exclusive scan of the totals) + (the
// digit's count in the earlier tiles).  The earlier tiles are split over the block's four 256-thread
// quarters, so the chain of dependent L2 reads is a quarter as long.  Called by every thread.
__device__ __forceinline__ uint32_t radix_fused_offset(const uint32_t *__restrict__ hist,
                                                       const uint32_t *__restrict__ totals) {
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t part[kRadixBlock / 256][256];
    const unsigned d = threadIdx.x & 255u, q = threadIdx.x >> 8, lane = d & 63u, w = d >> 6;
    constexpr unsigned kq = kRadixBlock / 256;
    uint32_t v = 0, x = 0;
    if (q == 0) {  // waves 0-3: the scan of the totals
        v = totals[d];
        x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
    }
    const unsigned t = blockIdx.x;
    uint32_t pre = 0;
    unsigned i = q;
    for (; i + 3 * kq < t; i += 4 * kq)
        pre += hist[(size_t)i * 256 + d] + hist[(size_t)(i + kq) * 256 + d] + hist[(size_t)(i + 2 * kq) * 256 + d] +
               hist[(size_t)(i + 3 * kq) * 256 + d];
    for (; i < t; i += kq) pre += hist[(size_t)i * 256 + d];
    part[q][d] = pre;
    __syncthreads();
    if (q != 0) return 0u;
    uint32_t base = x - v;
    for (unsigned j = 0; j < w; j++) base += wsum[j];
#pragma unroll
    for (unsigned j = 0; j < kq; j++) base += part[j][d];
    return base;
}

template <typename K, bool FUSED>
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const K *__restrict__ kin,
                                                               const uint32_t *__restrict__ vin,
                                                               K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                               size_t n, int shift, const uint32_t *__restrict__ offs,
                                                               unsigned ntiles, const uint32_t *__restrict__ totals) {
    __shared__ uint32_t cnt[kRadixWaves][256];
    __shared__ uint32_t base[kRadixWaves][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kRadixWaves * 256; i += kRadixBlock) (&cnt[0][0])[i] = 0;
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
    }
#pragma unroll
    for (int c = 0; c < kRadixItems; c++) {
        const size_t idx = wbase + (size_t)c * 64 + lane;
        const bool ok = idx < n;
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
    // FUSED: the offsets are computed by every thread (barrier inside) and valid below 256
    const uint32_t off0 = FUSED ? radix_fused_offset(offs, totals) : 0u;
    __syncthreads();
    if (threadIdx.x < 256) {
        const uint32_t d = threadIdx.x;
        uint32_t run = FUSED ? off0 : offs[(size_t)d * ntiles + blockIdx.x];
#pragma unroll
        for (int ww = 0; ww < kRadixWaves; ww++) {
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
        const bool fused = nt <= kFuseTiles;
        uint32_t *totals = scratch + (size_t)256 * nt;
        const int passes = (key_bits + 7) / 8;
        if (passes > kMaxPasses) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_radix_zero<uint32_t>, dim3(passes), dim3(kScanBlock), 0, stream, totals,
                           (unsigned)(256 * passes));
        for (int shift = 0, pass = 0; shift < key_bits; shift += 8, pass++) {
            if (fused) {
                uint32_t *tp = totals + 256 * pass;
                hipLaunchKernelGGL((k_radix_hist<K, true>), dim3(nt), dim3(kRadixBlock), 0, stream, ki, n, shift,
                                   scratch, nt, tp);
                hipLaunchKernelGGL((k_radix_scatter<K, true>), dim3(nt), dim3(kRadixBlock), 0, stream, ki, vi, ko,
                                   vo, n, shift, scratch, nt, tp);
            } else {
                uint32_t *tp = totals + 256 * pass;
                hipLaunchKernelGGL((k_radix_hist<K, false>), dim3(nt), dim3(kRadixBlock), 0, stream, ki, n, shift,
                                   scratch, nt, tp);
                hipLaunchKernelGGL(k_radix_scan_rows<uint32_t>, dim3(256), dim3(kScanBlock), 0, stream, scratch, nt,
                                   tp);
                hipLaunchKernelGGL((k_radix_scatter<K, false>), dim3(nt), dim3(kRadixBlock), 0, stream, ki, vi, ko,
                                   vo, n, shift, scratch, nt, (const uint32_t *)nullptr);
            }
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
