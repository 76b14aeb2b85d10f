// rt_sort.h — stable LSD radix sort of (uint32 key, uint32 value) pairs on the device, used by the
// scene build (Morton codes) and by the per-vertex normals (vertex ids). Each pass of 8 bits is a
// per-tile histogram -> per-digit row scan -> scatter; the scatter block scans the 256 digit totals
// itself, ranks keys with wave64 ballot/mbcnt multi-split so equal digits keep input order, groups
// its tile by digit in LDS and writes each digit's run of the tile contiguously.
#pragma once

#include <utility>

#include "rt_common.h"

namespace rt {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;                       // items per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__global__ void __launch_bounds__(kSortThreads) k_sort_hist(const uint32_t* __restrict__ keys,
                                                            uint32_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            uint32_t nblk) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Row scan: block d turns row d of the digit-major histogram (counts of digit d in tiles
// 0..nblk-1) into its exclusive prefix sums, in place, and writes the row total to totals[d].
// Thread t takes a contiguous run of the row: its sum, one workgroup scan of the 256 sums, then
// the run's prefixes from its offset.
__global__ void __launch_bounds__(kSortThreads) k_sort_rowscan(uint32_t* __restrict__ hist,
                                                               uint32_t nblk,
                                                               uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kSortThreads];
    uint32_t* row = hist + (uint64_t)blockIdx.x * nblk;
    const uint32_t per = (nblk + kSortThreads - 1) / kSortThreads;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = b + per < nblk ? b + per : nblk;  // run [b, e), empty past the row's end
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += row[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kSortThreads; off <<= 1) {  // inclusive scan of the run sums
        const uint32_t x = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == kSortThreads - 1) totals[blockIdx.x] = s[kSortThreads - 1];
}
// Stable scatter of one tile of kSortTile items. Items are visited in rounds of 256 (item order =
// (round, thread)); each round is ranked by wave64 multi-split (8 ballots give the lanes sharing
// this lane's digit) and written into LDS grouped by digit, so the tile ends up in LDS in
// (digit, input order). The tile is then written out with consecutive threads taking consecutive
// items: each digit's run of the tile lands in one contiguous span of the output.
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals, uint32_t nblk,
    uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];    // output position of this tile's first item of digit d
    __shared__ uint32_t lstart[kRadix];  // LDS position of this tile's first item of digit d
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcount[kWaves][kRadix];
    __shared__ uint32_t sk[kSortTile];
    __shared__ uint32_t sv[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t blk = blockIdx.x;
    // the row scan left exclusive per-tile prefixes: this tile's count of digit tid is the step
    // to the next tile's prefix (to the digit total for the last tile)
    const uint32_t h0 = hist[(uint64_t)tid * nblk + blk];
    const uint32_t cnt = (blk + 1 < nblk ? hist[(uint64_t)tid * nblk + blk + 1] : totals[tid]) - h0;
    run[tid] = totals[tid];
    lstart[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {  // inclusive scans of the totals and tile counts
        const uint32_t x = tid >= off ? run[tid - off] : 0u;
        const uint32_t y = tid >= off ? lstart[tid - off] : 0u;
        __syncthreads();
        run[tid] += x;
        lstart[tid] += y;
        __syncthreads();
    }
    base[tid] = run[tid] - totals[tid] + h0;
    lstart[tid] -= cnt;
    run[tid] = 0;
    for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const uint64_t tile = (uint64_t)blk * kSortTile;
    const uint32_t m = (uint32_t)(n - tile < (uint64_t)kSortTile ? n - tile : (uint64_t)kSortTile);
    for (int r = 0; r < kSortItems; ++r) {
        const uint64_t i = tile + (uint64_t)r * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (key >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t mb = __ballot(bit);
            peers &= bit ? mb : ~mb;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid && lane == leader) wcount[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = lstart[d] + run[d] + rank;
            for (int w = 0; w < wave; ++w) off += wcount[w][d];
            if (off < (uint32_t)kSortTile) {  // always true for a consistent histogram
                sk[off] = key;
                sv[off] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += wcount[w][tid];
            wcount[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
    for (uint32_t j = tid; j < m; j += kSortThreads) {
        const uint32_t key = sk[j];
        const uint32_t d = (key >> shift) & (kRadix - 1);
        const uint32_t off = base[d] + (j - lstart[d]);
        if (off < n) {  // always true for a consistent histogram; keeps writes in bounds
            kout[off] = key;
            vout[off] = sv[j];
        }
    }
}

// Bytes of the histogram scratch for n keys.
inline uint64_t sort_hist_bytes(uint64_t n) {
    const uint64_t nblk = (n + kSortTile - 1) / kSortTile;
    return 4ull * kRadix * (nblk ? nblk : 1);
}

// Sorts n pairs by the low `bits` bits of the keys in ceil(bits / 8) passes, ping-ponging between
// (ki, vi) and (ko, vo); on return ki/vi point at the sorted arrays. hist: sort_hist_bytes(n),
// totals: kRadix words.
inline void sort_pairs(uint32_t*& ki, uint32_t*& vi, uint32_t* ko, uint32_t* vo, uint32_t n, int bits,
                       uint32_t* hist, uint32_t* totals, hipStream_t s) {
    const uint32_t nblk = (uint32_t)((n + kSortTile - 1) / kSortTile);
    for (int shift = 0; shift < bits; shift += kRadixBits) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortThreads), 0, s, ki, n, shift, hist, nblk);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(kRadix), dim3(kSortThreads), 0, s, hist, nblk, totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(kSortThreads), 0, s, ki, vi, n, shift, hist,
                           totals, nblk, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
}

}  // namespace
}  // namespace rt
