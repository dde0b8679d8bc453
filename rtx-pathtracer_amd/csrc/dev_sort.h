// Block scans and a stable LSD radix sort of (32-bit key, 32-bit value)
// pairs, 6-bit digits: per-block digit histograms, one scan over them, and a
// scatter that ranks equal digits within a wave with one ballot per digit
// bit.  Used by the guiding update (samples by region) and the irradiance
// cache's lookup index (cell items by bucket).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "dev_math.h"

namespace rtx {

constexpr int kSortBlock = 1024;
constexpr int kDigitBits = 6;
constexpr int kDigits = 1 << kDigitBits;

// ------------------------------------------------------------ block scans
template <int BLOCK> RTX_DI uint32_t scanExclusive(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            uint32_t t = sh[w];
            sh[w] = acc;
            acc += t;
        }
        sh[BLOCK / 64] = acc;
    }
    __syncthreads();
    const uint32_t r = sh[wave] + x - v;
    total = sh[BLOCK / 64];
    __syncthreads();
    return r;
}

// One block: exclusive scan of n values in place; the total to *totalOut
// (may be NULL).  (Kernels here are templates so that every file that uses
// them gets its own copy.)
template <int UNUSED>
__global__ __launch_bounds__(1024) void kScanInPlace(uint32_t *a, uint32_t n, uint32_t *totalOut) {
    __shared__ uint32_t sh[1024 / 64 + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = scanExclusive<1024>(v, sh, total);
        if (i < n) a[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && totalOut) *totalOut = carry;
}

// ------------------------------------------------------------ radix sort
template <int UNUSED>
__global__ __launch_bounds__(kSortBlock) void kRadixHist(const uint32_t *keys, const uint32_t *count, int shift,
                                                        uint32_t nblk, uint32_t *hist) {
    __shared__ uint32_t h[kDigits];
    if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kDigits - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kDigits) hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: an item's place among the block's items of its digit is
// the count in earlier waves plus the count in earlier lanes of its wave
// (lanes of a digit matched with one ballot per digit bit).
template <int UNUSED>
__global__ __launch_bounds__(kSortBlock) void kRadixScatter(const uint32_t *keysIn, const uint32_t *valsIn,
                                                           uint32_t *keysOut, uint32_t *valsOut,
                                                           const uint32_t *count, int shift, uint32_t nblk,
                                                           const uint32_t *hist) {
    constexpr int kWaves = kSortBlock / 64;
    __shared__ uint32_t cnt[kWaves][kDigits];
    for (int t = threadIdx.x; t < kWaves * kDigits; t += kSortBlock) (&cnt[0][0])[t] = 0;
    __syncthreads();
    const uint32_t n = *count;
    const uint32_t i = blockIdx.x * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keysIn[i] : 0u;
    const uint32_t d = (key >> shift) & (kDigits - 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint64_t below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    const uint32_t rankInWave = (uint32_t)__popcll(below);
    if (valid && rankInWave == 0) cnt[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x < kDigits) { // exclusive prefix over the waves, per digit
        uint32_t acc = 0;
        for (int w = 0; w < kWaves; ++w) {
            uint32_t t = cnt[w][threadIdx.x];
            cnt[w][threadIdx.x] = acc;
            acc += t;
        }
    }
    __syncthreads();
    if (valid) {
        const uint32_t pos = hist[d * nblk + blockIdx.x] + cnt[wave][d] + rankInWave;
        keysOut[pos] = key;
        valsOut[pos] = valsIn[i];
    }
}

// Sorts the first *count (device) pairs of keysA/valsA by the low `bits`
// bits of the keys, stably, ping-ponging with keysB/valsB; n bounds *count.
// hist holds ceil(n / kSortBlock) * kDigits words.  On return the sorted
// pairs are in keysA/valsA (the pointers are swapped as needed).
inline void launchRadixSortPairs(hipStream_t s, uint32_t *&keysA, uint32_t *&valsA, uint32_t *&keysB,
                                 uint32_t *&valsB, const uint32_t *count, uint32_t n, int bits, uint32_t *hist) {
    const uint32_t nblk = (n + kSortBlock - 1) / kSortBlock;
    for (int shift = 0; shift < bits; shift += kDigitBits) {
        hipLaunchKernelGGL(kRadixHist<0>, dim3(nblk), dim3(kSortBlock), 0, s, keysA, count, shift, nblk, hist);
        hipLaunchKernelGGL(kScanInPlace<0>, dim3(1), dim3(1024), 0, s, hist, nblk * kDigits, (uint32_t *)nullptr);
        hipLaunchKernelGGL(kRadixScatter<0>, dim3(nblk), dim3(kSortBlock), 0, s, keysA, valsA, keysB, valsB, count,
                           shift, nblk, hist);
        std::swap(keysA, keysB);
        std::swap(valsA, valsB);
    }
}

} // namespace rtx
