// Device helpers for frontiers held as bitmaps (one 64-bit word per
// wavefront of vertices) with a summary bitmap of their non-empty chunks of
// kBsChunkWords words.
#pragma once

#include "gi_internal.h"

namespace gi {

constexpr int kBsChunkWords = 64;                  // chunk: 64 words = 4096 vertices
constexpr int kBsChunkVerts = kBsChunkWords * 64;
constexpr int kBsMaxSumWords = 256;                // summary words (V up to 2^26)

// One of three pointers without dynamic indexing into a private array
// (which would live in scratch memory).
template <class T>
__device__ __forceinline__ T bs_pick3(T p0, T p1, T p2, int32_t i) {
    return i == 0 ? p0 : (i == 1 ? p1 : p2);
}

__device__ __forceinline__ bool bs_test(const uint64_t *bm, int32_t v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// Marks the chunk holding `word` in the summary (test first: most calls find
// the bit already set).
__device__ __forceinline__ void bs_mark_chunk(uint64_t *sum, int32_t word) {
    const int32_t ch = word / kBsChunkWords;
    const uint64_t bit = 1ull << (ch & 63);
    if (!(sum[ch >> 6] & bit)) atomicOr((unsigned long long *)&sum[ch >> 6], bit);
}

// Exclusive prefix of set bits over the summary words into LDS `pre`;
// returns the number of non-empty chunks. Every thread of the block calls it.
__device__ __forceinline__ int32_t bs_summary_prefix(const uint64_t *sum, int32_t sum_words,
                                                     int32_t *pre) {
    __shared__ int32_t total_s;
    if (threadIdx.x < 64) {
        int32_t carry = 0;
        const int32_t lane = lane_id();
        for (int32_t base = 0; base < sum_words; base += 64) {
            const int32_t j = base + lane;
            const int32_t pc = j < sum_words ? __popcll(sum[j]) : 0;
            int32_t incl = pc;
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            if (j < sum_words) pre[j] = carry + incl - pc;
            carry += __shfl(incl, 63);
        }
        if (lane == 0) total_s = carry;
    }
    __syncthreads();
    return total_s;
}

// Chunk id of the k-th non-empty chunk.
__device__ __forceinline__ int32_t bs_kth_chunk(const uint64_t *sum, const int32_t *pre,
                                                int32_t sum_words, int32_t k) {
    const int32_t j = upper_bound_i32(pre, 0, sum_words, k) - 1;
    uint64_t w = sum[j];
    for (int32_t r = k - pre[j]; r > 0; --r) w &= w - 1;
    return j * 64 + (__ffsll((unsigned long long)w) - 1);
}

// Clears the words of every non-empty chunk of (bm, sum); the summary itself
// is left to the caller. Every thread of the block calls it.
__device__ __forceinline__ void bs_clear_chunks(uint64_t *bm, const uint64_t *sum, int32_t words,
                                                int32_t sum_words, int32_t *pre) {
    const int32_t nz = bs_summary_prefix(sum, sum_words, pre);
    for (int32_t k = blockIdx.x; k < nz; k += gridDim.x) {
        const int32_t ch = bs_kth_chunk(sum, pre, sum_words, k);
        if (threadIdx.x < kBsChunkWords) {
            const int32_t w = ch * kBsChunkWords + (int32_t)threadIdx.x;
            if (w < words) bm[w] = 0;
        }
    }
    __syncthreads();
}

// Block sums of up to three per-thread counters into this block's slots
// (slot arrays of stride `stride`). Every thread of the block calls it.
template <int Block>
__device__ __forceinline__ void bs_flush3(int32_t *slots, int32_t stride, int32_t a, int32_t b,
                                          int32_t c) {
    __shared__ int32_t part[3][Block / 64];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    if (lane_id() == 0) {
        part[0][threadIdx.x >> 6] = a;
        part[1][threadIdx.x >> 6] = b;
        part[2][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int32_t t = 0;
        for (int i = 0; i < Block / 64; ++i) t += part[threadIdx.x][i];
        slots[threadIdx.x * stride + blockIdx.x] = t;
    }
}

}  // namespace gi
