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

// Chunk flags: a kernel that makes a chunk non-empty stores 1 into the
// chunk's byte, and one workgroup later rebuilds the summary from the flags
// (bs_build_summary). Marking the summary from every workgroup instead puts
// all their atomics on the summary's few cache lines, where they serialize.
__device__ __forceinline__ void bs_flag_chunk(uint8_t *flags, int32_t word) {
    flags[word / kBsChunkWords] = 1;
}

// Bytes of flag storage for `chunks` chunks (whole 16-byte groups).
__host__ __device__ constexpr int64_t bs_flag_bytes(int64_t chunks) { return (chunks + 15) / 16 * 16; }

// 4 flag bytes (each 0 or 1) -> 4-bit mask.
__device__ __forceinline__ uint32_t bs_byte_mask(uint32_t x) {
    return (x & 1u) | ((x >> 7) & 2u) | ((x >> 14) & 4u) | ((x >> 21) & 8u);
}

// Rewrites every word of summary `sum` from the chunk flags and clears the
// flags; one workgroup of Block threads, every thread calls it. Thread i
// reads the 16 flags of chunks 16i..16i+15 and four neighbouring threads make
// one 64-bit summary word.
template <int Block>
__device__ __forceinline__ void bs_build_summary(uint8_t *flags, int32_t chunks, uint64_t *sum) {
    static_assert(Block % 64 == 0, "whole wavefronts");
    uint4 *f4 = reinterpret_cast<uint4 *>(flags);
    const int32_t n16 = (chunks + 15) / 16;
    for (int32_t base = 0; base < n16; base += Block) {  // uniform
        const int32_t i = base + (int32_t)threadIdx.x;
        uint4 q = make_uint4(0u, 0u, 0u, 0u);
        if (i < n16) q = f4[i];
        const uint32_t m16 = bs_byte_mask(q.x) | (bs_byte_mask(q.y) << 4) | (bs_byte_mask(q.z) << 8) |
                             (bs_byte_mask(q.w) << 12);
        if (m16) f4[i] = make_uint4(0u, 0u, 0u, 0u);
        uint64_t word = (uint64_t)m16 << (16 * (threadIdx.x & 3));
        word |= __shfl_xor(word, 1);
        word |= __shfl_xor(word, 2);
        if ((threadIdx.x & 3) == 0 && i < n16) sum[i >> 2] = word;
    }
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
