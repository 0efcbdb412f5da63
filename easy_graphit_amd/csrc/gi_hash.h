// Counter-based generator spec shared by the device graph generators.
// SplitMix64 output #counter of state seed ^ mix(stream); R-MAT quadrant
// thresholds floor(p * 2^32) for a=0.57, a+b=0.76, a+b+c=0.95
// (runtime/infra_gapbs/generator.h:75-98 draws the same quadrants from
// std::mt19937, which is not reproducible across libraries).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gi {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t hash64(uint64_t seed, uint64_t stream,
                                                    uint64_t counter) {
    return mix64((seed ^ mix64(stream)) + (counter + 1) * 0x9e3779b97f4a7c15ULL);
}

__host__ __device__ __forceinline__ uint32_t uniform_below(uint64_t h, uint32_t n) {
    return (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

// Bijection of [0, 2^scale); the four round constants come from stream 2.
struct IdPermutation {
    uint32_t m[4], c[4], mask;
    int shift;
    int scale;
    __host__ void init(int s, uint64_t seed) {
        scale = s;
        mask = (s >= 32) ? 0xffffffffu : ((1u << s) - 1u);
        shift = (s + 1) / 2;
        for (int r = 0; r < 4; ++r) {
            m[r] = (uint32_t)hash64(seed, 2, 2 * r) | 1u;
            c[r] = (uint32_t)hash64(seed, 2, 2 * r + 1);
        }
    }
    __host__ __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        if (scale <= 0) return 0;
        for (int r = 0; r < 4; ++r) {
            x = (x * m[r] + c[r]) & mask;
            x ^= x >> shift;
        }
        return x;
    }
};

constexpr uint32_t kRmatA = 2448131358u;
constexpr uint32_t kRmatAB = 3264175144u;
constexpr uint32_t kRmatABC = 4080218931u;

__host__ __device__ __forceinline__ void rmat_edge(int scale, uint64_t seed, uint64_t e,
                                                   uint32_t *src, uint32_t *dst) {
    uint64_t words = (uint64_t)((scale + 1) / 2);
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; ++l) {
        if ((l & 1) == 0) h = hash64(seed, 1, e * words + (uint64_t)(l / 2));
        uint32_t r = (l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
        s <<= 1;
        d <<= 1;
        if (r < kRmatAB) {
            if (r >= kRmatA) d++;
        } else {
            s++;
            if (r >= kRmatABC) d++;
        }
    }
    *src = s;
    *dst = d;
}

__host__ __device__ __forceinline__ int32_t edge_weight(uint64_t seed, int32_t u, int32_t v,
                                                        int32_t wmin, int32_t wmax) {
    uint32_t a = (uint32_t)(u < v ? u : v), b = (uint32_t)(u < v ? v : u);
    uint64_t h = hash64(seed, 3, ((uint64_t)a << 32) | b);
    return wmin + (int32_t)uniform_below(h, (uint32_t)(wmax - wmin + 1));
}

}  // namespace gi
