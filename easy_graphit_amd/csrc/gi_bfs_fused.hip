// Fused BFS engine: the level loop of apps/bfs.cpp with every schedule
// decision on the device.
//
// The reference's fused mode (fuse_kernel, operators.cpp:651-672) runs the
// whole while-loop inside one cooperative kernel with grid barriers; its
// hybrid schedule decides push vs pull on the host from a D2H copy of the
// frontier size (operators.cpp:526). Here each level is two launches on one
// stream: a persistent step kernel that switches on the chosen mode, and a
// one-workgroup decide kernel (GraphIt's hybrid rule or the MLP selector,
// from device counters). Very wide push vertices (hubs) are listed when they
// are discovered, and the next TWC step spreads their edges over its whole
// grid before it walks its chunks.
//
// The decide stays a launch of its own: run by the last workgroup of a step
// to finish, it would need every workgroup to publish its partial sums with
// a device-scope release, and on this GPU (one L2 per XCD) each such fence
// writes the XCD's L2 back — measured at ~110 us per level on RMAT-24,
// against ~5 us for the extra launch.
//
// Frontiers are bitmaps only (one 64-bit word per wavefront of vertices),
// each with a summary bitmap of its non-empty 64-word chunks:
//   * `visited` holds every vertex claimed so far (and, from the start, the
//     vertices no edge can reach, plus the padding bits past V). A push step
//     tests it with one plain load per edge and claims with one atomicOr, so
//     only edges to still-unclaimed vertices cost an atomic; the winner then
//     sets the vertex in the OUTPUT bitmap;
//   * push steps visit only the non-empty chunks of the input (from its
//     summary), handed to workgroups round-robin, and list their vertices in
//     LDS — no global queue, no shared cursor;
//   * pull steps give each wave 64-vertex words; a vertex first probes one
//     precomputed candidate parent (its widest in-neighbour, stored
//     vertex-major so the probe is a coalesced read), then scans its
//     in-neighbours lane-serially (PULL_THREAD) or cooperatively with the
//     other unresolved vertices of its word (PULL_GROUP);
//   * |F'| and the sum of its degrees are reduced per workgroup into a slot
//     of a partial-sum array (plain stores) that the decide kernel adds up,
//     so no counter is shared by the whole grid.
#include <cstring>
#include <string>
#include <vector>

#include "gi_bfs_fused.h"
#include "gi_bitset.h"
#include "gi_lb.h"

namespace gi {

__device__ __forceinline__ int32_t deg_of(const int32_t *off, int32_t v) {
    return off[v + 1] - off[v];
}

constexpr int kStepBlock = 256;
constexpr int kDecideBlock = 1024;  // one workgroup: few loads of the slot rows per thread
constexpr int kChunkWords = 64;                  // chunk: 64 words = 4096 vertices
constexpr int kChunkVerts = kChunkWords * 64;
constexpr int kCandScan = 128;                   // in-neighbours examined per candidate

struct StepArgs {
    BfsState *st;
    const int32_t *off;         // out-edges (push)
    const int32_t *nbr;
    const int32_t *poff;        // in-edges (pull); the same arrays on a symmetric graph
    const int32_t *pnbr;
    const int32_t *cand;        // candidate parent per vertex, -1 when none
    int32_t cand_first;         // 1: the candidate is the vertex's first in-neighbour
    int32_t V;
    int32_t words;
    int32_t chunks;
    int32_t sum_words;
    int32_t *depth;
    int32_t *parent;
    uint64_t *visited;
    uint64_t *bm[3];
    uint64_t *sum[3];
    uint8_t *cflag;             // chunk flags of the output bitmap (bs_flag_chunk)
    int32_t *pcount;            // [kMaxGrid]: one slot per step workgroup
    unsigned long long *pfdeg;  // same layout
};

struct StepCtx {
    int32_t level;
    int32_t lane_max;          // degree < lane_max: one lane walks the vertex
    int32_t hub_min;           // TWC: degree >= hub_min goes to the grid bin
    int32_t want_fdeg;
    int32_t want_hubs;         // list the hubs of F' as they are discovered
    int32_t hub_out;           // their list: hubs[(level + 1) & 1]
    int32_t hubs_listed;       // TWC step whose hubs (of F) are all listed: walked grid-wide first
    const uint64_t *bm_in;
    uint64_t *bm_out;
    uint8_t *cflag;
    int32_t cnt;               // per-lane |F'| contribution
    unsigned long long fdeg;   // per-lane sum of degrees of F'
};

// Index one of three pointers without dynamic indexing into a private array
// (which would live in scratch memory).
template <class T>
__device__ __forceinline__ T pick3(T p0, T p1, T p2, int32_t i) {
    return i == 0 ? p0 : (i == 1 ? p1 : p2);
}

template <class T>
__device__ __forceinline__ T pick4(const T *x, int32_t i) {
    return i == 0 ? x[0] : (i == 1 ? x[1] : (i == 2 ? x[2] : x[3]));
}

__device__ __forceinline__ bool in_bitmap(const uint64_t *bm, int32_t v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// A discovered vertex's degree, read when the frontier degree sum or the hub
// list needs it.
__device__ __forceinline__ int32_t note_degree(const StepArgs &a, StepCtx &c, int32_t d) {
    if (!c.want_fdeg && !c.want_hubs) return 0;
    const int32_t dd = deg_of(a.off, d);
    if (c.want_fdeg) c.fdeg += (unsigned long long)dd;
    return dd;
}

// Appends d to the hub list of F' (one lane).
__device__ __forceinline__ void note_hub(const StepArgs &a, const StepCtx &c, int32_t d) {
    const int32_t p = atomicAdd(&a.st->hub_n[c.hub_out], 1);
    if (p < kMaxHubs) a.st->hubs[c.hub_out][p] = d;
}

// Appends the lanes' hubs to the list of F' with one atomic per wave (every
// lane of the wave calls it).
__device__ __forceinline__ void note_hubs_wave(const StepArgs &a, const StepCtx &c, int32_t d, bool hub) {
    const int32_t p = wave_atomic_add(&a.st->hub_n[c.hub_out], hub);
    if (hub && p < kMaxHubs) a.st->hubs[c.hub_out][p] = d;
}

// Push discovery: test, claim in `visited`, record in the output bitmap.
__device__ __forceinline__ void push_edge(const StepArgs &a, StepCtx &c, int32_t s, int32_t e,
                                          bool active) {
    if (!active) return;
    int32_t d = a.nbr[e];
    uint64_t bit = 1ull << (d & 63);
    uint64_t *vw = a.visited + (d >> 6);
    if (*vw & bit) return;  // a stale word only costs the atomic below
    unsigned long long old = atomicOr((unsigned long long *)vw, bit);
    if (old & bit) return;
    unsigned long long o2 = atomicOr((unsigned long long *)&c.bm_out[d >> 6], bit);
    if (o2 == 0) bs_flag_chunk(c.cflag, d >> 6);
    if (a.parent) a.parent[d] = s;  // the depth is written in word order by the next step
    c.cnt += 1;
    const int32_t dd = note_degree(a, c, d);
    if (c.want_hubs && dd >= c.hub_min) note_hub(a, c, d);
}

// Push discovery by a whole wavefront at once (every lane calls it; `active`
// lanes carry an edge). Lanes whose targets share a bitmap word — runs of
// neighbours of a sorted list, common among hubs in the degree-ordered
// layout — combine their bits with a segmented OR, and the last lane of each
// run claims them with one atomicOr, then tells the run which bits it won.
__device__ __forceinline__ void push_edge_wave(const StepArgs &a, StepCtx &c, int32_t s, int32_t e,
                                               bool active) {
    const int32_t lane = lane_id();
    const int32_t d = active ? a.nbr[e] : -1;
    const int32_t key = d >= 0 ? (d >> 6) : -1;
    const uint64_t bit = d >= 0 ? 1ull << (d & 63) : 0ull;
    const bool want = d >= 0 && !(a.visited[key] & bit);  // a stale word only costs the atomic
    // runs: maximal groups of adjacent lanes with the same word (the same
    // word may come back in a later run: that run claims on its own)
    const int32_t kp = __shfl_up(key, 1), kn = __shfl_down(key, 1);
    const bool head = lane == 0 || kp != key;
    const bool end = lane == 63 || kn != key;
    const uint64_t heads = __ballot(head);
    const int32_t my_head = 63 - __clzll((unsigned long long)(heads & ((2ull << lane) - 1)));
    uint64_t x = want ? bit : 0ull;
    for (int o = 1; o < 64; o <<= 1) {  // segmented OR inside the lane's own run
        const uint64_t y = __shfl_up(x, o);
        if (lane - o >= my_head) x |= y;
    }
    uint64_t won = 0;
    if (end && key >= 0 && x) {
        const unsigned long long old = atomicOr((unsigned long long *)&a.visited[key], x);
        won = x & ~old;
        if (won) {
            const unsigned long long o2 = atomicOr((unsigned long long *)&c.bm_out[key], won);
            if (o2 == 0) bs_flag_chunk(c.cflag, key);
        }
    }
    const uint64_t ends = __ballot(end);
    const int32_t leader = __ffsll((unsigned long long)(ends & (~0ull << lane))) - 1;
    const uint64_t run_won = __shfl(won, leader);
    // bits of the lanes before this one in its run: a target repeated in a
    // run is recorded once, by its first lane
    const uint64_t xp = __shfl_up(x, 1);
    const uint64_t before = lane - 1 >= my_head ? xp : 0ull;
    const bool mine = want && (run_won & bit) && !(before & bit);
    int32_t dd = 0;
    if (mine) {
        if (a.parent) a.parent[d] = s;  // the depth is written in word order by the next step
        c.cnt += 1;
        dd = note_degree(a, c, d);
    }
    if (c.want_hubs) note_hubs_wave(a, c, d, mine && dd >= c.hub_min);  // uniform
}

__device__ __forceinline__ void lane_walk(const StepArgs &a, StepCtx &c, int32_t s, int32_t e,
                                          int32_t e1) {
    for (; e < e1; ++e) push_edge(a, c, s, e, true);
}

__device__ __forceinline__ void wave_walk(const StepArgs &a, StepCtx &c, int32_t s, int32_t e0,
                                          int32_t e1) {
    int32_t lane = lane_id();
    for (int32_t b = e0; b < e1; b += 64) push_edge_wave(a, c, s, b + lane, b + lane < e1);
}

// Workgroup reduction of the lane counters into this block's slot.
__device__ __forceinline__ void flush_counts(int32_t *pcount, unsigned long long *pfdeg,
                                             StepCtx &c) {
    __shared__ int32_t wc[kStepBlock / 64];
    __shared__ unsigned long long wf[kStepBlock / 64];
    int32_t cnt = c.cnt;
    unsigned long long fd = c.fdeg;
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o);
        fd += __shfl_xor(fd, o);
    }
    if (lane_id() == 0) {
        wc[threadIdx.x >> 6] = cnt;
        wf[threadIdx.x >> 6] = fd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tc = 0;
        unsigned long long tf = 0;
        for (int i = 0; i < kStepBlock / 64; ++i) {
            tc += wc[i];
            tf += wf[i];
        }
        pcount[blockIdx.x] = tc;
        pfdeg[blockIdx.x] = tf;
    }
}

// Prefix of set bits over the summary words (LDS), for k-th-chunk lookup.
// Returns the number of non-empty chunks. Every thread of the block calls it.
__device__ int32_t summary_prefix(const uint64_t *sum, int32_t sum_words, int32_t *pre) {
    __shared__ int32_t total_s;
    // one wave scans 64 words at a time
    if (threadIdx.x < 64) {
        int32_t carry = 0;
        int32_t lane = lane_id();
        for (int32_t base = 0; base < sum_words; base += 64) {
            int32_t j = base + lane;
            int32_t pc = j < sum_words ? __popcll(sum[j]) : 0;
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
__device__ __forceinline__ int32_t kth_chunk(const uint64_t *sum, const int32_t *pre,
                                             int32_t sum_words, int32_t k) {
    int32_t j = upper_bound_i32(pre, 0, sum_words, k) - 1;
    uint64_t w = sum[j];
    for (int32_t r = k - pre[j]; r > 0; --r) w &= w - 1;
    return j * 64 + (__ffsll((unsigned long long)w) - 1);
}

// One batch (<= kStepBlock vertices of the block's LDS list) in push mode.
__device__ void push_batch(const StepArgs &a, StepCtx &c, int32_t mode, const int32_t *list,
                           int32_t nb) {
    __shared__ int32_t soff[kStepBlock], sscan[kStepBlock + 1];
    const int32_t tid = threadIdx.x, wid = tid >> 6;
    const bool has = tid < nb;
    const int32_t s = has ? list[tid] : 0;
    const int32_t e0 = has ? a.off[s] : 0, e1 = has ? a.off[s + 1] : 0;
    const int32_t dg = e1 - e0;
    if (mode == BFS_PUSH_THREAD) {
        lane_walk(a, c, s, e0, e1);
    } else if (mode == BFS_PUSH_WAVE) {
        for (int32_t j = wid; j < nb; j += kStepBlock / 64) {
            int32_t v = list[j];
            wave_walk(a, c, v, a.off[v], a.off[v + 1]);
        }
    } else if (mode == BFS_PUSH_TWC) {
        // lane bin, wave bin; the grid bin (listed hubs) was walked at the
        // start of the step, unless the list overflowed
        const bool todo = has && !(c.hubs_listed && dg >= c.hub_min);
        bool wide = todo && dg >= c.lane_max;
        uint64_t wm = __ballot(wide);
        while (wm) {
            int32_t l = __ffsll((unsigned long long)wm) - 1;
            wm &= wm - 1;
            int32_t ws = __shfl(s, l), w0 = __shfl(e0, l), w1 = __shfl(e1, l);
            wave_walk(a, c, ws, w0, w1);
        }
        if (todo && dg < c.lane_max) lane_walk(a, c, s, e0, e1);
    } else {
        // TWCE and CM need the workgroup's degree scan
        int32_t total = 0;
        int32_t ex = block_exclusive_scan(dg, &total);
        sscan[tid] = ex;
        soff[tid] = e0;
        __syncthreads();
        if (mode == BFS_PUSH_CM) {
            for (int32_t k = tid; k - tid < total; k += kStepBlock) {
                bool act = k < total;
                int32_t v = 0, e = 0;
                if (act) {
                    int32_t j = upper_bound_i32(sscan, 0, nb, k) - 1;
                    v = list[j];
                    e = soff[j] + (k - sscan[j]);
                }
                push_edge_wave(a, c, v, e, act);
            }
        } else {
            // TWCE: workgroup granule (degree >= block), wave granule (>= 64), lanes
            for (int32_t j = 0; j < nb; ++j) {
                int32_t dj = (j + 1 < nb ? sscan[j + 1] : total) - sscan[j];
                if (dj < kStepBlock) continue;  // uniform
                int32_t v = list[j], o = soff[j];
                for (int32_t b = 0; b < dj; b += kStepBlock) push_edge_wave(a, c, v, o + b + tid, b + tid < dj);
            }
            for (int32_t j = wid; j < nb; j += kStepBlock / 64) {
                int32_t dj = (j + 1 < nb ? sscan[j + 1] : total) - sscan[j];
                if (dj < 64 || dj >= kStepBlock) continue;  // wave-uniform
                wave_walk(a, c, list[j], soff[j], soff[j] + dj);
            }
            if (has && dg < 64) lane_walk(a, c, s, e0, e1);
        }
        __syncthreads();
    }
}

// 64-bit OR over the wave.
__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    return x;
}

constexpr int kPullWords = 4;     // words a wave works on at once (memory-level parallelism; pick4)
constexpr int kPullUnit = kPullWords * (kStepBlock / 64);  // words per workgroup unit
static_assert(kChunkWords % kPullUnit == 0, "a pull unit lies inside one chunk");
constexpr int kPullLaneSteps = 4; // PULL_GROUP: lane-serial edges before the cooperative scan (4: RMAT-24 group levels 120 -> 115 us against 8; 2, 3, 5, 6, 16, 32 no better)

// Cooperative scan of the open vertices of one word: their remaining in-edge
// lists [e, e1) concatenated, 64 edges per pass, until every open vertex has
// a parent in the frontier or its list is exhausted. Returns the resolved
// lanes.
__device__ __forceinline__ uint64_t pull_coop(const StepArgs &a, const StepCtx &c, int32_t w,
                                              int32_t e, int32_t e1) {
    const int32_t lane = lane_id();
    const int32_t len = e1 > e ? e1 - e : 0;
    int32_t incl = len;
    for (int o = 1; o < 64; o <<= 1) {
        int32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int32_t total = __shfl(incl, 63);
    const int32_t pre = incl - len;
    const uint64_t open_mask = __ballot(len > 0);
    uint64_t res = 0;
    for (int32_t base = 0; base < total; base += 64) {
        const int32_t g = base + lane;
        // owner: the last lane whose exclusive prefix is <= g
        int32_t j = 0;
        for (int32_t step = 32; step > 0; step >>= 1) {
            int32_t t = j + step;
            int32_t pt = __shfl(pre, t & 63);
            if (t < 64 && pt <= g) j = t;
        }
        const int32_t ej = __shfl(e, j), pj = __shfl(pre, j);
        bool hit = false;
        int32_t u = 0;
        if (g < total && !((res >> j) & 1ull)) {
            u = a.pnbr[ej + (g - pj)];
            hit = in_bitmap(c.bm_in, u);
        }
        if (__ballot(hit)) {
            if (hit && a.parent) a.parent[w * 64 + j] = u;  // any hit is a valid parent
            res |= wave_or(hit ? (1ull << j) : 0ull);
            if ((res | ~open_mask) == ~0ull) break;  // every open vertex resolved
        }
    }
    return res;
}

// Pull over words [w0, w0 + kPullWords) by one wave; lane l owns vertex
// 64 w + l of each. The kPullWords vertices of a lane advance together so
// their loads are in flight at the same time.
__device__ __forceinline__ bool pull_words(const StepArgs &a, StepCtx &c, int32_t mode, int32_t w0,
                                           const uint64_t *vis) {
    const int32_t lane = lane_id();
    int32_t par[kPullWords], e[kPullWords], e1[kPullWords], cd[kPullWords];
    // candidate and in-edge range of every lane's vertex, loaded without
    // waiting for `visited` so the loads overlap (past V: a clamped index,
    // never used since padding bits are set)
#pragma unroll
    for (int i = 0; i < kPullWords; ++i) {
        int32_t v = (w0 + i) * 64 + lane;
        v = v < a.V ? v : a.V - 1;
        cd[i] = a.cand[v];
        e[i] = a.poff[v];
        e1[i] = a.poff[v + 1];
    }
#pragma unroll
    for (int i = 0; i < kPullWords; ++i) {
        const bool want = !((vis[i] >> lane) & 1ull);  // padding bits past V are set
        par[i] = (want && cd[i] >= 0 && in_bitmap(c.bm_in, cd[i])) ? cd[i] : -1;
        if (!want || par[i] >= 0) e1[i] = e[i];  // closed: nothing to scan
        else if (a.cand_first && cd[i] >= 0) ++e[i];  // the candidate is the first in-neighbour: probed
    }
    bool any = false;
    // lane-serial scan, the lane's vertices interleaved
    const int32_t steps = mode == BFS_PULL_THREAD ? 0x7fffffff : kPullLaneSteps;
    for (int32_t it = 0; it < steps; ++it) {
        bool more = false;
#pragma unroll
        for (int i = 0; i < kPullWords; ++i) more |= e[i] < e1[i];
        if (!__ballot(more)) break;
        int32_t u[kPullWords];
#pragma unroll
        for (int i = 0; i < kPullWords; ++i) u[i] = e[i] < e1[i] ? a.pnbr[e[i]] : -1;
#pragma unroll
        for (int i = 0; i < kPullWords; ++i) {
            if (u[i] >= 0) {
                if (in_bitmap(c.bm_in, u[i])) {
                    par[i] = u[i];
                    e1[i] = e[i];
                } else {
                    ++e[i];
                }
            }
        }
    }
    // one word at a time from here (not unrolled: the cooperative scan is
    // inlined once); the arrays are read through selects, not indexing
#pragma unroll 1
    for (int i = 0; i < kPullWords; ++i) {
        const int32_t w = w0 + i;
        const int32_t ei = pick4(e, i), e1i = pick4(e1, i), pi = pick4(par, i);
        uint64_t coop = 0;
        if (mode != BFS_PULL_THREAD && __ballot(ei < e1i)) coop = pull_coop(a, c, w, ei, e1i);
        const int32_t v = w * 64 + lane;
        const bool found_self = pi >= 0 || ((coop >> lane) & 1ull);
        if (found_self) {  // (pull steps list no hubs)
            if (a.parent && pi >= 0) a.parent[v] = pi;
            if (c.want_fdeg) c.fdeg += (unsigned long long)deg_of(a.off, v);
        }
        const uint64_t found = __ballot(found_self);
        if (found) {
            any = true;
            if (lane == 0) {
                a.visited[w] = pick4(vis, i) | found;  // this wave owns word w in a pull step
                c.bm_out[w] = found;
                c.cnt += __popcll(found);
            }
        }
    }
    return any;
}

// Grid bin of PUSH_TWC, at the start of the step: the edges of F's listed
// hubs (degree >= hub_min, listed when they were discovered) spread evenly
// over the grid. Every workgroup scans the hub degrees into LDS (`lds`: the
// step's vertex list, free at this point) and walks its share of the
// concatenated edge range. Every thread of the workgroup calls it.
static_assert(kChunkVerts >= 2 * kMaxHubs + 1, "hub scan fits the vertex list");
constexpr int32_t kHubListShare = 64;  // list hubs during push steps of <= V/64 vertices, <= E/64 edges
__device__ __forceinline__ void hub_phase(const StepArgs &a, StepCtx &c, int32_t nh, int32_t *lds) {
    const int32_t *hubs = a.st->hubs[c.level & 1];
    int32_t *hv = lds, *hscan = lds + kMaxHubs;
    const int32_t per = (nh + kStepBlock - 1) / kStepBlock;
    const int32_t lo = threadIdx.x * per, hi = lo + per < nh ? lo + per : nh;
    int32_t part = 0;
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = hubs[i];
        hv[i] = s;
        part += deg_of(a.off, s);
    }
    int32_t total = 0;
    int32_t ex = block_exclusive_scan(part, &total);
    for (int32_t i = lo; i < hi; ++i) {
        hscan[i] = ex;
        ex += deg_of(a.off, hv[i]);
    }
    __syncthreads();
    const int32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const int32_t k0 = blockIdx.x * chunk, k1 = k0 + chunk < total ? k0 + chunk : total;
    for (int32_t kb = k0; kb < k1; kb += kStepBlock) {  // block-uniform trip count
        const int32_t k = kb + (int32_t)threadIdx.x;
        const bool act = k < k1;
        int32_t s = 0, e = 0;
        if (act) {
            const int32_t j = upper_bound_i32(hscan, 0, nh, k) - 1;
            s = hv[j];
            e = a.off[s] + (k - hscan[j]);
        }
        push_edge_wave(a, c, s, e, act);
    }
    __syncthreads();  // the list LDS is reused by the chunk loop
}

// ---------------------------------------------------------------- step
// One level's step: the push or the pull body, as the state's mode says.
// `pre` is LDS of kMaxSumWords + 1 words owned by the calling kernel, so a
// kernel that holds both step bodies allocates it once.
enum StepKind : int { kPushStep = 0, kPullStep = 1 };

template <int Kind>
__device__ __forceinline__ void step_body(const StepArgs &a0, int32_t *pre) {
    BfsState *st = a0.st;
    StepArgs a = a0;
    a.depth = st->depth;
    a.parent = st->parent;
    StepCtx c;
    c.level = st->level;
    c.lane_max = st->lane_max;
    c.hub_min = st->hub_min;
    c.want_fdeg = st->want_fdeg;
    const int32_t mode = st->mode;
    // F''s hubs are listed during small push steps only: after a large or a
    // pull step the next step is a pull step or has too many hubs for the
    // list (and the same-address atomics of listing them would cost more
    // than the list saves); such a step voids the list instead
    {
        const int32_t p = c.level & 1;
        c.want_hubs = st->want_hubs && mode < BFS_PULL_THREAD && st->count[p] <= st->hub_list_count &&
                      (!c.want_fdeg || (long long)st->fdeg[p] <= st->hub_list_fdeg);
    }
    c.hub_out = (c.level + 1) & 1;
    if (!c.want_hubs && blockIdx.x == 0 && threadIdx.x == 0) st->hub_n[c.hub_out] = kMaxHubs + 1;
    const int32_t nh = st->hub_n[c.level & 1];
    c.hubs_listed = mode == BFS_PUSH_TWC && st->want_hubs && nh <= kMaxHubs;  // else void or never kept
    const int32_t iin = c.level % 3, iout = (c.level + 1) % 3, iclr = (c.level + 2) % 3;
    c.bm_in = pick3(a0.bm[0], a0.bm[1], a0.bm[2], iin);
    c.bm_out = pick3(a0.bm[0], a0.bm[1], a0.bm[2], iout);
    c.cflag = a0.cflag;
    c.cnt = 0;
    c.fdeg = 0;
    const int32_t lane = lane_id();

    // the bitmap two levels ahead is free: clear its non-empty chunks (its
    // summary is cleared by the decide after this step)
    {
        const uint64_t *sc = pick3(a0.sum[0], a0.sum[1], a0.sum[2], iclr);
        uint64_t *bc = pick3(a0.bm[0], a0.bm[1], a0.bm[2], iclr);
        int32_t nz = summary_prefix(sc, a.sum_words, pre);
        for (int32_t k = blockIdx.x; k < nz; k += gridDim.x) {
            int32_t ch = kth_chunk(sc, pre, a.sum_words, k);
            if (threadIdx.x < kChunkWords) {
                int32_t w = ch * kChunkWords + (int32_t)threadIdx.x;
                if (w < a.words) bc[w] = 0;
            }
        }
        __syncthreads();
    }

    if constexpr (Kind == kPushStep) {
        __shared__ int32_t list[kChunkVerts];
        __shared__ int32_t n_s;
        if (c.hubs_listed && nh > 0) hub_phase(a, c, nh, list);  // uniform
        const uint64_t *si = pick3(a0.sum[0], a0.sum[1], a0.sum[2], iin);
        int32_t nz = summary_prefix(si, a.sum_words, pre);
        for (int32_t k = blockIdx.x; k < nz; k += gridDim.x) {
            int32_t chunk = kth_chunk(si, pre, a.sum_words, k);
            if (threadIdx.x < 64) {
                // wave 0 lists the chunk's members
                int32_t w = chunk * kChunkWords + (int32_t)threadIdx.x;
                uint64_t word = w < a.words ? c.bm_in[w] : 0ull;
                int32_t pc = __popcll(word);
                int32_t incl = pc;
                for (int o = 1; o < 64; o <<= 1) {
                    int32_t y = __shfl_up(incl, o);
                    if (lane >= o) incl += y;
                }
                int32_t pos = incl - pc;
                while (word) {
                    int32_t b = __ffsll((unsigned long long)word) - 1;
                    word &= word - 1;
                    list[pos++] = w * 64 + b;
                    a.depth[w * 64 + b] = c.level;  // members of F_level get their depth here
                }
                if (threadIdx.x == 63) n_s = incl;
            }
            __syncthreads();
            const int32_t n = n_s;
            for (int32_t b0 = 0; b0 < n; b0 += kStepBlock)
                push_batch(a, c, mode, list + b0, n - b0 < kStepBlock ? n - b0 : kStepBlock);
            __syncthreads();
        }
    } else {
        // pull over all vertices in units of kPullUnit words, one unit per
        // workgroup at a time (each wave takes kPullWords of them), so a
        // unit's summary bit costs at most one atomic
        const int32_t units = (a.words + kPullUnit - 1) / kPullUnit;
        const int32_t wid = threadIdx.x >> 6;
        for (int32_t un = blockIdx.x; un < units; un += gridDim.x) {
            const int32_t w0 = un * kPullUnit + wid * kPullWords;
            uint64_t vis[kPullWords];
            uint64_t all = ~0ull;
#pragma unroll
            for (int i = 0; i < kPullWords; ++i) {
                const bool in_range = w0 + i < a.words;
                vis[i] = in_range ? a.visited[w0 + i] : ~0ull;
                all &= vis[i];
                // members of F_level get their depth here
                const uint64_t fin = in_range ? c.bm_in[w0 + i] : 0ull;
                if ((fin >> lane) & 1ull) a.depth[(w0 + i) * 64 + lane] = c.level;
            }
            bool any = false;
            if (all != ~0ull) any = pull_words(a, c, mode, w0, vis);  // wave-uniform
            if (__syncthreads_or(any) && threadIdx.x == 0) bs_flag_chunk(c.cflag, un * kPullUnit);
        }
    }
    flush_counts(a.pcount, a.pfdeg, c);
}

// ---------------------------------------------------------------- decide
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}

// The decide kernel: one workgroup, after step `level` (or after seeding,
// level = -1), adds up the step's partial sums, closes its bookkeeping, ends
// the run on an empty frontier, else picks the next step's mode and clears
// what the next step writes. The MLP runs on wave 0; every thread passes the
// same barriers.
struct DecideArgs {
    BfsState *st;
    int32_t V;
    int64_t E;
    int32_t *pcount;
    unsigned long long *pfdeg;
    int32_t slots;          // step workgroups that wrote a slot
    uint64_t *sum[3];
    int32_t sum_words;
    uint8_t *cflag;
    int32_t chunks;
    int32_t *host_flag;     // pinned: {done, level, levels} when the run ends
    const int32_t *off;
    uint64_t *visited;
    uint64_t *bm0;          // the bitmap of F_0
    int32_t seed;           // the decide after the reset: a single source to seed, or -1
};

// The selector's 16 inputs in FP32 with the hardware log2 (v_log_f32): the
// values of bfs_raw_feature (double, used by the host to record training
// data) to within FP32 rounding, at a few dozen instructions instead of the
// hundreds of FP64 ones its logarithms and divisions take on one wavefront.
__device__ __forceinline__ void bfs_features_f32(float *f, int32_t n, unsigned long long fdeg, long long unv,
                                                 long long unv_e, int32_t level, int32_t prev_mode,
                                                 int32_t prev_n, int32_t V, int64_t E) {
    const float dV = V > 0 ? (float)V : 1.f, dE = E > 0 ? (float)E : 1.f;
    const float nn = (float)n, fd = (float)fdeg, u = (float)unv, ue = (float)unv_e;
    const float l_n = __log2f(1.f + nn);
    f[0] = nn / dV;
    f[1] = fd / dE;
    f[2] = u / dV;
    f[3] = ue / dE;
    f[4] = l_n / 32.f;
    f[5] = __log2f(1.f + fd) / 40.f;
    f[6] = __log2f(1.f + fd / (n > 0 ? nn : 1.f)) / 20.f;
    f[7] = __log2f(1.f + dE / dV) / 20.f;
    f[8] = __log2f(dV) / 32.f;
    f[9] = level < 64 ? (float)level / 64.f : 1.f;
    f[10] = prev_mode >= BFS_PULL_THREAD ? 1.f : 0.f;
    f[11] = (l_n - __log2f(1.f + (float)prev_n)) / 16.f;
    f[12] = fd / (1.f + ue);
    f[13] = nn / (1.f + u);
    f[14] = __log2f(1.f + u) / 32.f;
    f[15] = 1.f;
}

// BF16 weight k (unrolled loops make k a constant) of a row held in registers.
__device__ __forceinline__ float bf16_in(const uint4 *q, int k) {
    const uint4 v = q[k >> 3];
    const int j = (k >> 1) & 3;
    const uint32_t word = j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    return __uint_as_float((k & 1) ? (word & 0xffff0000u) : (word << 16));
}

template <int Block>
__device__ __forceinline__ void decide_body(const DecideArgs &d, const SelectorView &sel) {
    BfsState *st = d.st;
    if (st->done) return;  // uniform
    const int32_t V = d.V;
    const int64_t E = d.E;
    int32_t *pcount = d.pcount;
    unsigned long long *pfdeg = d.pfdeg;
    int32_t *host_flag = d.host_flag;
    // after step `level`, the bitmap two levels ahead of it was cleared by the
    // step; its summary is cleared here
    uint64_t *sum_clear = pick3(d.sum[0], d.sum[1], d.sum[2], (st->level + 2 + 3) % 3);
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    const int32_t next_level = st->level + 1;
    const int32_t p = next_level & 1;
    const bool learned = sel.mode == GI_SELECT_LEARNED && sel.w1 != nullptr;
    // the network runs on wave 0, one lane per hidden unit; each lane's
    // weights are loaded first, so their latency hides behind the partial sums
    static_assert(kSelMaxOut == 8, "one 16-byte row of output weights per hidden unit");
    uint4 w1q[kSelInputs / 8], w2q[kSelHidden / 8], w3q[1];
    uint16_t b1v = 0, b2v = 0, b3v = 0;
#pragma unroll
    for (int i = 0; i < kSelInputs / 8; ++i) w1q[i] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int i = 0; i < kSelHidden / 8; ++i) w2q[i] = make_uint4(0u, 0u, 0u, 0u);
    w3q[0] = make_uint4(0u, 0u, 0u, 0u);
    if (learned && wid == 0) {
        const uint4 *r1 = reinterpret_cast<const uint4 *>(sel.w1 + lane * kSelInputs);
        const uint4 *r2 = reinterpret_cast<const uint4 *>(sel.w2 + lane * kSelHidden);
#pragma unroll
        for (int i = 0; i < kSelInputs / 8; ++i) w1q[i] = r1[i];
#pragma unroll
        for (int i = 0; i < kSelHidden / 8; ++i) w2q[i] = r2[i];
        w3q[0] = *reinterpret_cast<const uint4 *>(sel.w3 + lane * kSelMaxOut);
        b1v = sel.b1[lane];
        b2v = sel.b2[lane];
        if (lane < sel.n_out) b3v = sel.b3[lane];
    }
    __shared__ int32_t rn[Block / 64];
    __shared__ unsigned long long rf[Block / 64];
    int32_t n = 0;
    unsigned long long fdeg = 0;
    if (next_level == 0 && d.seed >= 0) {
        // a single source: seeded here (the reset cleared its entries); every
        // thread counts it
        const int32_t v = d.seed, dv = d.off[v + 1] - d.off[v];
        n = 1;
        fdeg = (unsigned long long)dv;
        if (threadIdx.x == 0) {
            const uint64_t bit = 1ull << (v & 63);
            const int32_t ch = (v >> 6) / kChunkWords;
            st->depth[v] = 0;
            if (st->parent) st->parent[v] = v;
            d.visited[v >> 6] |= bit;
            d.bm0[v >> 6] |= bit;
            d.sum[0][ch >> 6] |= 1ull << (ch & 63);
            if (st->want_hubs && dv >= st->hub_min) {  // the first step's grid bin
                st->hubs[0][0] = v;
                st->hub_n[0] = 1;
            }
        }
    } else if (next_level == 0) {  // seeded by the seed kernel
        n = st->count[0];
        fdeg = st->fdeg[0];
    } else {
        for (int32_t i = threadIdx.x; i < d.slots; i += Block) {
            n += pcount[i];
            fdeg += pfdeg[i];
            pcount[i] = 0;
            pfdeg[i] = 0;
        }
        for (int o = 32; o > 0; o >>= 1) {
            n += __shfl_xor(n, o);
            fdeg += __shfl_xor(fdeg, o);
        }
        if (lane == 0) {
            rn[wid] = n;
            rf[wid] = fdeg;
        }
        for (int32_t i = threadIdx.x; i < d.sum_words; i += Block) sum_clear[i] = 0;
        // the summary of F', from the flags its chunks set
        bs_build_summary<Block>(d.cflag, d.chunks, pick3(d.sum[0], d.sum[1], d.sum[2], next_level % 3));
        __syncthreads();
        n = 0;
        fdeg = 0;
        for (int i = 0; i < Block / 64; ++i) {
            n += rn[i];
            fdeg += rf[i];
        }
    }
    const int32_t prev_n = next_level > 0 ? st->count[p ^ 1] : 0;
    const long long unv = st->unvisited - n;
    const long long unv_e = st->unvisited_edges - (long long)fdeg;
    if (n == 0) {  // uniform
        if (threadIdx.x == 0) {
            st->done = 1;
            st->levels = next_level;
            if (host_flag) {  // the host waits on the replay's event, then reads these
                host_flag[2] = next_level;
                host_flag[1] = next_level;
                host_flag[0] = 1;
            }
        }
        return;
    }
    int32_t mode;
    if (!learned) {
        if (sel.mode == GI_SELECT_S1) mode = sel.mode_s1;
        else if (sel.mode == GI_SELECT_S2) mode = sel.mode_s2;
        else mode = ((double)n <= (double)sel.threshold * (double)V) ? sel.mode_s1 : sel.mode_s2;
    } else {
        if (wid != 0) return;  // the state is written by thread 0
        // every lane computes the whole input vector (the same values, so no
        // lane waits on another's branch of the feature switch)
        float feats[kSelInputs];
        bfs_features_f32(feats, n, fdeg, unv, unv_e, next_level, st->mode, prev_n, V, E);
        float a1 = bf16_to_f32(b1v);
#pragma unroll
        for (int k = 0; k < kSelInputs; ++k)
            a1 += (feats[k] - sel.feat_mean[k]) * sel.feat_scale[k] * bf16_in(w1q, k);
        const float h1 = a1 > 0.f ? a1 : 0.f;
        float a2 = bf16_to_f32(b2v);
#pragma unroll
        for (int k = 0; k < kSelHidden; ++k) a2 += __shfl(h1, k) * bf16_in(w2q, k);
        const float h2 = a2 > 0.f ? a2 : 0.f;
        // output o = b3[o] + sum over the hidden units of h2 * w3[unit][o]
        mode = sel.mode_s1;
        float best = 3.4e38f;
#pragma unroll
        for (int o = 0; o < kSelMaxOut; ++o) {
            if (o >= sel.n_out) break;  // uniform
            float t = h2 * bf16_in(w3q, o);
            for (int x = 32; x > 0; x >>= 1) t += __shfl_xor(t, x);
            const float yo = t + bf16_to_f32((uint16_t)__shfl((int32_t)b3v, o));
            if (((sel.allowed >> o) & 1) && yo < best) {
                best = yo;
                mode = o;
            }
        }
    }
    if (threadIdx.x == 0) {
        st->count[p] = n;
        st->fdeg[p] = fdeg;
        st->unvisited = unv;
        st->unvisited_edges = unv_e;
        st->mode = mode;
        st->level = next_level;
        st->hub_n[(next_level + 1) & 1] = 0;  // the list the step walked: the next step refills it
        if (mode >= BFS_PULL_THREAD) st->pull_levels += 1;
        if (next_level < kBfsMaxLevels) {
            st->hist_mode[next_level] = mode;
            st->hist_next[next_level] = n;
        }
    }
}

__global__ void __launch_bounds__(kDecideBlock) bfs_decide_kernel(DecideArgs d, SelectorView sel) {
    decide_body<kDecideBlock>(d, sel);
}

// A level's step: the push or the pull body, as the decided mode says.
__device__ __forceinline__ void level_body(const StepArgs &a) {
    __shared__ int32_t pre[kMaxSumWords + 1];
    if (a.st->done) return;                                            // uniform: set by a decide kernel
    if (a.st->mode >= BFS_PULL_THREAD) step_body<kPullStep>(a, pre);  // uniform
    else step_body<kPushStep>(a, pre);
}

// The recorded level loop launches bfs_step. Traced levels (per-level
// timing) and sweeps launch the same code, traced levels under two more
// symbols, one per direction, so a profile tells their launches apart. Both
// step bodies together need 72 VGPRs: 7 waves per SIMD without spilling
// measured faster than 8 with a spill.
#define GI_STEP_KERNEL(name)                                                               \
    __global__ void __launch_bounds__(kStepBlock) __attribute__((amdgpu_waves_per_eu(7))) \
    name(StepArgs a) {                                                                    \
        level_body(a);                                                                     \
    }
GI_STEP_KERNEL(bfs_step)
GI_STEP_KERNEL(bfs_push_step_traced)
GI_STEP_KERNEL(bfs_pull_step_traced)
#undef GI_STEP_KERNEL

struct SeedArgs {
    int32_t *depth;
    int32_t *parent;
    int32_t lane_max;
    int32_t hub_min;
    int32_t want_fdeg;
    int32_t want_hubs;
    int32_t hub_list_count;
    long long hub_list_fdeg;
};

struct ResetArgs {
    int32_t *depth;
    int32_t depth_len;      // depth entries to clear (the rest is never read)
    int32_t *parent;
    uint64_t *visited;
    const uint64_t *dead;   // vertices without in-edges, and the padding past V
    uint64_t *bm[3];
    uint64_t *sum[3];
    int32_t V;
    int32_t words;
    int32_t sum_words;
};

// Clears every array of a run in one pass and initialises the state.
__global__ void bfs_reset_kernel(BfsState *st, SeedArgs sa, ResetArgs r, int64_t E) {
    const int32_t stride = gridDim.x * blockDim.x;
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // depths [0, depth_len) and parents [0, V): 16-byte stores where both
    // arrays are 16-byte aligned, then the remainder one by one
    const int32_t dlen = r.depth_len, plen = r.parent ? r.V : 0;
    const int32_t vlen = dlen > plen ? dlen : plen;
    const bool aligned = ((reinterpret_cast<uintptr_t>(r.depth) | reinterpret_cast<uintptr_t>(r.parent)) & 15) == 0;
    const int32_t head = aligned ? vlen / 4 * 4 : 0;
    const int4 none = make_int4(-1, -1, -1, -1);
    for (int32_t q = tid; q < head / 4; q += stride) {
        const int32_t i = 4 * q;
        if (i + 4 <= dlen) reinterpret_cast<int4 *>(r.depth)[q] = none;
        else
            for (int k = 0; k < 4; ++k)
                if (i + k < dlen) r.depth[i + k] = -1;
        if (i + 4 <= plen) reinterpret_cast<int4 *>(r.parent)[q] = none;
        else
            for (int k = 0; k < 4; ++k)
                if (i + k < plen) r.parent[i + k] = -1;
    }
    for (int32_t i = head + tid; i < vlen; i += stride) {
        if (i < dlen) r.depth[i] = -1;
        if (i < plen) r.parent[i] = -1;
    }
    const int32_t n = r.words > r.sum_words + 1 ? r.words : r.sum_words + 1;
    for (int32_t i = tid; i < n; i += stride) {
        if (i < r.words) {
            r.visited[i] = r.dead[i];
            r.bm[0][i] = 0;
            r.bm[1][i] = 0;
            r.bm[2][i] = 0;
        }
        if (i <= r.sum_words) {
            r.sum[0][i] = 0;
            r.sum[1][i] = 0;
            r.sum[2][i] = 0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->count[0] = 0;
        st->count[1] = 0;
        st->fdeg[0] = 0;
        st->fdeg[1] = 0;
        st->hub_n[0] = 0;
        st->hub_n[1] = 0;
        st->want_hubs = sa.want_hubs;
        st->hub_list_count = sa.hub_list_count;
        st->hub_list_fdeg = sa.hub_list_fdeg;
        st->done = 0;
        st->level = -1;
        st->levels = 0;
        st->pull_levels = 0;
        st->mode = BFS_PUSH_THREAD;
        st->unvisited = r.V;
        st->unvisited_edges = E;
        st->lane_max = sa.lane_max;
        st->hub_min = sa.hub_min;
        st->want_fdeg = sa.want_fdeg;
        st->depth = sa.depth;
        st->parent = sa.parent;
    }
}

// Pull-side tables of a graph: the dead mask (no in-edges, or past V) and
// the candidate parent of each vertex, its in-neighbour of largest out-degree
// among the first kCandScan (the neighbour most likely to be reached early).
__global__ void bfs_dead_kernel(const int32_t *poff, int32_t V, int32_t words, uint64_t *dead) {
    for (int32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x) {
        uint64_t m = 0;
        for (int b = 0; b < 64; ++b) {
            int32_t v = w * 64 + b;
            if (v >= V || poff[v + 1] == poff[v]) m |= 1ull << b;
        }
        dead[w] = m;
    }
}

__global__ void bfs_cand_kernel(const int32_t *poff, const int32_t *pnbr, const int32_t *off,
                                int32_t V, int32_t first, int32_t *cand) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        int32_t e0 = poff[v], e1 = poff[v + 1];
        if (first) {  // the first in-neighbour (a degree-ordered view lists the widest first)
            cand[v] = e0 < e1 ? pnbr[e0] : -1;
            continue;
        }
        if (e1 - e0 > kCandScan) e1 = e0 + kCandScan;
        int32_t best = -1, bd = -1;
        for (int32_t e = e0; e < e1; ++e) {
            int32_t u = pnbr[e];
            int32_t d = off[u + 1] - off[u];
            if (d > bd) {
                bd = d;
                best = u;
            }
        }
        cand[v] = best;
    }
}

// Seeds one or many sources (a multi-source BFS gives every vertex its
// distance to the nearest source; over a disjoint union of graphs with one
// source each it is every graph's own BFS). Repeated sources count once.
__global__ void bfs_seed_sources_kernel(BfsState *st, const int32_t *src, int32_t n,
                                        int32_t single, uint64_t *bm0, uint64_t *sum0,
                                        uint64_t *visited, const int32_t *off) {
    int32_t cnt = 0;
    unsigned long long fd = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int32_t v = src ? src[i] : single;
        uint64_t bit = 1ull << (v & 63);
        unsigned long long old = atomicOr((unsigned long long *)&bm0[v >> 6], bit);
        if (old & bit) continue;
        atomicOr((unsigned long long *)&visited[v >> 6], bit);
        st->depth[v] = 0;
        if (st->parent) st->parent[v] = v;
        int32_t ch = (v >> 6) / kChunkWords;
        atomicOr((unsigned long long *)&sum0[ch >> 6], 1ull << (ch & 63));
        cnt += 1;
        const int32_t dv = off[v + 1] - off[v];
        fd += (unsigned long long)dv;
        if (st->want_hubs && dv >= st->hub_min) {  // the first step's grid bin
            const int32_t p = atomicAdd(&st->hub_n[0], 1);
            if (p < kMaxHubs) st->hubs[0][p] = v;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o);
        fd += __shfl_xor(fd, o);
    }
    if (lane_id() == 0 && cnt) {
        atomicAdd(&st->count[0], cnt);
        atomicAdd(&st->fdeg[0], fd);
    }
}

__global__ void count_reached_fused_kernel(const int32_t *depth, int32_t V, int32_t *out) {
    int32_t cnt = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        cnt += depth[v] >= 0;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane_id() == 0 && cnt) atomicAdd(out, cnt);
}

// ---------------------------------------------------------------- host side
struct FusedWork {
    BfsState *st = nullptr;
    uint64_t *bm[3] = {nullptr, nullptr, nullptr};
    uint64_t *sum[3] = {nullptr, nullptr, nullptr};
    uint8_t *cflag = nullptr;        // chunk flags (bs_flag_chunk), zero between steps
    int32_t *pcount = nullptr;
    unsigned long long *pfdeg = nullptr;
    int32_t *reached = nullptr;
    uint64_t *dead = nullptr;        // pull-side tables (see bfs_dead_kernel)
    int32_t *cand = nullptr;
    int32_t cand_first = 0;          // cand holds first in-neighbours (degree-ordered views)
    const int32_t *poff = nullptr;   // in-edge CSR (the graph itself when symmetric)
    const int32_t *pnbr = nullptr;
    SelectorView last_sel{};         // selector of the current run (passed to the decides)
    int32_t *d_src = nullptr;        // sources of a multi-source run
    int32_t src_cap = 0;
    int32_t *host_flag = nullptr;    // pinned, coherent {done, level, levels, mode}: written by the
                                     // decide that ends a run, or copied in by the traced path
    hipEvent_t flag_ev = nullptr;    // recorded after each replay of a block of levels
    int32_t sum_words = 0;
    int32_t lane_max = 64;           // bin thresholds (gi_bfs_set_bins)
    int32_t hub_min = kHubDegree;
    int32_t blocks_per_cu = 8;
    // level loop recorded as a hipGraph, one exec per level count
    // recorded level loops by (level count, selector view bytes): the decide
    // kernel takes the view by value
    std::map<std::pair<int32_t, std::string>, hipGraphExec_t> execs;
    int32_t recent_levels[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // depths of the last runs (ring)
    int32_t recent_n = 0;
};

static PlanRegistry<FusedWork> g_fused;

static void drop_execs(FusedWork *w) {
    for (auto &kv : w->execs) (void)hipGraphExecDestroy(kv.second);
    w->execs.clear();
}

static void free_work(FusedWork *w) {
    drop_execs(w);
    dfree(w->st);
    for (auto p : w->bm) dfree(p);
    for (auto p : w->sum) dfree(p);
    dfree(w->pcount);
    dfree(w->pfdeg);
    dfree(w->cflag);
    dfree(w->reached);
    dfree(w->dead);
    dfree(w->cand);
    dfree(w->d_src);
    if (w->host_flag) (void)hipHostFree(w->host_flag);
    if (w->flag_ev) (void)hipEventDestroy(w->flag_ev);
    delete w;
}

static int ensure_fused(gi_graph *g, FusedWork **out) {
    if ((*out = g_fused.find(g)) != nullptr) return GI_OK;
    int64_t words = bitmap_words(g->V);
    int64_t chunks = ceil_div(words, kChunkWords);
    int64_t sum_words = ceil_div(chunks, 64);
    if (sum_words > kMaxSumWords)
        return fail(GI_ERR_RANGE, "fused BFS supports up to 2^26 vertices");
    const gi_graph *pg = g;  // in-edges for pull steps
    if (g->directed) {
        gi_graph *t = nullptr;
        GI_TRY(gi_graph_transpose(g, &t));
        pg = t;
    }
    FusedWork *w = new FusedWork();
    w->cand_first = g->degree_view ? 1 : 0;
    w->sum_words = (int32_t)sum_words;
    w->poff = pg->d_off;
    w->pnbr = pg->d_nbr;
    int rc = dalloc(&w->st, 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->bm[i], (size_t)words + 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->sum[i], (size_t)sum_words + 1);
    if (rc == GI_OK) rc = dalloc(&w->pcount, (size_t)kMaxGrid);
    if (rc == GI_OK) rc = dalloc(&w->pfdeg, (size_t)kMaxGrid);
    if (rc == GI_OK) rc = dalloc(&w->cflag, (size_t)bs_flag_bytes(chunks));
    if (rc == GI_OK && hipMemset(w->cflag, 0, (size_t)bs_flag_bytes(chunks)) != hipSuccess)
        rc = fail(GI_ERR_HIP, "flag init");
    if (rc == GI_OK) rc = dalloc(&w->reached, 1);
    if (rc == GI_OK) rc = dalloc(&w->dead, (size_t)words + 1);
    if (rc == GI_OK) rc = dalloc(&w->cand, (size_t)g->V + 1);
    if (rc == GI_OK && !g->d_visited) rc = dalloc(&g->d_visited, (size_t)words + 1);
    if (rc == GI_OK && hipHostMalloc((void **)&w->host_flag, sizeof(int32_t) * 4, hipHostMallocCoherent) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
    if (rc == GI_OK && hipEventCreateWithFlags(&w->flag_ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(GI_ERR_HIP, "event create failed");
    if (rc == GI_OK) rc = fill_i32(w->pcount, kMaxGrid, 0, 0);
    if (rc == GI_OK) rc = fill_u64((uint64_t *)w->pfdeg, kMaxGrid, 0, 0);
    if (rc == GI_OK) {
        int grid = (int)ceil_div(g->V, 256);
        grid = grid < 4096 ? (grid < 1 ? 1 : grid) : 4096;
        bfs_dead_kernel<<<grid, 256>>>(w->poff, g->V, (int32_t)words, w->dead);
        bfs_cand_kernel<<<grid, 256>>>(w->poff, w->pnbr, g->d_off, g->V, w->cand_first, w->cand);
        if (hipGetLastError() != hipSuccess) rc = fail(GI_ERR_HIP, "pull table launch failed");
    }
    if (rc == GI_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(GI_ERR_HIP, "init sync");
    if (rc != GI_OK) {
        free_work(w);
        return rc;
    }
    g_fused.put(g, w);
    *out = w;
    return GI_OK;
}

void release_fused(const gi_graph *g) {
    if (FusedWork *p = g_fused.take(g)) free_work(p);
}

int32_t fused_mode_of(const gi_schedule &s) {
    if (s.direction == GI_PULL) return s.load_balancing == GI_LB_VERTEX_BASED ? BFS_PULL_THREAD : BFS_PULL_GROUP;
    switch (s.load_balancing) {
        case GI_LB_VERTEX_BASED: return BFS_PUSH_THREAD;
        case GI_LB_WM: return BFS_PUSH_WAVE;
        case GI_LB_TWCE: return BFS_PUSH_TWCE;
        case GI_LB_CM: return BFS_PUSH_CM;
        default: return BFS_PUSH_TWC;  // TWC, STRICT, EDGE_ONLY: grid-binned
    }
}

// Persistent grids: as many workgroups as can be resident at once (the
// step kernels assign work statically, so a second round of workgroups
// would serialise), capped by the tuned blocks per CU.
static int grid_of(const FusedWork *w, int resident) {
    int per = w->blocks_per_cu < resident ? w->blocks_per_cu : resident;
    int g = device_info().num_cus * per;
    return g < kMaxGrid ? g : kMaxGrid;
}

// Every level, traced or not, runs on the same persistent grid; each of its
// workgroups writes one partial-sum slot.
static int step_grid(const FusedWork *w) {
    static const int resident = resident_per_cu((const void *)bfs_step, kStepBlock);
    return grid_of(w, resident);
}

static StepArgs step_args(gi_graph *g, FusedWork *w) {
    StepArgs a;
    a.st = w->st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.poff = w->poff;
    a.pnbr = w->pnbr;
    a.cand = w->cand;
    a.cand_first = w->cand_first;
    a.V = g->V;
    a.words = (int32_t)bitmap_words(g->V);
    a.chunks = (int32_t)ceil_div(a.words, kChunkWords);
    a.sum_words = w->sum_words;
    a.depth = nullptr;   // read from the state inside the kernels
    a.parent = nullptr;
    a.visited = g->d_visited;
    for (int i = 0; i < 3; ++i) {
        a.bm[i] = w->bm[i];
        a.sum[i] = w->sum[i];
    }
    a.cflag = w->cflag;
    a.pcount = w->pcount;
    a.pfdeg = w->pfdeg;
    return a;
}

static DecideArgs decide_args(gi_graph *g, FusedWork *w) {
    DecideArgs d;
    d.st = w->st;
    d.V = g->V;
    d.E = g->E;
    d.pcount = w->pcount;
    d.pfdeg = w->pfdeg;
    d.slots = step_grid(w);
    for (int i = 0; i < 3; ++i) d.sum[i] = w->sum[i];
    d.sum_words = w->sum_words;
    d.cflag = w->cflag;
    d.chunks = (int32_t)ceil_div(bitmap_words(g->V), kChunkWords);
    d.host_flag = w->host_flag;
    d.off = g->d_off;
    d.visited = g->d_visited;
    d.bm0 = w->bm[0];
    d.seed = -1;
    return d;
}

static void launch_decide(gi_graph *g, FusedWork *w, Launcher &L, int32_t seed = -1) {
    // the kernel reads the level from the device state
    DecideArgs d = decide_args(g, w);
    d.seed = seed;
    L(bfs_decide_kernel, 1, kDecideBlock, d, w->last_sel);
}

static void launch_decide(gi_graph *g, FusedWork *w, hipStream_t s, int32_t seed = -1) {
    Launcher L(s);
    launch_decide(g, w, L, seed);
}

// A recorded level: the step kernel (push or pull body as decided, the TWC
// grid bin included), the decide kernel.
static void launch_level(gi_graph *g, FusedWork *w, const StepArgs &a, Launcher &L) {
    L(bfs_step, step_grid(w), kStepBlock, a);
    launch_decide(g, w, L);
}

// A traced level with the mode already known on the host: the step code
// under its per-direction symbol alone between the two events, then the
// decide kernel.
static void launch_level_traced(gi_graph *g, FusedWork *w, const StepArgs &a, int32_t mode, hipStream_t s,
                                hipEvent_t *ev) {
    (void)hipEventRecord(ev[0], s);
    if (mode >= BFS_PULL_THREAD) bfs_pull_step_traced<<<step_grid(w), kStepBlock, 0, s>>>(a);
    else bfs_push_step_traced<<<step_grid(w), kStepBlock, 0, s>>>(a);
    (void)hipEventRecord(ev[1], s);
    launch_decide(g, w, s);
}

// Seed: clear the arrays, set the state, mark the sources, decide level 0.
// sources == null seeds the single vertex `single`.
static int seed_run(gi_graph *g, FusedWork *w, const int32_t *h_sources, int32_t n_sources,
                    int32_t single, int32_t *depth, int32_t *parent, const SelectorView &sel,
                    bool want_fdeg, hipStream_t s, int32_t depth_len = -1) {
    w->last_sel = sel;  // recorded level graphs are kept per selector view (level_graph)
    // hubs are listed at discovery when a TWC step may use the list: the
    // selector may pick TWC, or the degrees are read anyway
    const bool learned = sel.mode == GI_SELECT_LEARNED && sel.w1 != nullptr;
    const bool may_twc = learned ? ((sel.allowed >> BFS_PUSH_TWC) & 1) != 0
                                 : (sel.mode_s1 == BFS_PUSH_TWC || sel.mode_s2 == BFS_PUSH_TWC);
    // ... during push steps of at most V/64 vertices and E/64 edges
    SeedArgs sa{depth,
                parent,
                w->lane_max,
                w->hub_min,
                want_fdeg ? 1 : 0,
                (want_fdeg || may_twc) ? 1 : 0,
                g->V / kHubListShare,
                (long long)(g->E / kHubListShare)};
    w->host_flag[0] = 0;  // no run of this graph's engine is in flight: set again by the final decide
    ResetArgs r;
    r.depth = depth;
    r.depth_len = depth_len >= 0 && depth_len < g->V ? depth_len : g->V;
    r.parent = parent;
    r.visited = g->d_visited;
    r.dead = w->dead;
    for (int i = 0; i < 3; ++i) {
        r.bm[i] = w->bm[i];
        r.sum[i] = w->sum[i];
    }
    r.V = g->V;
    r.words = (int32_t)bitmap_words(g->V);
    r.sum_words = w->sum_words;
    int grid = (int)ceil_div(g->V > w->sum_words + 1 ? g->V : w->sum_words + 1, 256);
    grid = grid < 4096 ? grid : 4096;
    bfs_reset_kernel<<<grid, 256, 0, s>>>(w->st, sa, r, g->E);
    if (h_sources) {
        if (n_sources > w->src_cap) {
            dfree(w->d_src);
            w->d_src = nullptr;
            w->src_cap = 0;
            GI_TRY(dalloc(&w->d_src, (size_t)n_sources));
            w->src_cap = n_sources;
        }
        GI_HIP(hipMemcpyAsync(w->d_src, h_sources, sizeof(int32_t) * n_sources, hipMemcpyHostToDevice, s));
        int sg = (int)ceil_div(n_sources, 256);
        bfs_seed_sources_kernel<<<sg < 1024 ? sg : 1024, 256, 0, s>>>(w->st, w->d_src, n_sources, 0, w->bm[0],
                                                                      w->sum[0], g->d_visited, g->d_off);
    }
    launch_decide(g, w, s, h_sources ? -1 : single);  // a single source is seeded by this decide
    GI_LAUNCH_CHECK();
    if (h_sources) GI_HIP(hipStreamSynchronize(s));  // h_sources may go away on return
    return GI_OK;
}

// A hipGraph of `levels` consecutive levels (step, hub, decide each); every
// kernel reads its level from the device state, so one recording serves every
// run on this graph.
static int level_graph(gi_graph *g, FusedWork *w, int32_t levels, hipGraphExec_t *out) {
    const std::pair<int32_t, std::string> key(
        levels, std::string(reinterpret_cast<const char *>(&w->last_sel), sizeof(SelectorView)));
    auto it = w->execs.find(key);
    if (it != w->execs.end()) {
        *out = it->second;
        return GI_OK;
    }
    StepArgs a = step_args(g, w);
    hipGraphExec_t exec = nullptr;
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t l = 0; l < levels; ++l) launch_level(g, w, a, L);
        },
        &exec));
    if (w->execs.size() >= 16) drop_execs(w);
    w->execs[key] = exec;
    *out = exec;
    return GI_OK;
}

static int run_levels(gi_graph *g, FusedWork *w, hipStream_t s, FusedTrace *trace, bool timed,
                      const FusedEpilogue *epilogue, int32_t *levels) {
    StepArgs a = step_args(g, w);
    int32_t *done_flag = &w->st->done;
    static_assert(offsetof(BfsState, levels) == offsetof(BfsState, done) + 2 * sizeof(int32_t),
                  "done, level, levels are contiguous");
    if (trace && timed) {
        // per-level launches with events around each step; the host reads
        // {done, level, levels, mode} after every decide
        static_assert(offsetof(BfsState, mode) == offsetof(BfsState, done) + 3 * sizeof(int32_t),
                      "done, level, levels, mode are contiguous");
        // a BFS has at most V levels: the bound only guards against a state
        // that never reports done
        for (int64_t level = 0; level <= (int64_t)g->V + 1; ++level) {
            GI_HIP(hipMemcpyAsync(w->host_flag, done_flag, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            GI_HIP(hipStreamSynchronize(s));
            if (w->host_flag[0]) {
                *levels = w->host_flag[2];
                return epilogue ? (*epilogue)(s) : GI_OK;
            }
            hipEvent_t e0, e1;
            GI_HIP(hipEventCreate(&e0));
            GI_HIP(hipEventCreate(&e1));
            trace->ev.push_back(e0);
            trace->ev.push_back(e1);
            launch_level_traced(g, w, a, w->host_flag[3], s, &trace->ev[trace->ev.size() - 2]);
            GI_LAUNCH_CHECK();
        }
        return fail(GI_ERR_HIP, "BFS state never reported done");
    }
    // Graph mode: replay a recorded block of levels sized by the deepest of
    // the last runs on this graph (a level past the end costs one step launch
    // that returns at once, a block too short a whole extra replay), wait for
    // each replay's event (the epilogue is already queued behind it) and read
    // the {done, level, levels} the final decide wrote to pinned memory,
    // continue with blocks of 8, 16, 32, 48 levels for deeper runs.
    int32_t block = 0;
    for (int32_t l : w->recent_levels) block = l > block ? l : block;
    if (block <= 0) block = 8;
    if (block > 48) block = 48;
    int32_t next_block = 8;
    int64_t level = 0;
    while (level <= (int64_t)g->V + 1) {  // no depth limit: a BFS has at most V levels
        hipGraphExec_t exec = nullptr;
        GI_TRY(level_graph(g, w, block, &exec));
        GI_HIP(hipGraphLaunch(exec, s));
        level += block;
        // the decide that ends the run writes {done, level, levels}
        // into the pinned host_flag itself (the host cleared done before the run)
        GI_HIP(hipEventRecord(w->flag_ev, s));
        if (epilogue) GI_TRY((*epilogue)(s));
        GI_HIP(hipEventSynchronize(w->flag_ev));
        if (w->host_flag[0]) {
            *levels = w->host_flag[2];
            return GI_OK;
        }
        block = next_block;
        next_block = next_block < 48 ? (next_block * 2 < 48 ? next_block * 2 : 48) : 48;
    }
    return fail(GI_ERR_HIP, "BFS state never reported done");
}

int bfs_fused_run_multi(gi_graph *g, const int32_t *h_sources, int32_t n_sources, int32_t single,
                        const gi_hybrid_schedule *h, SelectorView sel, int32_t *d_depth,
                        int32_t *d_parent, hipStream_t s, FusedTrace *trace, bool timed,
                        int32_t *levels_out, const FusedEpilogue *epilogue, int32_t depth_len) {
    FusedWork *w = nullptr;
    GI_TRY(ensure_fused(g, &w));
    sel.mode_s1 = fused_mode_of(h->s1);
    sel.mode_s2 = fused_mode_of(h->s2);
    sel.threshold = h->threshold;
    if (sel.mode != GI_SELECT_LEARNED) sel.mode = h->selector;
    GI_TRY(seed_run(g, w, h_sources, n_sources, single, d_depth, d_parent, sel,
                    sel.mode == GI_SELECT_LEARNED, s, depth_len));
    int32_t levels = 0;
    GI_TRY(run_levels(g, w, s, trace, timed, epilogue, &levels));
    w->recent_levels[w->recent_n++ % 8] = levels;
    if (levels_out) *levels_out = levels;
    if (trace) {
        int32_t n = levels < kBfsMaxLevels ? levels : kBfsMaxLevels;
        trace->modes.resize(n);
        trace->sizes.resize(n);
        GI_HIP(hipMemcpy(trace->modes.data(), w->st->hist_mode, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        GI_HIP(hipMemcpy(trace->sizes.data(), w->st->hist_next, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        GI_HIP(hipMemcpy(&trace->pull_levels, &w->st->pull_levels, sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    return GI_OK;
}

int bfs_fused_run(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, SelectorView sel,
                  int32_t *d_depth, int32_t *d_parent, hipStream_t s, FusedTrace *trace,
                  bool timed, int32_t *levels_out, const FusedEpilogue *epilogue, int32_t depth_len) {
    return bfs_fused_run_multi(g, nullptr, 0, source, h, sel, d_depth, d_parent, s, trace, timed,
                               levels_out, epilogue, depth_len);
}

int fused_reached(gi_graph *g, const int32_t *depth, int32_t *out, hipStream_t s) {
    FusedWork *w = nullptr;
    GI_TRY(ensure_fused(g, &w));
    GI_TRY(fill_i32(w->reached, 1, 0, s));
    count_reached_fused_kernel<<<step_grid(w), 256, 0, s>>>(depth, g->V, w->reached);
    GI_LAUNCH_CHECK();
    return read_count(w->reached, out, s);
}

int fused_set_bins(gi_graph *g, int32_t lane_max, int32_t hub_min, int32_t blocks_per_cu) {
    FusedWork *w = nullptr;
    GI_TRY(ensure_fused(g, &w));
    if (lane_max < 1 || lane_max > 4096 || hub_min < 64 || blocks_per_cu < 1 || blocks_per_cu > 8)
        return fail(GI_ERR_RANGE, "bins out of range");
    w->lane_max = lane_max;
    w->hub_min = hub_min;
    if (w->blocks_per_cu != blocks_per_cu) drop_execs(w);  // grid size is baked into the graphs
    w->blocks_per_cu = blocks_per_cu;
    return GI_OK;
}

// ---------------------------------------------------------------- sweep
__global__ void bfs_force_mode_kernel(BfsState *st, int32_t mode) {
    st->mode = mode;
    st->hub_n[(st->level + 1) & 1] = 0;  // every replay lists the hubs of F' afresh
}

__global__ void zero_slots_kernel(int32_t *pcount, unsigned long long *pfdeg) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMaxGrid; i += gridDim.x * blockDim.x) {
        pcount[i] = 0;
        pfdeg[i] = 0;
    }
}

struct SweepHost {
    int32_t count[2];
    unsigned long long fdeg[2];
    int32_t done;
    int32_t level;
    int32_t levels;
    int32_t mode;
    long long unvisited;
    long long unvisited_edges;
};

}  // namespace gi

using namespace gi;

extern "C" {

int gi_bfs_set_bins(gi_graph *g, int32_t lane_max, int32_t hub_min, int32_t blocks_per_cu) {
    if (!g) return fail(GI_ERR_INVALID, "null graph");
    // the degree-ordered view is where the fused engine runs when there is one
    if (g->layout) GI_TRY(fused_set_bins(layout_view(g), lane_max, hub_min, blocks_per_cu));
    return fused_set_bins(g, lane_max, hub_min, blocks_per_cu);
}

int gi_bfs_multi(gi_graph *g, const int32_t *sources, int32_t num_sources,
                 const gi_hybrid_schedule *h, int32_t *d_depth, int32_t *d_parent,
                 int32_t *levels, void *stream) {
    if (!g || !h || !d_depth || (num_sources > 0 && !sources)) return fail(GI_ERR_INVALID, "null argument");
    if (num_sources < 1) return fail(GI_ERR_RANGE, "no sources");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    for (int32_t i = 0; i < num_sources; ++i)
        if (sources[i] < 0 || sources[i] >= g->V) return fail(GI_ERR_RANGE, "source %d out of range", sources[i]);
    GI_TRY(validate_schedule(&h->s1));
    GI_TRY(validate_schedule(&h->s2));
    SelectorView sel{};
    if (h->selector == GI_SELECT_LEARNED) sel = selector_view();
    if (!sel.w1) sel.mode = h->selector == GI_SELECT_LEARNED ? GI_SELECT_HEURISTIC : h->selector;
    if (!g->layout)
        return bfs_fused_run_multi(g, sources, num_sources, 0, h, sel, d_depth, d_parent,
                                   (hipStream_t)stream, nullptr, false, levels);
    // degree-ordered view: map the sources in, the results out
    std::vector<int32_t> mapped((size_t)num_sources);
    for (int32_t i = 0; i < num_sources; ++i) mapped[i] = layout_map_source(g, sources[i]);
    hipStream_t s = (hipStream_t)stream;
    int32_t *rdepth = layout_tmp(g, 0), *rparent = d_parent ? layout_tmp(g, 1) : nullptr;
    GI_TRY(bfs_fused_run_multi(layout_view(g), mapped.data(), num_sources, 0, h, sel, rdepth, rparent, s,
                               nullptr, false, levels));
    GI_TRY(layout_map_back(g, rdepth, d_depth, false, s));
    if (d_parent) GI_TRY(layout_map_back(g, rparent, d_parent, true, s));
    return GI_OK;
}

int gi_bfs_sweep(gi_graph *g, int32_t source, int32_t mode_mask, int32_t follow, int32_t reps,
                 float *features, float *times_ms, int32_t max_levels, int32_t *num_levels,
                 void *stream) {
    if (!g || !features || !times_ms || !num_levels) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (source < 0 || source >= g->V) return fail(GI_ERR_RANGE, "source %d out of range", source);
    if ((mode_mask & ((1 << BFS_NUM_MODES) - 1)) == 0) return fail(GI_ERR_INVALID, "empty mode mask");
    if (follow >= BFS_NUM_MODES) return fail(GI_ERR_RANGE, "bad follow mode");
    if (reps < 1) reps = 1;
    hipStream_t s = (hipStream_t)stream;
    if (g->layout) {  // time the modes where the engine runs: on the degree-ordered view
        source = layout_map_source(g, source);
        g = layout_view(g);
    }
    FusedWork *w = nullptr;
    GI_TRY(ensure_fused(g, &w));
    int64_t words = bitmap_words(g->V);
    int32_t *depth = nullptr, *depth_bak = nullptr;
    uint64_t *vis_bak = nullptr;
    GI_TRY(dalloc(&depth, (size_t)g->V));
    int rc = dalloc(&depth_bak, (size_t)g->V);
    if (rc == GI_OK) rc = dalloc(&vis_bak, (size_t)words + 1);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == GI_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(GI_ERR_HIP, "event create");
    SelectorView sel{};
    sel.mode = GI_SELECT_S1;
    sel.mode_s1 = BFS_PUSH_THREAD;
    sel.mode_s2 = BFS_PUSH_THREAD;
    if (rc == GI_OK) rc = seed_run(g, w, nullptr, 0, source, depth, nullptr, sel, true, s);
    StepArgs a = step_args(g, w);
    int32_t nl = 0, prev_mode = BFS_PUSH_THREAD, prev_n = 0;
    // One level: snapshot the state, time every mode of the mask from it with
    // the production step kernel (bfs_step), then advance by the followed
    // mode. Returns 1 when the BFS is done; errors return
    // through rc so the buffers below are freed.
    auto one_level = [&]() -> int {
        SweepHost h;
        GI_HIP(hipMemcpyAsync(&h, w->st, sizeof(h), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        if (h.done) return 1;
        int32_t p = h.level & 1;
        bfs_raw_features(features + (int64_t)nl * kSelInputs, h.count[p], h.fdeg[p], h.unvisited,
                         h.unvisited_edges, h.level, prev_mode, prev_n, g->V, g->E);
        const int32_t iout = (h.level + 1) % 3;
        // the step clears bitmap (level + 2) % 3, already empty after the first run
        GI_HIP(hipMemcpyAsync(depth_bak, depth, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice, s));
        GI_HIP(hipMemcpyAsync(vis_bak, g->d_visited, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, s));
        float best = 3.4e38f;
        int32_t best_mode = -1;
        for (int32_t m = 0; m <= BFS_NUM_MODES; ++m) {
            bool final_run = m == BFS_NUM_MODES;
            if (!final_run && !((mode_mask >> m) & 1)) {
                times_ms[(int64_t)nl * BFS_NUM_MODES + m] = -1.f;
                continue;
            }
            int32_t run = final_run ? (follow >= 0 ? follow : best_mode) : m;
            float tmin = 3.4e38f;
            for (int32_t r = 0; r < (final_run ? 1 : reps); ++r) {
                GI_HIP(hipMemcpyAsync(depth, depth_bak, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice, s));
                GI_HIP(hipMemcpyAsync(g->d_visited, vis_bak, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, s));
                GI_TRY(fill_u64(w->bm[iout], words, 0, s));
                GI_TRY(fill_u64(w->sum[iout], w->sum_words, 0, s));
                GI_TRY(fill_u64(reinterpret_cast<uint64_t *>(w->cflag), bs_flag_bytes(a.chunks) / 8, 0, s));
                zero_slots_kernel<<<8, 256, 0, s>>>(w->pcount, w->pfdeg);
                bfs_force_mode_kernel<<<1, 1, 0, s>>>(w->st, run);
                GI_HIP(hipEventRecord(e0, s));
                bfs_step<<<step_grid(w), kStepBlock, 0, s>>>(a);
                GI_HIP(hipEventRecord(e1, s));
                GI_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                GI_HIP(hipEventElapsedTime(&ms, e0, e1));
                tmin = ms < tmin ? ms : tmin;
            }
            if (!final_run) {
                times_ms[(int64_t)nl * BFS_NUM_MODES + m] = tmin;
                if (tmin < best) {
                    best = tmin;
                    best_mode = m;
                }
            } else {
                launch_decide(g, w, s);
                GI_LAUNCH_CHECK();
                prev_mode = run;
            }
        }
        prev_n = h.count[p];
        ++nl;
        return GI_OK;
    };
    while (rc == GI_OK && nl < max_levels) {
        int r = one_level();
        if (r == 1) break;
        rc = r;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(depth);
    dfree(depth_bak);
    dfree(vis_bak);
    *num_levels = nl;
    return rc;
}

}  // extern "C"
