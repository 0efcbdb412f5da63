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
constexpr int kChunkWords = 64;                  // chunk: 64 words = 4096 vertices
constexpr int kChunkVerts = kChunkWords * 64;
constexpr int kCandScan = 128;                   // in-neighbours examined per candidate
static_assert(kChunkWords == kBsChunkWords, "chunk flags mark the step's chunks");
// A level's |F| and degree sum are added up in kStripes stripes, each on a
// 128-byte line of its own (one atomic per workgroup and counter, spread
// over the lines by workgroup), read back by the next launch.
constexpr int kStripes = 64;                     // one per lane of the reading wave
constexpr int kStripeI32 = 32;
constexpr int kStripeU64 = 16;
constexpr int kCountSets = 3;                    // by level % 3: read, written, cleared
constexpr int kFlagSets = 4;                     // by level % 4 (see level_body)

struct StepArgs {
    BfsState *st;
    const int32_t *off;         // out-edges (push)
    const int32_t *nbr;
    const int32_t *poff;        // in-edges (pull); the same arrays on a symmetric graph
    const int32_t *pnbr;
    const int32_t *cand;        // candidate parent per vertex, -1 when none
    int32_t cand_first;         // 1: the candidate is the vertex's first in-neighbour
    int32_t V;
    int64_t E;
    int32_t words;
    int32_t chunks;
    int32_t *depth;
    int32_t *parent;
    uint64_t *visited;
    uint64_t *bm[3];            // F_l in bm[l % 3]
    uint8_t *flags;             // kFlagSets sets of chunk flags (flag_chunk), flag_bytes apart:
    int32_t flag_bytes;         // those of F_l in set (l - 1) % 4
    int32_t *cnt;               // [kCountSets][kStripes] lines: |F_l| in set (l - 1) % 3
    unsigned long long *fdg;    // the same for the degree sums (in the same lines)
    int32_t *host_flag;         // pinned {done, level, levels}: set when a run ends
    const uint16_t *selw;       // the engine's copy of the selector's BF16 blob (kSelBlob entries)
};

struct StepCtx {
    int32_t level;
    int32_t lane_max;          // degree < lane_max: one lane walks the vertex
    int32_t hub_min;           // TWC: degree >= hub_min goes to the grid bin
    int32_t want_fdeg;
    int32_t want_hubs;         // list the hubs of F' as they are discovered
    int32_t want_giants;       // list the giant hubs of F' (every step of a run that may use TWC)
    int32_t giant_min;         // degree of a giant hub
    int32_t hub_in;            // the hub lists of F: hubs[level % 3], giants[level % 3]
    int32_t hub_out;           // the lists of F': [(level + 1) % 3]
    int32_t hubs_listed;       // TWC step whose hubs (of F) are all listed: walked grid-wide first
    int32_t giants_listed;     // TWC step whose giant hubs are all listed: likewise
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

__device__ __forceinline__ uint8_t *flag_set(const StepArgs &a, int32_t i) {
    return a.flags + (size_t)i * a.flag_bytes;
}

// Workgroup b of a step launch of G workgroups owns the chunks b + j G (j <
// 64); their flags are stored at b * 64 + j, so that a workgroup reads its
// 64 with one line.
__device__ __forceinline__ int32_t flag_pos(int32_t chunk, int32_t G) { return (chunk % G) * 64 + chunk / G; }

// The flag of the chunk holding bitmap word `word`, from a launch of G
// workgroups' layout.
__device__ __forceinline__ void flag_chunk(uint8_t *flags, int32_t word, int32_t G) {
    flags[flag_pos(word / kChunkWords, G)] = 1;
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
    if (!c.want_fdeg && !c.want_hubs && !c.want_giants) return 0;
    const int32_t dd = deg_of(a.off, d);
    if (c.want_fdeg) c.fdeg += (unsigned long long)dd;
    return dd;
}

// Appends d (degree dd) to the giant or the hub list of F', as the step
// keeps them (one lane; listed hubs are rare, one atomic each).
__device__ __forceinline__ void note_hub(const StepArgs &a, const StepCtx &c, int32_t d, int32_t dd) {
    if (c.want_giants && dd >= c.giant_min) {
        const int32_t p = atomicAdd(&a.st->giant_n[c.hub_out], 1);
        if (p < kMaxGiants) a.st->giants[c.hub_out][p] = d;
    } else if (c.want_hubs && dd >= c.hub_min && dd < c.giant_min) {
        const int32_t p = atomicAdd(&a.st->hub_n[c.hub_out], 1);
        if (p < kMaxHubs) a.st->hubs[c.hub_out][p] = d;
    }
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
    if (o2 == 0) flag_chunk(c.cflag, d >> 6, (int32_t)gridDim.x);
    if (a.parent) a.parent[d] = s;  // the depth is written in word order by the next step
    c.cnt += 1;
    const int32_t dd = note_degree(a, c, d);
    if (dd >= c.hub_min) note_hub(a, c, d, dd);
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
            if (o2 == 0) flag_chunk(c.cflag, key, (int32_t)gridDim.x);
        }
    }
    const uint64_t ends = __ballot(end);
    const int32_t leader = __ffsll((unsigned long long)(ends & (~0ull << lane))) - 1;
    const uint64_t run_won = __shfl(won, leader);
    // bits of the lanes before this one in its run: a target repeated in a
    // run is recorded once, by its first lane
    const uint64_t xp = __shfl_up(x, 1);
    const uint64_t before = lane - 1 >= my_head ? xp : 0ull;
    if (want && (run_won & bit) && !(before & bit)) {
        if (a.parent) a.parent[d] = s;  // the depth is written in word order by the next step
        c.cnt += 1;
        const int32_t dd = note_degree(a, c, d);
        if (dd >= c.hub_min) note_hub(a, c, d, dd);  // rare: one atomic per listed hub
    }
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

// Workgroup reduction of the lane counters into the level's stripes.
__device__ __forceinline__ void flush_counts(const StepArgs &a, int32_t set, StepCtx &c) {
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
        const int32_t line = set * kStripes + (int32_t)(blockIdx.x % kStripes);
        if (tc) atomicAdd(&a.cnt[line * kStripeI32], tc);
        if (tf) atomicAdd(&a.fdg[line * kStripeU64], tf);
    }
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
        const bool todo = has && !(c.giants_listed && dg >= c.giant_min) &&
                          !(c.hubs_listed && dg >= c.hub_min && dg < c.giant_min);
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
        if (found_self) {  // (pull steps list no hubs: giant hubs are found by the first, push levels)
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
// hubs (degree >= hub_min, listed when they were discovered: the giant
// ones, then the others) spread evenly over the grid. Every workgroup scans
// the hub degrees into LDS (`lds`: the step's vertex list, free at this
// point) and walks its share of the concatenated edge range. Every thread of
// the workgroup calls it.
static_assert(kChunkVerts >= 2 * (kMaxHubs + kMaxGiants) + 1, "hub scan fits the vertex list");
constexpr int32_t kHubListShare = 64;  // list hubs during push steps of <= V/64 vertices, <= E/64 edges
__device__ __forceinline__ void hub_phase(const StepArgs &a, StepCtx &c, int32_t ng, int32_t nhub, int32_t *lds) {
    const int32_t *giants = a.st->giants[c.hub_in];
    const int32_t *hubs = a.st->hubs[c.hub_in];
    const int32_t nh = ng + nhub;
    int32_t *hv = lds, *hscan = lds + kMaxHubs + kMaxGiants;
    const int32_t per = (nh + kStepBlock - 1) / kStepBlock;
    const int32_t lo = threadIdx.x * per, hi = lo + per < nh ? lo + per : nh;
    int32_t part = 0;
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = i < ng ? giants[i] : hubs[i - ng];
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
// Step l: F_l = bm[l % 3] expanded by the push or the pull body into
// F_l+1 = bm[(l + 1) % 3], whose chunk flags go to flags[l % 4]; bm[(l + 2)
// % 3] (F_l-1, done with) is cleared on the way. Chunk c belongs to
// workgroup c % G (G the grid): a workgroup clears and lists only its own
// chunks, from the bit masks `fl_clear` (flags of F_l-1) and `fl_list`
// (flags of F_l) of its chunks, bit j for chunk blockIdx + j G, which the
// caller loaded with the level's other inputs. `list` is LDS of the
// calling kernel. `ph` is l mod 12 (the rotations' indices).
enum StepKind : int { kPushStep = 0, kPullStep = 1 };

template <int Kind>
__device__ __forceinline__ void step_body(const StepArgs &a0, int32_t *list, int32_t l, int32_t ph, int32_t mode,
                                          int32_t n, unsigned long long nfdeg, uint64_t fl_clear,
                                          uint64_t fl_list) {
    BfsState *st = a0.st;
    StepArgs a = a0;
    a.depth = st->depth;
    a.parent = st->parent;
    const int32_t p3 = ph % 3;
    StepCtx c;
    c.level = l;
    c.lane_max = st->lane_max;
    c.hub_min = st->hub_min;
    c.want_fdeg = st->want_fdeg;
    // F''s hubs are listed during small push steps only: after a large or a
    // pull step the next step is a pull step or has too many hubs for the
    // list (and the same-address atomics of listing them would cost more
    // than the list saves); such a step voids the list instead
    c.want_hubs = st->want_hubs && mode < BFS_PULL_THREAD && n <= st->hub_list_count &&
                  (!c.want_fdeg || (long long)nfdeg <= st->hub_list_fdeg);
    c.want_giants = st->want_hubs && mode < BFS_PULL_THREAD;
    c.giant_min = st->hub_min * kGiantFactor;
    c.hub_in = p3;
    c.hub_out = (p3 + 1) % 3;
    if (!c.want_hubs && blockIdx.x == 0 && threadIdx.x == 0) st->hub_n[c.hub_out] = kMaxHubs + 1;
    const int32_t nh = st->hub_n[p3], ng = st->giant_n[p3];
    c.hubs_listed = mode == BFS_PUSH_TWC && st->want_hubs && nh <= kMaxHubs;  // else void or never kept
    c.giants_listed = mode == BFS_PUSH_TWC && st->want_hubs && ng <= kMaxGiants;
    c.bm_in = pick3(a0.bm[0], a0.bm[1], a0.bm[2], p3);
    c.bm_out = pick3(a0.bm[0], a0.bm[1], a0.bm[2], (p3 + 1) % 3);
    c.cflag = flag_set(a0, ph & 3);
    c.cnt = 0;
    c.fdeg = 0;
    const int32_t lane = lane_id();

    // the bitmap two levels ahead is free: clear this workgroup's non-empty
    // chunks of it
    {
        uint64_t *bc = pick3(a0.bm[0], a0.bm[1], a0.bm[2], (p3 + 2) % 3);
        for (uint64_t m = fl_clear; m; m &= m - 1) {  // uniform
            const int32_t ch = (int32_t)blockIdx.x + (__ffsll((unsigned long long)m) - 1) * (int32_t)gridDim.x;
            if (threadIdx.x < kChunkWords) {
                const int32_t w = ch * kChunkWords + (int32_t)threadIdx.x;
                if (w < a.words) bc[w] = 0;
            }
        }
    }

    if constexpr (Kind == kPushStep) {
        __shared__ int32_t n_s;
        const int32_t nhw = c.hubs_listed ? nh : 0, ngw = c.giants_listed ? ng : 0;
        if (nhw + ngw > 0) hub_phase(a, c, ngw, nhw, list);  // uniform
        for (uint64_t m = fl_list; m; m &= m - 1) {  // uniform
            const int32_t chunk = (int32_t)blockIdx.x + (__ffsll((unsigned long long)m) - 1) * (int32_t)gridDim.x;
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
            const int32_t nl = n_s;
            for (int32_t b0 = 0; b0 < nl; b0 += kStepBlock)
                push_batch(a, c, mode, list + b0, nl - b0 < kStepBlock ? nl - b0 : kStepBlock);
            __syncthreads();
        }
    } else {
        // pull over all vertices in units of kPullUnit words, one unit per
        // workgroup at a time (each wave takes kPullWords of them), so a
        // unit's chunk flag costs at most one store
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
            if (__syncthreads_or(any) && threadIdx.x == 0) flag_chunk(c.cflag, un * kPullUnit, (int32_t)gridDim.x);
        }
    }
    flush_counts(a, p3, c);
}

// ---------------------------------------------------------------- decide
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}

// The selector's 16 inputs in FP32 with the hardware log2 (v_log_f32): the
// values of bfs_raw_feature (double, used by the host to record training
// data) to within FP32 rounding, at a few dozen instructions instead of the
// hundreds of FP64 ones its logarithms and divisions take on one wavefront.
__device__ __forceinline__ void bfs_features_f32(float *f, int32_t n, unsigned long long fdeg, long long unv,
                                                 long long unv_e, int32_t level, int32_t prev_mode,
                                                 int32_t prev_n, int32_t V, int64_t E) {
    // divisions as reciprocal-multiplies (__fdividef, a few ulp): the network
    // runs at the start of every step, on its critical path
    const float dV = V > 0 ? (float)V : 1.f, dE = E > 0 ? (float)E : 1.f;
    const float nn = (float)n, fd = (float)fdeg, u = (float)unv, ue = (float)unv_e;
    const float l_n = __log2f(1.f + nn);
    f[0] = __fdividef(nn, dV);
    f[1] = __fdividef(fd, dE);
    f[2] = __fdividef(u, dV);
    f[3] = __fdividef(ue, dE);
    f[4] = l_n * (1.f / 32.f);
    f[5] = __log2f(1.f + fd) * (1.f / 40.f);
    f[6] = __log2f(1.f + __fdividef(fd, n > 0 ? nn : 1.f)) * (1.f / 20.f);
    f[7] = __log2f(1.f + __fdividef(dE, dV)) * (1.f / 20.f);
    f[8] = __log2f(dV) * (1.f / 32.f);
    f[9] = level < 64 ? (float)level * (1.f / 64.f) : 1.f;
    f[10] = prev_mode >= BFS_PULL_THREAD ? 1.f : 0.f;
    f[11] = (l_n - __log2f(1.f + (float)prev_n)) * (1.f / 16.f);
    f[12] = __fdividef(fd, 1.f + ue);
    f[13] = __fdividef(nn, 1.f + u);
    f[14] = __log2f(1.f + u) * (1.f / 32.f);
    f[15] = 1.f;
}

// BF16 weight k (unrolled loops make k a constant) of a row held in registers.
__device__ __forceinline__ float bf16_in(const uint4 *q, int k) {
    const uint4 v = q[k >> 3];
    const int j = (k >> 1) & 3;
    const uint32_t word = j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    return __uint_as_float((k & 1) ? (word & 0xffff0000u) : (word << 16));
}

// |F_l| and the sum of its degrees, from the stripes the previous launch
// (or the seed) added them up in, count set `set` = (l - 1) % 3: lane i
// reads stripe i. Every lane of the calling wave gets the totals.
__device__ __forceinline__ void level_counts(const StepArgs &a, int32_t set, int32_t *n, unsigned long long *fdeg) {
    static_assert(kStripes == 64, "one stripe per lane");
    const int32_t line = set * kStripes + lane_id();
    int32_t c = a.cnt[line * kStripeI32];
    unsigned long long f = a.fdg[line * kStripeU64];
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        f += __shfl_xor(f, o);
    }
    *n = c;
    *fdeg = f;
}

// The non-empty chunks among this workgroup's (chunk blockIdx + j G at bit
// j) in one set of chunk flags.
__device__ __forceinline__ uint64_t wg_chunk_mask(const uint8_t *flags, int32_t chunks) {
    const int32_t j = lane_id();
    const int32_t ch = (int32_t)blockIdx.x + j * (int32_t)gridDim.x;
    return __ballot(ch < chunks && flags[(int32_t)blockIdx.x * 64 + j] != 0);
}

// The selector network's parameters, loaded at the start of a launch with
// everything else the decision reads (the engine's copy of the blob always
// holds a model or zeros, so the loads need not wait on the mode): lane u of
// wave w holds hidden unit u's second-layer weights over inputs [16 w, 16 w
// + 16); lane u of the deciding wave also its first-layer row, its row of w3 (the 8
// outputs it feeds), b1[u] and b2[u] (the low and high halves of b), and
// b3[u] for u < kSelMaxOut.
struct SelHead {
    uint4 w1[kSelInputs / 8];
    uint4 w3[1];
    uint32_t b;
    uint32_t b3;
};

static_assert(kStepBlock == 4 * 64 && kSelHidden == 64 && kSelInputs == 16, "the network's split over 4 waves");
__device__ __forceinline__ void load_sel_slice(const uint16_t *blob, uint4 *w2) {
    const int32_t u = lane_id(), wid = threadIdx.x >> 6;
    const uint4 *r2 = reinterpret_cast<const uint4 *>(blob + kSelOffW2 + u * kSelHidden + 16 * wid);
    w2[0] = r2[0];
    w2[1] = r2[1];
}

__device__ __forceinline__ void load_sel_head(const uint16_t *blob, SelHead *q) {
    const int32_t u = lane_id();
    const uint4 *r1 = reinterpret_cast<const uint4 *>(blob + kSelOffW1 + u * kSelInputs);
    q->w1[0] = r1[0];
    q->w1[1] = r1[1];
    q->w3[0] = *reinterpret_cast<const uint4 *>(blob + kSelOffW3 + u * kSelMaxOut);
    q->b = (uint32_t)blob[kSelOffB1 + u] | ((uint32_t)blob[kSelOffB2 + u] << 16);
    q->b3 = u < kSelMaxOut ? (uint32_t)blob[kSelOffB3 + u] : 0u;
}

__device__ __forceinline__ bool sel_learned(const SelectorView &sel) {
    return sel.mode == GI_SELECT_LEARNED && sel.w1 != nullptr;
}

// GraphIt's hybrid rule or one of its schedules (not the network).
__device__ __forceinline__ int32_t rule_mode(const SelectorView &sel, int32_t n, int32_t V) {
    if (sel.mode == GI_SELECT_S1) return sel.mode_s1;
    if (sel.mode == GI_SELECT_S2) return sel.mode_s2;
    return ((double)n <= (double)sel.threshold * (double)V) ? sel.mode_s1 : sel.mode_s2;
}

// The network's first layer (the deciding wave; lane u: hidden unit u). Every lane
// computes the whole input vector (the same values, so no lane waits on
// another's branch of the feature switch).
__device__ __forceinline__ float sel_layer1(const SelectorView &sel, const SelHead &q, int32_t n,
                                            unsigned long long fdeg, long long unv, long long unv_e, int32_t l,
                                            int32_t prev_mode, int32_t prev_n, int32_t V, int64_t E) {
    float feats[kSelInputs];
    bfs_features_f32(feats, n, fdeg, unv, unv_e, l, prev_mode, prev_n, V, E);
    float a1 = bf16_to_f32((uint16_t)(q.b & 0xffffu));
#pragma unroll
    for (int k = 0; k < kSelInputs; ++k) a1 += (feats[k] - sel.feat_mean[k]) * sel.feat_scale[k] * bf16_in(q.w1, k);
    return a1 > 0.f ? a1 : 0.f;
}

// Wave w's part of the second layer over inputs [16 w, 16 w + 16) (h1 from
// LDS, in input order).
__device__ __forceinline__ float sel_layer2_part(const float *h1, const uint4 *w2) {
    const int32_t wid = threadIdx.x >> 6;
    float p = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) p += h1[16 * wid + k] * bf16_in(&w2[0], k);
#pragma unroll
    for (int k = 0; k < 8; ++k) p += h1[16 * wid + 8 + k] * bf16_in(&w2[1], k);
    return p;
}

// The deciding wave: the second layer's sum (its parts in wave order) and
// the output layer; the fastest allowed mode. Every lane of the wave gets it.
__device__ __forceinline__ int32_t sel_output(const SelectorView &sel, const SelHead &q, const float *part) {
    static_assert(kSelMaxOut == 8, "one 16-byte row of output weights per hidden unit");
    const int32_t lane = lane_id();
    const float a2 = bf16_to_f32((uint16_t)(q.b >> 16)) + part[lane] + part[64 + lane] + part[128 + lane] +
                     part[192 + lane];
    const float h2 = a2 > 0.f ? a2 : 0.f;
    // output o = b3[o] + sum over the hidden units of h2 * w3[unit][o]; the
    // eight butterfly sums run side by side (w3 is zero past n_out)
    float t[kSelMaxOut];
#pragma unroll
    for (int o = 0; o < kSelMaxOut; ++o) t[o] = h2 * bf16_in(q.w3, o);
#pragma unroll
    for (int x = 32; x > 0; x >>= 1)
#pragma unroll
        for (int o = 0; o < kSelMaxOut; ++o) t[o] += __shfl_xor(t[o], x);
    int32_t mode = sel.mode_s1;
    float best = 3.4e38f;
#pragma unroll
    for (int o = 0; o < kSelMaxOut; ++o) {
        const float yo = t[o] + bf16_to_f32((uint16_t)__builtin_amdgcn_readlane((int32_t)q.b3, o));
        if (o < sel.n_out && ((sel.allowed >> o) & 1) && yo < best) {
            best = yo;
            mode = o;
        }
    }
    return mode;
}

// The end of a run: F_l is empty. One thread.
__device__ __forceinline__ void finish_run(const StepArgs &a, int32_t l, const BfsLevel &prev) {
    BfsState *st = a.st;
    st->levels = l;
    st->pull_levels = prev.pull_levels;
    st->done = 1;
    if (a.host_flag) {  // the host waits on the replay's event, then reads these
        a.host_flag[2] = l;
        a.host_flag[1] = l;
        a.host_flag[0] = 1;
    }
}

// One level, one launch. Every workgroup first decides step l itself, from
// what the previous launch left (the counts of F_l in the stripes, the
// record of level l - 1) -- the same inputs and the same code, hence the
// same mode in every workgroup -- so no workgroup waits on another and no
// decide launch sits between two steps. Everything the decision and the
// step's set-up read is loaded at once: the rotations' indices come from
// the phase ph = l mod 12, a launch argument, not from the state. The first
// workgroup writes the record of level l and clears what the next launch
// adds into (count set (l + 1) % 3, hub list (l + 2) % 3), each workgroup
// its flags of set (l + 1) % 4: each was last read by the launch before this
// one and is next written by the one after. An empty F_l ends the run. Graph replays run launch k of
// a block as level base + k (base kept in the state); eager launches pass
// their level with base 0.
// The decision of step l = base + k (every thread of the workgroup calls
// it; the result is uniform). One wave loads what the decision and the
// step's set-up read (each of these loads touches up to 64 lines: one wave
// per workgroup loading them keeps the launch's requests to the L2 down) and
// decides; the network's second layer is split over the four waves. With
// `step`, the first workgroup writes the level's record and clears the count
// set the next launch adds into, and each workgroup its flags of the set
// the next launch writes; otherwise (the traced levels' decision launch)
// thread 0 stores the mode as the one the traced step launch runs with. An
// empty F_l ends the run. `scratch`: 320 floats of LDS.
struct LevelDec {
    int32_t go, mode, n, l;
    unsigned long long fdeg;
    uint64_t fl_clear, fl_list;
};

__device__ __forceinline__ void level_decide(const StepArgs &a, int32_t k, int32_t ph, bool step, float *scratch,
                                             LevelDec *d) {
    __shared__ int32_t go_s, net_s, mode_s, n_s, l_s;
    __shared__ unsigned long long fdeg_s;
    __shared__ uint64_t mask_s[2];
    float *h1_s = scratch, *part_s = scratch + 64;
    BfsState *st = a.st;
    const int32_t p2 = ph & 1, p3 = ph % 3, p4 = ph & 3;
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    constexpr int32_t dw = 0;  // the deciding wave
    const SelectorView &sel = st->sel;
    uint4 w2[2];
    load_sel_slice(a.selw, w2);
    SelHead q;
    BfsLevel prev;
    int32_t n = 0, mode = 0, l = 0;
    unsigned long long fdeg = 0;
    long long unv = 0, unv_e = 0;
    if (wid == dw) {
        load_sel_head(a.selw, &q);
        const bool learned = sel_learned(sel);
        const int32_t done = st->done, base = step ? st->base : 0, forced = step ? st->force_mode : -1;
        prev = st->lv[p2 ^ 1];
        level_counts(a, (p3 + 2) % 3, &n, &fdeg);
        const uint64_t fl_clear = step ? wg_chunk_mask(flag_set(a, (p4 + 2) & 3), a.chunks) : 0ull;
        const uint64_t fl_list = step ? wg_chunk_mask(flag_set(a, (p4 + 3) & 3), a.chunks) : 0ull;
        l = base + k;
        int32_t go = 0, net = 0;
        if (!done && n == 0) {  // wave-uniform: F_l is empty
            if (blockIdx.x == 0 && lane == 0) finish_run(a, l, prev);
        } else if (!done) {
            go = 1;
            unv = prev.unvisited - n;
            unv_e = prev.unvisited_edges - (long long)fdeg;
            if (forced >= 0) {
                mode = forced;
            } else if (!learned) {
                mode = rule_mode(sel, n, a.V);
            } else {
                net = 1;
                h1_s[lane] = sel_layer1(sel, q, n, fdeg, unv, unv_e, l, prev.mode, prev.count, a.V, a.E);
            }
            if (step) {
                if (blockIdx.x == 0) {
                    static_assert(kStripes == 64, "one stripe per lane");
                    const int32_t line = ((p3 + 1) % kCountSets) * kStripes + lane;
                    a.cnt[line * kStripeI32] = 0;
                    a.fdg[line * kStripeU64] = 0;
                }
                if (lane < 4)  // this workgroup's flags of the set the next launch writes
                    reinterpret_cast<uint4 *>(flag_set(a, (p4 + 1) & 3) + (size_t)blockIdx.x * 64)[lane] =
                        make_uint4(0u, 0u, 0u, 0u);
            }
        }
        if (lane == 0) {
            go_s = go;
            net_s = net;
            mode_s = mode;
            n_s = n;
            l_s = l;
            fdeg_s = fdeg;
            mask_s[0] = fl_clear;
            mask_s[1] = fl_list;
        }
    }
    __syncthreads();
    d->go = go_s;
    if (!d->go) return;  // uniform
    if (net_s) {  // uniform: the network's second layer over the four waves, then its output in the deciding wave
        part_s[wid * 64 + lane] = sel_layer2_part(h1_s, w2);
        __syncthreads();
        if (wid == dw) {
            mode = sel_output(sel, q, part_s);
            if (lane == 0) mode_s = mode;
        }
        __syncthreads();
    }
    d->mode = mode_s;
    d->n = n_s;
    d->l = l_s;
    d->fdeg = fdeg_s;
    d->fl_clear = mask_s[0];
    d->fl_list = mask_s[1];
    if (wid == dw && lane == 0) {
        mode = d->mode;
        if (!step) {
            st->force_mode = mode;
        } else if (blockIdx.x == 0) {
            BfsLevel &cur = st->lv[p2];
            cur.level = l;
            cur.mode = mode;
            cur.count = n;
            cur.fdeg = fdeg;
            cur.unvisited = unv;
            cur.unvisited_edges = unv_e;
            cur.pull_levels = prev.pull_levels + (mode >= BFS_PULL_THREAD ? 1 : 0);
            if (l < kBfsMaxLevels) {
                st->hist_mode[l] = mode;
                st->hist_next[l] = n;
            }
            st->hub_n[(p3 + 2) % 3] = 0;
            st->giant_n[(p3 + 2) % 3] = 0;
        }
    }
}

__device__ __forceinline__ void level_body(const StepArgs &a, int32_t k, int32_t ph) {
    __shared__ int32_t list[kChunkVerts];  // push: a chunk's vertices; first, the network's activations
    static_assert(sizeof(int32_t) * kChunkVerts >= sizeof(float) * 5 * 64, "the activations fit");
    LevelDec d;
    level_decide(a, k, ph, true, reinterpret_cast<float *>(list), &d);
    if (!d.go) return;  // uniform: the run has ended
    if (d.mode >= BFS_PULL_THREAD)  // uniform
        step_body<kPullStep>(a, list, d.l, ph, d.mode, d.n, d.fdeg, d.fl_clear, d.fl_list);
    else
        step_body<kPushStep>(a, list, d.l, ph, d.mode, d.n, d.fdeg, d.fl_clear, d.fl_list);
}

// The recorded level loop launches bfs_step, block launch k as level
// base + k with phase (base + k) mod 12. Traced levels (per-level timing)
// and sweeps launch the same code with their level and base 0, traced
// levels under two more symbols, one per direction, so a profile tells
// their launches apart. Both step bodies together need 72 VGPRs: 7 waves
// per SIMD without spilling measured faster than 8 with a spill.
#define GI_STEP_KERNEL(name)                                                               \
    __global__ void __launch_bounds__(kStepBlock) __attribute__((amdgpu_waves_per_eu(7))) \
    name(StepArgs a, int32_t k, int32_t ph) {                                             \
        level_body(a, k, ph);                                                              \
    }
GI_STEP_KERNEL(bfs_step)
GI_STEP_KERNEL(bfs_push_step_traced)
GI_STEP_KERNEL(bfs_pull_step_traced)
#undef GI_STEP_KERNEL

// Traced levels: the decision of step l alone (the same code as the step
// launches'), stored as the forced mode the traced step launch then runs
// with, so the host knows the direction before it launches; an empty F_l
// ends the run.
__global__ void __launch_bounds__(kStepBlock) bfs_decide_kernel(StepArgs a, int32_t l) {
    __shared__ float scratch[5 * 64];
    LevelDec d;
    level_decide(a, l, l % 12, false, scratch, &d);
}

// Sweeps: the selector inputs of step l.
__global__ void __launch_bounds__(64) bfs_peek_kernel(StepArgs a, int32_t l, BfsPeek *out) {
    const BfsLevel prev = a.st->lv[(l + 1) & 1];
    int32_t n;
    unsigned long long fdeg;
    level_counts(a, (l + 2) % kCountSets, &n, &fdeg);
    if (threadIdx.x == 0) {
        out->n = n;
        out->fdeg = fdeg;
        out->unvisited = prev.unvisited - n;
        out->unvisited_edges = prev.unvisited_edges - (long long)fdeg;
        out->prev_mode = prev.mode;
        out->prev_n = prev.count;
        out->level = l;
    }
}

struct SeedArgs {
    SelectorView sel;
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
    int32_t source;         // >= 0: the run's one source, seeded here (F_0)
    int32_t step_grid;      // workgroups of the step launches (the chunk flags' layout)
    const int32_t *off;     // out-edge offsets (the source's degree)
    int32_t *depth;
    int32_t depth_len;      // depth entries to clear (the rest is never read)
    int32_t *parent;
    uint64_t *visited;
    const uint64_t *dead;   // vertices without in-edges, and the padding past V
    uint64_t *bm[3];
    uint64_t *flags;        // every set, as 64-bit words
    int32_t flag_words;
    int32_t *cnt;
    unsigned long long *fdg;
    int32_t V;
    int32_t words;
};

// Clears every array of a run in one pass and initialises the state.
// A single source (r.source >= 0) is seeded in the same pass by the threads
// that clear its entries (depth 0, parent itself, its bit in visited and in
// F_0, its chunk flag in set 3, its count and degree in count set 2 (what
// step 0 reads), the first step's hub lists).
__global__ void bfs_reset_kernel(BfsState *st, SeedArgs sa, ResetArgs r, int64_t E) {
    const int32_t stride = gridDim.x * blockDim.x;
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t src = r.source;
    const int32_t sw = src >= 0 ? src >> 6 : -1;                // its bitmap word
    const uint64_t sbit = src >= 0 ? 1ull << (src & 63) : 0ull;
    // its chunk flag in set 3, as a 64-bit word of the flag array and a byte in it
    const int64_t sfb = src >= 0 ? 3LL * (r.flag_words / kFlagSets) * 8 + flag_pos(sw / kChunkWords, r.step_grid) : -1;
    // depths [0, depth_len) and parents [0, V): 16-byte stores where both
    // arrays are 16-byte aligned, then the remainder one by one
    const int32_t dlen = r.depth_len, plen = r.parent ? r.V : 0;
    const int32_t vlen = dlen > plen ? dlen : plen;
    const bool aligned = ((reinterpret_cast<uintptr_t>(r.depth) | reinterpret_cast<uintptr_t>(r.parent)) & 15) == 0;
    const int32_t head = aligned ? vlen / 4 * 4 : 0;
    const int4 none = make_int4(-1, -1, -1, -1);
    const int32_t sq = src >= 0 && src < head ? src >> 2 : -1;  // the source's 16 bytes
    for (int32_t q = tid; q < head / 4; q += stride) {
        const int32_t i = 4 * q;
        int4 dv = none, pv = none;
        if (q == sq) {  // the source: depth 0, parent itself
            const int32_t j = src & 3;
            dv.x = j == 0 ? 0 : -1;
            dv.y = j == 1 ? 0 : -1;
            dv.z = j == 2 ? 0 : -1;
            dv.w = j == 3 ? 0 : -1;
            pv.x = j == 0 ? src : -1;
            pv.y = j == 1 ? src : -1;
            pv.z = j == 2 ? src : -1;
            pv.w = j == 3 ? src : -1;
        }
        if (i + 4 <= dlen) reinterpret_cast<int4 *>(r.depth)[q] = dv;
        else
            for (int k = 0; k < 4; ++k)
                if (i + k < dlen) r.depth[i + k] = i + k == src ? 0 : -1;
        if (i + 4 <= plen) reinterpret_cast<int4 *>(r.parent)[q] = pv;
        else
            for (int k = 0; k < 4; ++k)
                if (i + k < plen) r.parent[i + k] = i + k == src ? src : -1;
    }
    for (int32_t i = head + tid; i < vlen; i += stride) {
        if (i < dlen) r.depth[i] = i == src ? 0 : -1;
        if (i < plen) r.parent[i] = i == src ? src : -1;
    }
    if (tid == 0 && src >= dlen && src < r.V) r.depth[src] = 0;  // a source past the cleared depths
    const int32_t lines = kCountSets * kStripes;
    int32_t n = r.words > r.flag_words ? r.words : r.flag_words;
    n = n > lines ? n : lines;
    for (int32_t i = tid; i < n; i += stride) {
        if (i < r.words) {
            const uint64_t b = i == sw ? sbit : 0ull;
            r.visited[i] = r.dead[i] | b;
            r.bm[0][i] = b;
            r.bm[1][i] = 0;
            r.bm[2][i] = 0;
        }
        if (i < r.flag_words) r.flags[i] = i == (int32_t)(sfb >> 3) ? 1ull << (8 * (sfb & 7)) : 0ull;
        if (i < lines) {
            const bool seed = src >= 0 && i == 2 * kStripes;  // count set 2 (level -1), first stripe
            r.cnt[i * kStripeI32] = seed ? 1 : 0;
            r.fdg[i * kStripeU64] = seed ? (unsigned long long)(r.off[src + 1] - r.off[src]) : 0ull;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int i = 0; i < 3; ++i) {
            st->hub_n[i] = 0;
            st->giant_n[i] = 0;
        }
        if (src >= 0 && sa.want_hubs) {  // the first step's grid bin
            const int32_t dv = r.off[src + 1] - r.off[src];
            if (dv >= sa.hub_min * kGiantFactor) {
                st->giants[0][0] = src;
                st->giant_n[0] = 1;
            } else if (dv >= sa.hub_min) {
                st->hubs[0][0] = src;
                st->hub_n[0] = 1;
            }
        }
        st->want_hubs = sa.want_hubs;
        st->hub_list_count = sa.hub_list_count;
        st->hub_list_fdeg = sa.hub_list_fdeg;
        st->done = 0;
        st->levels = 0;
        st->pull_levels = 0;
        st->force_mode = -1;
        st->base = 0;
        BfsLevel &before = st->lv[1];  // level -1: nothing reached yet
        before.level = -1;
        before.mode = BFS_PUSH_THREAD;
        before.count = 0;
        before.pull_levels = 0;
        before.fdeg = 0;
        before.unvisited = r.V;
        before.unvisited_edges = E;
        st->lane_max = sa.lane_max;
        st->hub_min = sa.hub_min;
        st->want_fdeg = sa.want_fdeg;
        st->depth = sa.depth;
        st->parent = sa.parent;
        st->sel = sa.sel;
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

// Seeds one or many sources into F_0 (a multi-source BFS gives every vertex
// its distance to the nearest source; over a disjoint union of graphs with
// one source each it is every graph's own BFS): the bitmap, the chunk flags
// step 0 reads (flags3 = flags[3]), the counts it reads (count set 2, first
// stripe), the hub list. Repeated sources count once.
__global__ void bfs_seed_sources_kernel(BfsState *st, const int32_t *src, int32_t n, int32_t single,
                                        uint64_t *bm0, uint8_t *flags3, uint64_t *visited, const int32_t *off,
                                        int32_t *cnt2, unsigned long long *fdg2, int32_t step_grid) {
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
        flag_chunk(flags3, v >> 6, step_grid);
        cnt += 1;
        const int32_t dv = off[v + 1] - off[v];
        fd += (unsigned long long)dv;
        if (st->want_hubs && dv >= st->hub_min * kGiantFactor) {  // the first step's grid bin
            const int32_t p = atomicAdd(&st->giant_n[0], 1);
            if (p < kMaxGiants) st->giants[0][p] = v;
        } else if (st->want_hubs && dv >= st->hub_min) {
            const int32_t p = atomicAdd(&st->hub_n[0], 1);
            if (p < kMaxHubs) st->hubs[0][p] = v;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o);
        fd += __shfl_xor(fd, o);
    }
    if (lane_id() == 0 && cnt) {
        atomicAdd(cnt2, cnt);
        atomicAdd(fdg2, fd);
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
    uint8_t *flags = nullptr;        // kFlagSets sets of chunk flags, flag_bytes apart
    int32_t flag_bytes = 0;
    int32_t *cnt = nullptr;          // level counts: kCountSets x kStripes lines
    unsigned long long *fdg = nullptr;
    BfsPeek *peek = nullptr;         // sweeps: a level's selector inputs
    int32_t *reached = nullptr;
    uint64_t *dead = nullptr;        // pull-side tables (see bfs_dead_kernel)
    int32_t *cand = nullptr;
    int32_t cand_first = 0;          // cand holds first in-neighbours (degree-ordered views)
    const int32_t *poff = nullptr;   // in-edge CSR (the graph itself when symmetric)
    const int32_t *pnbr = nullptr;
    int32_t *d_src = nullptr;        // sources of a multi-source run
    int32_t src_cap = 0;
    int32_t *host_flag = nullptr;    // pinned, coherent: {done, level, levels} written by the launch
                                     // that ends a run; [4] the next block's base level (copied in)
    hipEvent_t flag_ev = nullptr;    // recorded after each replay of a block of levels
    int32_t lane_max = 64;           // bin thresholds (gi_bfs_set_bins)
    int32_t hub_min = kHubDegree;
    int32_t blocks_per_cu = 8;
    // recorded level loops by 12 x level count + phase of the first level
    std::map<int32_t, hipGraphExec_t> execs;
    int32_t recent_levels[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // depths of the last runs (ring)
    int32_t recent_n = 0;
    uint16_t *selw = nullptr;                 // copy of the selector blob the step launches read
    unsigned long long sel_version = ~0ull;   // the model it was copied from
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
    dfree(w->flags);
    dfree(w->cnt);  // (fdg points into it)
    dfree(w->peek);
    dfree(w->selw);
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
    // a step workgroup owns at most 64 chunks (one bit each of a wave's mask)
    // and the grid has at least one workgroup per CU
    const int64_t max_chunks = 64LL * device_info().num_cus;
    if (chunks > max_chunks)
        return fail(GI_ERR_RANGE, "fused BFS supports up to %lld vertices on this GPU", (long long)max_chunks * kChunkVerts);
    const gi_graph *pg = g;  // in-edges for pull steps
    if (g->directed) {
        gi_graph *t = nullptr;
        GI_TRY(gi_graph_transpose(g, &t));
        pg = t;
    }
    FusedWork *w = new FusedWork();
    w->cand_first = g->degree_view ? 1 : 0;
    w->poff = pg->d_off;
    w->pnbr = pg->d_nbr;
    // chunk flags laid out per workgroup (flag_pos) for the largest step grid
    // (gi_bfs_set_bins allows up to 8 workgroups per CU)
    w->flag_bytes = 64 * 8 * device_info().num_cus;  // a multiple of 16
    int rc = dalloc(&w->st, 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->bm[i], (size_t)words + 1);
    if (rc == GI_OK) rc = dalloc(&w->flags, (size_t)kFlagSets * w->flag_bytes);
    // a stripe's count (int 0) and degree sum (ints 2-3) share its line
    static_assert(kStripeI32 == 2 * kStripeU64, "stripes of both counters are lines of the same size");
    if (rc == GI_OK) rc = dalloc(&w->cnt, (size_t)kCountSets * kStripes * kStripeI32);
    if (rc == GI_OK) w->fdg = reinterpret_cast<unsigned long long *>(w->cnt + 2);
    if (rc == GI_OK) rc = dalloc(&w->selw, (size_t)kSelBlob);
    if (rc == GI_OK) rc = dalloc(&w->peek, 1);
    if (rc == GI_OK) rc = dalloc(&w->reached, 1);
    if (rc == GI_OK) rc = dalloc(&w->dead, (size_t)words + 1);
    if (rc == GI_OK) rc = dalloc(&w->cand, (size_t)g->V + 1);
    if (rc == GI_OK && !g->d_visited) rc = dalloc(&g->d_visited, (size_t)words + 1);
    if (rc == GI_OK && hipHostMalloc((void **)&w->host_flag, sizeof(int32_t) * 8, hipHostMallocCoherent) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
    if (rc == GI_OK && hipEventCreateWithFlags(&w->flag_ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(GI_ERR_HIP, "event create failed");
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

// Every level, traced or not, runs on the same persistent grid.
static int step_grid(const FusedWork *w) {
    static const int resident = resident_per_cu((const void *)bfs_step, kStepBlock);
    return grid_of(w, resident);
}

static StepArgs step_args(gi_graph *g, FusedWork *w) {
    // (ensure_fused checked that a workgroup owns at most 64 chunks)
    StepArgs a;
    a.st = w->st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.poff = w->poff;
    a.pnbr = w->pnbr;
    a.cand = w->cand;
    a.cand_first = w->cand_first;
    a.V = g->V;
    a.E = g->E;
    a.words = (int32_t)bitmap_words(g->V);
    a.chunks = (int32_t)ceil_div(a.words, kChunkWords);
    a.depth = nullptr;   // read from the state inside the kernels
    a.parent = nullptr;
    a.visited = g->d_visited;
    for (int i = 0; i < 3; ++i) a.bm[i] = w->bm[i];
    a.flags = w->flags;
    a.flag_bytes = w->flag_bytes;
    a.cnt = w->cnt;
    a.fdg = w->fdg;
    a.host_flag = w->host_flag;
    a.selw = w->selw;
    return a;
}

// A traced level: the decision alone (the host reads the mode it forces),
// then the step code under its per-direction symbol between two events.
// Returns 1 when the run has ended.
static int launch_level_traced(gi_graph *g, FusedWork *w, const StepArgs &a, int32_t l, hipStream_t s,
                               FusedTrace *trace) {
    bfs_decide_kernel<<<1, kStepBlock, 0, s>>>(a, l);
    GI_LAUNCH_CHECK();
    static_assert(offsetof(BfsState, force_mode) == 3 * sizeof(int32_t), "done, levels, pull_levels, force_mode");
    int32_t h[4];
    GI_HIP(hipMemcpyAsync(h, w->st, sizeof(h), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    if (h[0]) return 1;
    hipEvent_t e0, e1;
    GI_HIP(hipEventCreate(&e0));
    GI_HIP(hipEventCreate(&e1));
    trace->ev.push_back(e0);
    trace->ev.push_back(e1);
    GI_HIP(hipEventRecord(e0, s));
    if (h[3] >= BFS_PULL_THREAD) bfs_pull_step_traced<<<step_grid(w), kStepBlock, 0, s>>>(a, l, l % 12);
    else bfs_push_step_traced<<<step_grid(w), kStepBlock, 0, s>>>(a, l, l % 12);
    GI_HIP(hipEventRecord(e1, s));
    GI_LAUNCH_CHECK();
    return GI_OK;
}

// Seed: clear the arrays, set the state, mark the sources (F_0).
// sources == null seeds the single vertex `single`.
static int seed_run(gi_graph *g, FusedWork *w, const int32_t *h_sources, int32_t n_sources,
                    int32_t single, int32_t *depth, int32_t *parent, const SelectorView &sel,
                    bool want_fdeg, hipStream_t s, int32_t depth_len = -1) {
    // hubs are listed at discovery when a TWC step may use the list: the
    // selector may pick TWC, or the degrees are read anyway
    const bool learned = sel.mode == GI_SELECT_LEARNED && sel.w1 != nullptr;
    const bool may_twc = learned ? ((sel.allowed >> BFS_PUSH_TWC) & 1) != 0
                                 : (sel.mode_s1 == BFS_PUSH_TWC || sel.mode_s2 == BFS_PUSH_TWC);
    if (learned && w->sel_version != sel.version) {  // the step launches read the engine's copy
        GI_HIP(hipMemcpyAsync(w->selw, sel.w1, sizeof(uint16_t) * kSelBlob, hipMemcpyDeviceToDevice, s));
        w->sel_version = sel.version;
    }
    // ... during push steps of at most V/64 vertices and E/64 edges
    SeedArgs sa{sel,
                depth,
                parent,
                w->lane_max,
                w->hub_min,
                want_fdeg ? 1 : 0,
                (want_fdeg || may_twc) ? 1 : 0,
                g->V / kHubListShare,
                (long long)(g->E / kHubListShare)};
    w->host_flag[0] = 0;  // no run of this graph's engine is in flight: set again by the final launch
    ResetArgs r;
    r.source = h_sources ? -1 : single;
    r.step_grid = step_grid(w);
    r.off = g->d_off;
    r.depth = depth;
    r.depth_len = depth_len >= 0 && depth_len < g->V ? depth_len : g->V;
    r.parent = parent;
    r.visited = g->d_visited;
    r.dead = w->dead;
    for (int i = 0; i < 3; ++i) r.bm[i] = w->bm[i];
    r.flags = reinterpret_cast<uint64_t *>(w->flags);
    r.flag_words = kFlagSets * w->flag_bytes / 8;
    r.cnt = w->cnt;
    r.fdg = w->fdg;
    r.V = g->V;
    r.words = (int32_t)bitmap_words(g->V);
    int grid = (int)ceil_div(g->V, 256);
    grid = grid < 4096 ? (grid < 1 ? 1 : grid) : 4096;
    bfs_reset_kernel<<<grid, 256, 0, s>>>(w->st, sa, r, g->E);
    int32_t *cnt2 = w->cnt + (size_t)2 * kStripes * kStripeI32;  // count set 2 (level -1), first stripe
    unsigned long long *fdg2 = w->fdg + (size_t)2 * kStripes * kStripeU64;
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
        bfs_seed_sources_kernel<<<sg < 1024 ? sg : 1024, 256, 0, s>>>(
            w->st, w->d_src, n_sources, 0, w->bm[0], w->flags + 3 * w->flag_bytes, g->d_visited, g->d_off, cnt2, fdg2,
            step_grid(w));
    }  // (a single source is seeded by the reset)
    GI_LAUNCH_CHECK();
    if (h_sources) GI_HIP(hipStreamSynchronize(s));  // h_sources may go away on return
    return GI_OK;
}

// A hipGraph of `levels` consecutive step launches, launch k running level
// base + k (base from the device state) with phase (ph0 + k) mod 12, ph0 =
// base mod 12, so one recording serves every run on this graph (the
// selector is in the device state) and every block of a run that starts at
// the same phase.
static int level_graph(gi_graph *g, FusedWork *w, int32_t levels, int32_t ph0, hipGraphExec_t *out) {
    const int32_t key = levels * 12 + ph0;
    auto it = w->execs.find(key);
    if (it != w->execs.end()) {
        *out = it->second;
        return GI_OK;
    }
    StepArgs a = step_args(g, w);
    hipGraphExec_t exec = nullptr;
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t k = 0; k < levels; ++k)
                L(bfs_step, step_grid(w), kStepBlock, a, k, (ph0 + k) % 12);
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
    if (trace && timed) {
        // per-level launches with events around each step; a BFS has at most
        // V levels: the bound only guards against a state that never
        // reports done
        for (int64_t level = 0; level <= (int64_t)g->V + 1; ++level) {
            const int r = launch_level_traced(g, w, a, (int32_t)level, s, trace);
            if (r == 1) {
                GI_HIP(hipMemcpy(levels, &w->st->levels, sizeof(int32_t), hipMemcpyDeviceToHost));
                return epilogue ? (*epilogue)(s) : GI_OK;
            }
            GI_TRY(r);
        }
        return fail(GI_ERR_HIP, "BFS state never reported done");
    }
    // Graph mode: replay a recorded block of step launches sized by the
    // deepest of the last runs on this graph (a launch past the end returns at
    // once, a block too short costs a whole extra replay), wait for each
    // replay's event (the epilogue is already queued behind it) and read the
    // {done, level, levels} the final launch wrote to pinned memory; continue
    // with blocks of 8, 16, 32, 48 levels for deeper runs, each after copying
    // its base level into the state.
    int32_t block = 0;
    for (int32_t l : w->recent_levels) block = l + 1 > block ? l + 1 : block;
    if (block <= 1) block = 8;
    if (block > 48) block = 48;
    int32_t next_block = 8;
    int64_t level = 0;
    while (level <= (int64_t)g->V + 1) {  // no depth limit: a BFS has at most V levels
        hipGraphExec_t exec = nullptr;
        GI_TRY(level_graph(g, w, block, (int32_t)(level % 12), &exec));
        GI_HIP(hipGraphLaunch(exec, s));
        level += block;
        // the launch that ends the run writes {done, level, levels} into the
        // pinned host_flag itself (the host cleared done before the run)
        GI_HIP(hipEventRecord(w->flag_ev, s));
        if (epilogue) GI_TRY((*epilogue)(s));
        GI_HIP(hipEventSynchronize(w->flag_ev));
        if (w->host_flag[0]) {
            *levels = w->host_flag[2];
            return GI_OK;
        }
        if (level > (int64_t)INT32_MAX - 64) break;
        w->host_flag[4] = (int32_t)level;  // the next block's first level
        GI_HIP(hipMemcpyAsync(&w->st->base, &w->host_flag[4], sizeof(int32_t), hipMemcpyHostToDevice, s));
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
// Before each timed replay of step l: clear the counts it adds into and its
// hub list (every replay lists the hubs of F' afresh), and force its mode.
__global__ void __launch_bounds__(64) bfs_sweep_prep_kernel(StepArgs a, int32_t l, int32_t mode) {
    const int32_t line = (l % kCountSets) * kStripes + (int32_t)threadIdx.x;
    a.cnt[line * kStripeI32] = 0;
    a.fdg[line * kStripeU64] = 0;
    if (threadIdx.x == 0) {
        a.st->force_mode = mode;
        a.st->hub_n[(l + 1) % 3] = 0;
        a.st->giant_n[(l + 1) % 3] = 0;
    }
}

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
    int32_t nl = 0;
    // One level: snapshot the state, time every mode of the mask from it with
    // the production step kernel (bfs_step, its mode forced), then advance by
    // the followed mode. Returns 1 when the BFS is done; errors return
    // through rc so the buffers below are freed.
    auto one_level = [&]() -> int {
        const int32_t l = nl;  // level l: F_l from the previous level's final run
        BfsPeek h;
        bfs_peek_kernel<<<1, 64, 0, s>>>(a, l, w->peek);
        GI_LAUNCH_CHECK();
        GI_HIP(hipMemcpyAsync(&h, w->peek, sizeof(h), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        if (h.n == 0) return 1;
        bfs_raw_features(features + (int64_t)nl * kSelInputs, h.n, h.fdeg, h.unvisited, h.unvisited_edges, l,
                         h.prev_mode, h.prev_n, g->V, g->E);
        const int32_t iout = (l + 1) % 3;
        // the step clears bitmap (l + 2) % 3, already empty after the first run
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
                GI_TRY(fill_u64(reinterpret_cast<uint64_t *>(w->flags + (l % kFlagSets) * w->flag_bytes),
                                w->flag_bytes / 8, 0, s));
                bfs_sweep_prep_kernel<<<1, 64, 0, s>>>(a, l, run);
                GI_HIP(hipEventRecord(e0, s));
                bfs_step<<<step_grid(w), kStepBlock, 0, s>>>(a, l, l % 12);
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
            }
        }
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
