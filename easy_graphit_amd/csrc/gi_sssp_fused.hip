// Fused delta-stepping SSSP: the loop of apps/sssp.cpp:34-57 (dequeue the
// current priority window, relax its out-edges with writeMin, re-enqueue
// improved vertices) with the window bookkeeping on the device and the
// phases recorded as hipGraphs. Shortest-path distances are unique, so the
// order of relaxations does not change the result.
//
// The window is [lo, hi). Every vertex below lo is settled; the near
// frontier F holds the vertices improved into the window since their last
// expansion (three rotating bitmaps with chunk flags, as in the fused BFS).
// Vertices improved to a distance >= hi are pending: they are not kept in a
// set of their own (the reference's priority queue, gpu_priority_queue.h),
// because a vertex is pending exactly when hi <= dist < INT_MAX. Instead
// every unit of 1024 vertices keeps lower bounds of its pending distances
// (lowered with atomicMin when a relaxation leaves a vertex pending; most
// relaxations find the bound already lower and skip the atomic). A window
// scan reads dist only in the units whose bound falls inside the window,
// and rewrites their bounds exactly; the other units are skipped, so later
// windows read a small part of dist instead of all of it.
// Light and heavy edges (the delta-stepping split of Meyer and Sanders):
// each out-list is copied with its light edges (w < delta) first. An edge
// from the window [lo, hi) that is heavy lands at or beyond hi, so it can
// never put a vertex back into the window. A push step whose F has at least
// E / kSsDeferDiv out-edges relaxes light edges only and records F in the
// window set W; when the window is settled, one heavy step relaxes the heavy
// edges of W once, from final distances, instead of once per improvement of
// each vertex within the window. Steps with smaller F relax all edges (no
// extra step for windows that settle in one step).
// One launch per step (plus, on graphs with vertices of at least kSsHubMin
// edges, the step's grid-wide hub bin), recorded as hipGraph blocks. Every
// workgroup of a step decides the step itself, from the previous step's
// record and the counts it left in striped counters (the same inputs, hence
// the same decision everywhere; no decide launch):
//   relax      push: relax the out-edges of F (each workgroup over its own
//              chunks of F), spread evenly over the workgroup (vertices of at
//              least kSsHubMin edges go to the hub bin); an improvement below
//              hi joins F'; or pull, for a large F: each unsettled vertex
//              takes the minimum over its in-neighbours in F (no atomics on
//              distances);
//   heavy      the heavy edges of W, once per window: the same push, or,
//              on a symmetric graph whose W holds E / kSsHeavyPullDiv edges
//              or more, the pull over each list's heavy part;
//   scan       the next window [lo, hi): scan dist in the units whose
//              bound is below hi and relax the out-edges of the vertices in
//              [lo, hi) at once; also the least pending distance >= hi;
//   done       nothing pending.
// After a relax step with an empty F' the window is settled: heavy step if W
// holds vertices, else a scan of [hi, hi + delta). A scan that moved nothing
// is followed by a scan at the least pending distance
// (gpu_priority_queue.h:126-146 starts a window at the minimum remaining
// priority too). A window is thus its scan-and-relax launch plus the relax
// steps its improvements need.
#include <algorithm>
#include <climits>
#include <map>
#include <vector>

#include "gi_bitset.h"
#include "gi_lb.h"

namespace gi {

constexpr int kSsBlock = 256;
constexpr int32_t kSsHubMin = 4096;   // at least this wide: grid-wide bin
constexpr int32_t kSsMaxHubs = 1024;
constexpr int32_t kSsMaxGrid = 8192;  // workgroups of a step launch at most
constexpr int32_t kSsMaxBlock = 2048; // step launches of a recorded block at most
constexpr int32_t kSsPullDiv = 8;     // a step pulls when |F| > V / kSsPullDiv

constexpr int kSsPullUnit = 4 * (kSsBlock / 64);  // words per unit (pull work, window scan, pending bounds)

// F with E / kSsDeferDiv out-edges or more: heavy edges wait (RMAT-24,
// delta 16, four sources: E/16 5.49 ms, E/64 5.12, E/128 to E/4096 4.78 to
// 4.81; the grid's windows stay below all of these)
constexpr long long kSsDeferDiv = 256;
constexpr long long kSsHeavyPullDiv = 8;  // W with E / kSsHeavyPullDiv out-edges or more: the heavy step pulls

// A step's counts (|F'| or the vertices a scan moved, their out-edges, the
// least pending distance a scan saw) are added up in kSsStripes stripes,
// each on a 128-byte line of its own (holding all three), one atomic per
// workgroup and counter, in three rotating sets: read by the next step,
// written, zeroed.
constexpr int kSsStripes = 64;
constexpr int kSsLineI32 = 32;
constexpr int kSsSets = 3;
constexpr int kSsFlagSets = 4;

enum SsMode : int32_t { SS_RELAX = 0, SS_ADVANCE = 1, SS_HEAVY = 2, SS_DONE = 3 };

// One step's bookkeeping, by step parity: written by the first workgroup of
// the step's launch, read by the next launch (and by the step's hub bin).
struct SsRec {
    int32_t mode;
    int32_t lo;          // window [lo, hi)
    int32_t hi;
    int32_t pull;        // the step pulls (a large F, or a heavy step over a large W)
    int32_t defer;       // the push step relaxes light edges only and records F in W
    int32_t ws_any;      // W holds vertices whose heavy edges wait
    int32_t front;       // |F|
    int32_t steps;       // relax steps up to and including this one
    int32_t windows;     // windows entered after the first
    int32_t pulls;       // relax steps that pulled
    long long fdeg;      // out-edges of F
    long long ws_fdeg;   // out-edges of the F's recorded in W this window (W's size in edges, at most)
    int32_t uq;          // scans so far mod 2: the pending bounds this step lowers
    int32_t local;       // the step's workgroups relax their improvements below hi themselves
};

struct SsState {
    int32_t done;
    int32_t slots;       // step launches of the run, when done
    int32_t steps;       // relax steps, windows, pull steps, when done
    int32_t windows;
    int32_t pulls;
    int32_t base;        // step of the first launch of a replayed block
    int32_t delta;
    int32_t pull_min;    // |F| above which a step pulls
    int32_t *dist;       // the run's distances (recorded launches read it from here)
    SsRec rec[2];        // rec[1] holds step -1 after the reset
    int32_t hub_n[3];    // hubs listed by step s in hubs[s % 3]
    int32_t hubs[3][kSsMaxHubs];
};

struct SsArgs {
    SsState *st;
    const int32_t *off;
    const int32_t *nbr;   // out-lists
    const int32_t *wt;
    const int32_t *snbr;  // the same lists, light edges first (light or heavy steps)
    const int32_t *swt;
    const int32_t *lend;  // end of each light part (null: no light/heavy split)
    long long defer_min;  // out-edges of F from which its heavy edges wait
    long long heavy_pull_min;  // W of this many out-edges or more: the heavy step pulls (0: never)
    const int32_t *poff;  // in-edges with their weights (the graph itself when symmetric)
    const int32_t *pnbr;
    const int32_t *pwt;
    int32_t *dist;
    int32_t V;
    int32_t words;
    uint64_t *bm[3];      // F_s in bm[s % 3]
    uint8_t *flags;       // kSsFlagSets sets of chunk flags, flag_bytes apart: F_s's in set (s - 1) % 4
    int32_t flag_bytes;
    // Per unit of kSsPullUnit words, lower bounds of its pending distances
    // (dist >= hi): the least of ubase[u] (written by the window scans) and
    // ucl[q * ustride + u] for q = 0, 1 (lowered by the relaxations of the
    // steps from a scan with q = scans mod 2 up to the next scan, which
    // resets it for the units it scans: it read what those steps wrote).
    int32_t *ubase;
    int32_t *ucl;
    int32_t ustride;
    uint64_t *ws;         // W: vertices of the window whose heavy edges wait
    uint8_t *wsflag;      // its chunk flags
    int32_t chunks;
    int32_t *cnt;         // [kSsSets][kSsStripes] lines: |F'| (or moved)
    long long *fdg;       // the same, out-edges
    int32_t *mn;          // the same, least pending distance seen by a scan (INT_MAX: none; -1: it found vertices)
    int32_t *host_flag;   // pinned {done, slots, steps, windows, pulls}: written when the run ends
    int32_t wide_chunks;       // pull: chunks of the in-lists of wide vertices
    const int32_t *chunk_v;    // vertex of each chunk
    const int32_t *chunk_e0;   // first in-edge of each chunk
};

__device__ __forceinline__ uint8_t *ss_flag_set(const SsArgs &a, int32_t i) {
    return a.flags + (size_t)i * a.flag_bytes;
}

// Workgroup b of a launch of G workgroups owns the chunks b + j G (j < 64)
// and the units b + j G: their flags and pending bounds are stored at b * 64
// + j (units in groups of 64 G), so that a workgroup reads its 64 with one
// or two cache lines.
__device__ __forceinline__ int32_t ss_fpos(int32_t chunk, int32_t G) { return (chunk % G) * 64 + chunk / G; }
__device__ __forceinline__ int32_t ss_upos(int32_t unit, int32_t G) {
    const int32_t j = unit / G;
    return (j >> 6) * 64 * G + (unit % G) * 64 + (j & 63);
}
__device__ __forceinline__ int32_t ss_unit_of(int32_t v) { return v / (64 * kSsPullUnit); }

// The flag of the chunk holding bitmap word `word`, in a step launch.
__device__ __forceinline__ void ss_flag_chunk(uint8_t *flags, int32_t word) {
    flags[ss_fpos(word / kBsChunkWords, (int32_t)gridDim.x)] = 1;
}

// The non-empty chunks among this workgroup's (chunk blockIdx + j G at bit
// j) in one set of chunk flags.
__device__ __forceinline__ uint64_t ss_chunk_mask(const uint8_t *flags, int32_t chunks) {
    const int32_t j = lane_id();
    const int32_t ch = (int32_t)blockIdx.x + j * (int32_t)gridDim.x;
    return __ballot(ch < chunks && flags[(int32_t)blockIdx.x * 64 + j] != 0);
}

// Improvements below hi that the workgroup relaxes itself, in the same
// launch (a light-edge chain inside a small window would otherwise cost a
// launch per link): a queue in LDS of kSsQCap entries {vertex, distance,
// first edge, end}, for vertices of at most kSsLocalDeg out-edges. A full
// queue, or what is left after kSsDrainRounds rounds, joins F' instead.
constexpr int32_t kSsQCap = 256;
constexpr int32_t kSsLocalDeg = 32;
constexpr int kSsDrainRounds = 8;
constexpr int32_t kSsLocalMaxF = 16384;  // relax steps with |F| up to this ...
constexpr long long kSsLocalMaxE = 262144;  // ... and this many out-edges drain their improvements
constexpr int32_t kSsLocalScan = 256;    // a scan drains in the workgroups that found this many at most

struct SsCtx {
    int32_t hi;
    uint64_t *bm_out;
    uint8_t *cflag;  // chunk flags of bm_out
    int32_t *ucl;    // this step's pending bounds
    int32_t nf;      // per-lane: new members of F'
    int32_t fdeg;    // per-lane: their out-edges
    bool local;      // workgroup-uniform: improvements below hi go to the queue q
    int32_t *q;      // LDS: 4 * kSsQCap entries (vertex, distance, first edge, end)
    int32_t *qn;     // LDS: entries claimed in q (may exceed kSsQCap: those joined F')
};

__device__ __forceinline__ SsCtx ss_ctx(const SsArgs &a, const SsRec &cur, int32_t ph) {
    SsCtx c;
    c.hi = cur.hi;
    c.bm_out = bs_pick3(a.bm[0], a.bm[1], a.bm[2], (ph + 1) % 3);
    c.cflag = ss_flag_set(a, ph & 3);
    c.ucl = a.ucl + cur.uq * a.ustride;
    c.nf = 0;
    c.fdeg = 0;
    c.local = false;
    c.q = nullptr;
    c.qn = nullptr;
    return c;
}

// Workgroup totals of the lane counters into the step's stripes (set `set`):
// one atomic per workgroup and counter.
__device__ __forceinline__ void ss_flush(const SsArgs &a, int32_t set, int32_t nf, long long fdeg, int32_t mn) {
    __shared__ int32_t rn[kSsBlock / 64], rm[kSsBlock / 64];
    __shared__ long long rf[kSsBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        nf += __shfl_xor(nf, o);
        fdeg += __shfl_xor(fdeg, o);
        const int32_t y = __shfl_xor(mn, o);
        mn = y < mn ? y : mn;
    }
    if (lane_id() == 0) {
        rn[threadIdx.x >> 6] = nf;
        rf[threadIdx.x >> 6] = fdeg;
        rm[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tn = 0, tm = INT_MAX;
        long long tf = 0;
        for (int i = 0; i < kSsBlock / 64; ++i) {
            tn += rn[i];
            tf += rf[i];
            tm = rm[i] < tm ? rm[i] : tm;
        }
        const int32_t line = set * kSsStripes + (int32_t)(blockIdx.x % kSsStripes);
        if (tn) atomicAdd(&a.cnt[line * kSsLineI32], tn);
        if (tf) atomicAdd((unsigned long long *)&a.fdg[line * (kSsLineI32 / 2)], (unsigned long long)tf);
        if (tm < INT_MAX) atomicMin(&a.mn[line * kSsLineI32], tm);
    }
}

// v joins F' (the bitmap of the next step), counted once.
__device__ __forceinline__ void ss_join(SsCtx &c, int32_t v, int32_t deg) {
    const uint64_t bit = 1ull << (v & 63);
    const unsigned long long o = atomicOr((unsigned long long *)&c.bm_out[v >> 6], bit);
    if (!(o & bit)) {
        c.nf += 1;
        c.fdeg += deg;
        if (o == 0) ss_flag_chunk(c.cflag, v >> 6);
    }
}

// writeMin (apps/sssp.cpp:24-27) of dist[v] to nd, known to be below the
// value last read in this step. Below hi the vertex is relaxed again, by
// this workgroup from its queue or by the next step from F', whether or not
// this atomic is the one that lowers dist (a lower value written meanwhile
// came from a relaxation that schedules the vertex too), so the atomic's
// result is not waited for, and the bitmap atomic and the degree loads are
// in flight with it. At or above hi the vertex is pending: its unit's bound
// for this step's parity is kept at or below nd.
// With `known`, e0 and e1 are v's edge range, loaded with dist[v].
__device__ __forceinline__ void ss_claim(const SsArgs &a, SsCtx &c, int32_t v, int32_t nd, bool known = false,
                                         int32_t e0 = 0, int32_t e1 = 0) {
    atomicMin(&a.dist[v], nd);
    if (nd >= c.hi) {
        int32_t *m = &c.ucl[ss_upos(ss_unit_of(v), (int32_t)gridDim.x)];
        if (nd < *m) atomicMin(m, nd);
        return;
    }
    if (!known) {
        e0 = a.off[v];
        e1 = a.off[v + 1];
    }
    if (c.local && e1 - e0 <= kSsLocalDeg) {
        const int32_t p = atomicAdd(c.qn, 1);
        if (p < kSsQCap) {
            c.q[p] = v;
            c.q[kSsQCap + p] = nd;
            c.q[2 * kSsQCap + p] = e0;
            c.q[3 * kSsQCap + p] = e1;
            return;
        }
    }
    ss_join(c, v, e1 - e0);
}

// Relaxation of edge e (of lists nbr, wt) from a vertex at distance du.
__device__ __forceinline__ void ss_relax(const SsArgs &a, SsCtx &c, const int32_t *nbr, const int32_t *wt,
                                         int32_t du, int32_t e, bool active) {
    if (!active) return;
    const int32_t v = nbr[e];
    const int32_t nd = du + wt[e];
    if (nd >= a.dist[v]) return;  // test before the atomic
    ss_claim(a, c, v, nd);
}

constexpr int kSsBatch = 4;  // edges per thread with their loads in flight together

// Relaxes the out-edges of n listed vertices, in rounds of kSsBlock
// vertices whose edges are spread evenly over the workgroup, each thread
// finding its edge's owner by binary search in LDS; a vertex with at least
// kSsHubMin edges in its range goes to the step's hub list (hubs[hl]) for
// the grid-wide bin instead. get(j, s, e0, e1, du) gives entry j: vertex,
// edge range [e0, e1) of (nbr, wt), distance. pre_s, e0_s, du_s hold
// kSsBlock entries of LDS. Every thread of the workgroup calls it.
template <class Get>
__device__ __forceinline__ void ss_rounds(const SsArgs &a, SsCtx &c, int32_t n, const int32_t *nbr,
                                          const int32_t *wt, int32_t hl, int32_t *pre_s, int32_t *e0_s,
                                          int32_t *du_s, Get get) {
    SsState *st = a.st;
    for (int32_t b0 = 0; b0 < n; b0 += kSsBlock) {
        const int32_t j = b0 + (int32_t)threadIdx.x;
        const bool has = j < n;
        int32_t s = 0, e0 = 0, e1 = 0, du = 0;
        if (has) get(j, s, e0, e1, du);
        const int32_t dg = e1 - e0;
        const bool hub = has && dg >= kSsHubMin;
        const int32_t hpos = wave_atomic_add(&st->hub_n[hl], hub);
        if (hub && hpos < kSsMaxHubs) st->hubs[hl][hpos] = s;
        const bool mine = has && (!hub || hpos >= kSsMaxHubs);
        int32_t total = 0;
        const int32_t ex = block_exclusive_scan(mine ? dg : 0, &total);
        pre_s[threadIdx.x] = ex;
        e0_s[threadIdx.x] = e0;
        du_s[threadIdx.x] = du;
        __syncthreads();
        for (int32_t k0 = 0; k0 < total; k0 += kSsBlock * kSsBatch) {
            int32_t v[kSsBatch], nd[kSsBatch], dv[kSsBatch];
#pragma unroll
            for (int p = 0; p < kSsBatch; ++p) {
                const int32_t k = k0 + p * kSsBlock + (int32_t)threadIdx.x;
                v[p] = -1;
                nd[p] = 0;
                if (k < total) {  // owner: the last entry whose prefix is <= k
                    const int32_t o = upper_bound_i32(pre_s, 0, kSsBlock, k) - 1;
                    const int32_t e = e0_s[o] + (k - pre_s[o]);
                    v[p] = nbr[e];
                    nd[p] = du_s[o] + wt[e];
                }
            }
            if (c.local) {  // uniform: few edges; the targets' edge ranges are loaded with their distances
                int32_t o0[kSsBatch], o1[kSsBatch];
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p) {
                    const bool ok = v[p] >= 0;
                    dv[p] = ok ? a.dist[v[p]] : INT_MIN;
                    o0[p] = ok ? a.off[v[p]] : 0;
                    o1[p] = ok ? a.off[v[p] + 1] : 0;
                }
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p)
                    if (nd[p] < dv[p]) ss_claim(a, c, v[p], nd[p], true, o0[p], o1[p]);
            } else {
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p) dv[p] = v[p] >= 0 ? a.dist[v[p]] : INT_MIN;
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p)
                    if (nd[p] < dv[p]) ss_claim(a, c, v[p], nd[p]);
            }
        }
        __syncthreads();
    }
}

// The workgroup relaxes the entries of its queue (c.q with its count in
// qn[0], filled by its claims in this step) in rounds, each round's improvements queued in the other
// buffer (q1, kSsQCap * 4 entries of LDS not otherwise in use by now), all
// edges of each vertex (a queued vertex has at most kSsLocalDeg). After
// kSsDrainRounds rounds what is still queued joins F'. Every thread of the
// workgroup calls it, with c.local set.
__device__ __forceinline__ void ss_drain(const SsArgs &a, SsCtx &c, int32_t hl, int32_t *q1, int32_t *qn,
                                         int32_t *pre_s, int32_t *e0_s, int32_t *du_s) {
    int32_t *cur = c.q, *nxt = q1;
    int32_t *ncur = qn, *nnxt = qn + 1;  // (c.qn is qn[0])
    for (int r = 0;; ++r) {
        __syncthreads();  // the claims that filled cur are done
        int32_t n = *ncur;
        n = n < kSsQCap ? n : kSsQCap;
        if (n == 0) break;  // uniform
        if (r == kSsDrainRounds) {  // what is left joins F'
            for (int32_t j = threadIdx.x; j < n; j += kSsBlock)
                ss_join(c, cur[j], cur[3 * kSsQCap + j] - cur[2 * kSsQCap + j]);
            break;
        }
        if (threadIdx.x == 0) *nnxt = 0;  // (nxt was consumed by the previous round)
        __syncthreads();
        c.q = nxt;
        c.qn = nnxt;
        ss_rounds(a, c, n, a.nbr, a.wt, hl, pre_s, e0_s, du_s,
                  [&](int32_t j, int32_t &v, int32_t &e0, int32_t &e1, int32_t &du) {
                      v = cur[j];
                      du = cur[kSsQCap + j];
                      e0 = cur[2 * kSsQCap + j];
                      e1 = cur[3 * kSsQCap + j];
                  });
        int32_t *t = cur;
        cur = nxt;
        nxt = t;
        int32_t *tn = ncur;
        ncur = nnxt;
        nnxt = tn;
    }
}

// Push relax step s (mode RELAX, or HEAVY: F is W and the edges are the
// heavy ones), over this workgroup's chunks of F (`fl_in`: their mask).
// `list` holds kBsChunkVerts entries (LDS).
__device__ __forceinline__ void ss_push_body(const SsArgs &a, const SsRec &cur, int32_t ph, uint64_t fl_in,
                                             int32_t *list, int32_t *pre_s, int32_t *e0_s, int32_t *du_s,
                                             int32_t *q0, int32_t *qn) {
    __shared__ int32_t n_s;
    const bool heavy = cur.mode == SS_HEAVY;  // F is W, edges from lend
    const bool defer = !heavy && cur.defer;   // edges up to lend; F joins W
    const int32_t *nbr = heavy || defer ? a.snbr : a.nbr;  // split lists only where the split is used
    const int32_t *wt = heavy || defer ? a.swt : a.wt;
    SsCtx c = ss_ctx(a, cur, ph);
    c.local = cur.local != 0;  // (a relax step of a small F: all edges, no hubs in the queue)
    c.q = q0;
    c.qn = qn;
    const uint64_t *bm_in = heavy ? a.ws : bs_pick3(a.bm[0], a.bm[1], a.bm[2], ph % 3);
    const int32_t lane = lane_id();
    for (uint64_t m = fl_in; m; m &= m - 1) {  // uniform
        const int32_t chunk = (int32_t)blockIdx.x + (__ffsll((unsigned long long)m) - 1) * (int32_t)gridDim.x;
        if (threadIdx.x < 64) {  // wave 0 lists the chunk's members
            const int32_t w = chunk * kBsChunkWords + (int32_t)threadIdx.x;
            uint64_t word = w < a.words ? bm_in[w] : 0ull;
            if (word) {  // this workgroup alone owns the chunk's words of W in this step
                if (heavy) {
                    a.ws[w] = 0;  // consumed
                } else if (defer) {
                    const uint64_t o = a.ws[w];
                    a.ws[w] = o | word;
                    if (!o) ss_flag_chunk(a.wsflag, w);
                }
            }
            const int32_t pc = __popcll(word);
            int32_t incl = pc;
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            int32_t pos = incl - pc;
            while (word) {
                const int32_t b = __ffsll((unsigned long long)word) - 1;
                word &= word - 1;
                list[pos++] = w * 64 + b;
            }
            if (threadIdx.x == 63) n_s = incl;
        }
        __syncthreads();
        const int32_t n = n_s;
        // W's chunk flag is consumed only now: every wave of the workgroup
        // took its ballot of the flags (at the start of the launch) before
        // this barrier
        if (heavy && threadIdx.x == 0) a.wsflag[ss_fpos(chunk, (int32_t)gridDim.x)] = 0;
        ss_rounds(a, c, n, nbr, wt, ph % 3, pre_s, e0_s, du_s,
                  [&](int32_t j, int32_t &s, int32_t &e0, int32_t &e1, int32_t &du) {
                      s = list[j];
                      e0 = heavy ? a.lend[s] : a.off[s];
                      e1 = defer ? a.lend[s] : a.off[s + 1];
                      du = a.dist[s];
                  });
        __syncthreads();
    }
    if (c.local) ss_drain(a, c, ph % 3, list, qn, pre_s, e0_s, du_s);
    ss_flush(a, ph % kSsSets, c.nf, c.fdeg, INT_MAX);
}

// Pull relax step for a large frontier: every vertex not yet settled below
// the window takes the minimum of dist[u] + w over its in-neighbours u in F
// (no atomics on dist: a vertex's word of 64 vertices belongs to one wave).
// Work is in units of 16 words per workgroup (4 per wave), so a unit's
// chunk flag costs at most one store.
constexpr int32_t kSsPullWide = 4096;            // in-degree at least this: chunked pull
constexpr int32_t kSsPullChunk = 16 * 1024;      // in-edges per chunk (a workgroup's item)
constexpr int kSsPasses = 4;                     // cooperative passes in flight per wave

// Pull relax step; `best_s` holds kSsBlock entries of LDS. The heavy step
// pulls the same way (symmetric graphs, whose in-lists are the split
// out-lists): F is W, the in-edges are each list's heavy part, and only
// vertices beyond hi can improve (a heavy edge from [lo, hi) lands at or
// beyond hi).
__device__ __forceinline__ void ss_pull_body(const SsArgs &a, const SsRec &cur, int32_t ph, int32_t *best_s) {
    const int32_t hi = cur.hi;
    const bool heavy = cur.mode == SS_HEAVY;
    const int32_t settled = heavy ? hi : cur.lo;  // vertices at or below this cannot improve
    const int32_t *pnbr = heavy ? a.snbr : a.pnbr;
    const int32_t *pwt = heavy ? a.swt : a.pwt;
    const uint64_t *bm_in = heavy ? a.ws : bs_pick3(a.bm[0], a.bm[1], a.bm[2], ph % 3);
    uint64_t *bm_out = bs_pick3(a.bm[0], a.bm[1], a.bm[2], (ph + 1) % 3);
    uint8_t *cflag = ss_flag_set(a, ph & 3);
    int32_t *ucl = a.ucl + cur.uq * a.ustride;  // this step's pending bounds
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    int32_t *best = best_s + 64 * wid;
    int32_t nf = 0, fdeg = 0;
    const int32_t units = (a.words + kSsPullUnit - 1) / kSsPullUnit;
    for (int32_t un = blockIdx.x; un < units; un += gridDim.x) {
        bool any = false;
        for (int32_t i = 0; i < 4; ++i) {
            const int32_t w = un * kSsPullUnit + wid * 4 + i;
            if (w >= a.words) break;  // wave-uniform
            const int32_t v = w * 64 + lane;
            const int32_t dv = v < a.V ? a.dist[v] : INT_MAX;
            bool open = v < a.V && dv > settled;
            const int32_t p0 = open ? a.poff[v] : 0, p1 = open ? a.poff[v + 1] : 0;
            if (p1 - p0 >= kSsPullWide) open = false;  // wide: the chunk items below handle it
            const int32_t e0 = open ? (heavy ? a.lend[v] : p0) : 0;
            const int32_t len = open ? p1 - e0 : 0;
            best[lane] = INT_MAX;
            int32_t incl = len;
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            const int32_t total = __shfl(incl, 63);
            const int32_t pre_l = incl - len;
            __builtin_amdgcn_wave_barrier();
            // kSsPasses passes of 64 edges at a time: their loads are in flight together
            for (int32_t base = 0; base < total; base += 64 * kSsPasses) {
                int32_t jj[kSsPasses], ee[kSsPasses], uu[kSsPasses], ww[kSsPasses];
#pragma unroll
                for (int p = 0; p < kSsPasses; ++p) {
                    const int32_t g = base + 64 * p + lane;
                    int32_t j = 0;  // owner: the last lane whose exclusive prefix is <= g
                    for (int32_t step = 32; step > 0; step >>= 1) {
                        const int32_t t = j + step;
                        const int32_t pt = __shfl(pre_l, t & 63);
                        if (t < 64 && pt <= g) j = t;
                    }
                    const int32_t ej = __shfl(e0, j), pj = __shfl(pre_l, j);
                    jj[p] = j;
                    ee[p] = g < total ? ej + (g - pj) : -1;
                }
#pragma unroll
                for (int p = 0; p < kSsPasses; ++p) {
                    uu[p] = ee[p] >= 0 ? pnbr[ee[p]] : -1;
                    ww[p] = ee[p] >= 0 ? pwt[ee[p]] : 0;
                }
#pragma unroll
                for (int p = 0; p < kSsPasses; ++p)
                    if (uu[p] >= 0 && bs_test(bm_in, uu[p])) atomicMin(&best[jj[p]], a.dist[uu[p]] + ww[p]);
                __builtin_amdgcn_wave_barrier();
            }
            const int32_t nd = best[lane];
            const bool imp = open && nd < dv;
            if (imp) a.dist[v] = nd;
            const bool join = imp && nd < hi;
            int32_t pend = imp && !join ? nd : INT_MAX;  // least pending result of the word
            for (int o = 32; o > 0; o >>= 1) {
                const int32_t y = __shfl_xor(pend, o);
                pend = y < pend ? y : pend;
            }
            int32_t *um = &ucl[ss_upos(un, (int32_t)gridDim.x)];
            if (lane == 0 && pend < *um) atomicMin(um, pend);
            const uint64_t near = __ballot(join);
            uint64_t old = 0;
            if (lane == 0 && near) {  // atomic: a wide vertex of the same word may join too
                old = atomicOr((unsigned long long *)&bm_out[w], near);
                nf += __popcll(near & ~old);
                any = true;
            }
            old = __shfl(old, 0);
            if (join && !((old >> lane) & 1ull)) fdeg += a.off[v + 1] - a.off[v];
            __builtin_amdgcn_wave_barrier();
        }
        if (__syncthreads_or(any) && threadIdx.x == 0) ss_flag_chunk(cflag, un * kSsPullUnit);
    }
    // wide vertices: each chunk of an in-list is a workgroup item; its minimum
    // joins dist[v] by atomicMin, and an improvement below hi joins F'
    const int32_t per_wave = kSsPullChunk / (kSsBlock / 64);
    for (int32_t k = blockIdx.x; k < a.wide_chunks; k += gridDim.x) {
        const int32_t v = a.chunk_v[k];
        const int32_t dv = a.dist[v];
        if (dv <= settled) continue;  // (a wave may see another workgroup's improvement: no barrier here)
        const int32_t end = a.poff[v + 1];
        int32_t e0 = a.chunk_e0[k] + wid * per_wave;
        const int32_t e1 = e0 + per_wave < end ? e0 + per_wave : end;
        if (heavy && e0 < a.lend[v]) e0 = a.lend[v];  // the heavy part of the list
        int32_t m = INT_MAX;
#pragma unroll 4
        for (int32_t e = e0 + lane; e < e1; e += 64) {
            const int32_t u = pnbr[e];
            if (bs_test(bm_in, u)) {
                const int32_t c = a.dist[u] + pwt[e];
                m = c < m ? c : m;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t y = __shfl_xor(m, o);
            m = y < m ? y : m;
        }
        if (lane == 0 && m < dv) {
            const int32_t old = atomicMin(&a.dist[v], m);
            if (m < old && m >= hi) {
                int32_t *um = &ucl[ss_upos(ss_unit_of(v), (int32_t)gridDim.x)];
                if (m < *um) atomicMin(um, m);
            }
            if (m < old && m < hi) {
                const uint64_t bit = 1ull << (v & 63);
                const unsigned long long o = atomicOr((unsigned long long *)&bm_out[v >> 6], bit);
                if (!(o & bit)) {
                    nf += 1;
                    fdeg += a.off[v + 1] - a.off[v];
                    if (o == 0) ss_flag_chunk(cflag, v >> 6);
                }
            }
        }
    }
    ss_flush(a, ph % kSsSets, nf, fdeg, INT_MAX);
}

// Window scan-and-relax (mode ADVANCE), over the units whose pending bound
// is below hi (a unit whose bound is at least hi holds nothing below hi and
// is skipped): each workgroup lists the vertices of those units with lo <=
// dist < hi and then relaxes all their out-edges, as a push step would in
// the next launch; improvements below hi join F_s+1 (or the workgroup's
// queue). Each wave scans whole units of its own (the r-th unit to scan
// goes to wave r mod 4), the 16 words of a unit with their loads in flight
// together, and lists what it finds through an LDS counter, with no barrier
// per unit; the edge ranges of the list are loaded afterwards, in one pass.
// The unit's base bound becomes its least distance >= hi and the bound
// lowered since the previous scan is reset (those relaxations ran in earlier
// launches, whose writes to dist this scan read); this step's own
// relaxations lower the other bound, which the scan leaves alone, so no
// bound is lost to a concurrent relaxation. The min stripes get -1 when the
// workgroup found vertices, else the least bound over its units (the least
// pending distance, exact when no vertex was found anywhere). lds holds
// kBsChunkVerts entries; pre_s, e0_s, du_s kSsBlock each; um_s the bounds
// of the workgroup's first 64 units, loaded at the start of the launch (the
// least of each unit's three bounds, whatever the parity).
__device__ __forceinline__ void ss_scan_body(const SsArgs &a, const SsRec &cur, int32_t ph, int32_t *lds,
                                             int32_t *pre_s, int32_t *e0_s, int32_t *du_s, int32_t *q0,
                                             int32_t *qn, int32_t *um_s) {
    __shared__ int32_t lcount;
    const int32_t lo = cur.lo, hi = cur.hi;
    SsCtx c = ss_ctx(a, cur, ph);
    int32_t *uold = a.ucl + (cur.uq ^ 1) * a.ustride;  // bounds lowered since the previous scan
    constexpr int32_t kCap = kBsChunkVerts / 4;          // found vertices listed before they are relaxed
    int32_t *lv = lds, *ldu = lds + kCap, *le0 = lds + 2 * kCap, *le1 = lds + 3 * kCap;
    static_assert(kSsPullUnit == 16, "a unit is 16 words: one wave loads it at once");
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    const int32_t units = (a.words + kSsPullUnit - 1) / kSsPullUnit;
    const int32_t G = (int32_t)gridDim.x;
    if (threadIdx.x == 0) lcount = 0;
    int32_t beyond = INT_MAX;  // per wave: least bound over its units and the units not scanned
    bool found = false;        // per wave
    bool spilled = false;      // per wave: a found vertex did not fit the list and joined F'
    for (int32_t g0 = 0; (int32_t)blockIdx.x + g0 * G < units; g0 += 64) {  // this workgroup's units, 64 at a time
        if (g0 > 0) {  // the bounds of the next 64 (the first 64 came with the launch)
            __syncthreads();  // every wave has read um_s
            if (threadIdx.x < 64) {
                const int32_t ul = (int32_t)blockIdx.x + (g0 + lane) * G;
                int32_t um = INT_MAX;
                if (ul < units) {
                    const int32_t up = ss_upos(ul, G);
                    const int32_t b0 = a.ubase[up], b1 = uold[up], b2 = c.ucl[up];
                    um = b0 < b1 ? b0 : b1;
                    um = b2 < um ? b2 : um;
                }
                um_s[lane] = um;
            }
        }
        __syncthreads();  // um_s is written (and lcount, the first time)
        const int32_t um = um_s[lane];
        const uint64_t todo = __ballot(um < hi);  // the same in every wave
        int32_t skip = um >= hi ? um : INT_MAX;   // the bounds of the units not scanned
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t y = __shfl_xor(skip, o);
            skip = y < skip ? y : skip;
        }
        beyond = skip < beyond ? skip : beyond;
        int32_t r = 0;
        for (uint64_t t = todo; t; t &= t - 1, ++r) {  // wave-uniform
            if ((r & 3) != wid) continue;
            const int32_t un = (int32_t)blockIdx.x + (g0 + __ffsll((unsigned long long)t) - 1) * G;
            const int32_t v0 = un * kSsPullUnit * 64 + lane;
            int32_t dv[kSsPullUnit];
#pragma unroll
            for (int i = 0; i < kSsPullUnit; ++i) {
                const int32_t v = v0 + 64 * i;
                dv[i] = v < a.V ? a.dist[v] : INT_MAX;
            }
            int32_t ub = INT_MAX;
#pragma unroll
            for (int i = 0; i < kSsPullUnit; ++i) {
                const bool in = dv[i] >= lo && dv[i] < hi;  // dv is INT_MAX past V
                if (dv[i] >= hi && dv[i] < ub) ub = dv[i];
                const uint64_t m = __ballot(in);
                if (m) {  // wave-uniform
                    found = true;
                    int32_t base = 0;
                    if (lane == 0) base = atomicAdd(&lcount, (int32_t)__popcll(m));
                    base = __shfl(base, 0);
                    if (in) {
                        const int32_t q = base + (int32_t)lanes_below(m);
                        const int32_t v = v0 + 64 * i;
                        if (q < kCap) {
                            lv[q] = v;
                            ldu[q] = dv[i];
                        } else {  // the list is full: the next step relaxes it from F'
                            ss_join(c, v, a.off[v + 1] - a.off[v]);
                            spilled = true;
                        }
                    }
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const int32_t y = __shfl_xor(ub, o);
                ub = y < ub ? y : ub;
            }
            if (lane == 0) {  // the unit's exact bound
                a.ubase[ss_upos(un, G)] = ub;
                uold[ss_upos(un, G)] = INT_MAX;
            }
            beyond = ub < beyond ? ub : beyond;
        }
    }
    __syncthreads();  // the list is complete
    int32_t listed = lcount;
    listed = listed < kCap ? listed : kCap;
    for (int32_t j = threadIdx.x; j < listed; j += kSsBlock) {  // the list's edge ranges, in one pass
        const int32_t v = lv[j];
        le0[j] = a.off[v];
        le1[j] = a.off[v + 1];
    }
    const bool any_spill = __syncthreads_or(spilled);
    if (listed > 0) {  // uniform
        // a workgroup that found few vertices relaxes what they improve
        // into the window itself, in this launch
        c.local = cur.local && !any_spill && listed <= kSsLocalScan;
        c.q = q0;
        c.qn = qn;
        ss_rounds(a, c, listed, a.nbr, a.wt, ph % 3, pre_s, e0_s, du_s,
                  [&](int32_t j, int32_t &v, int32_t &e0, int32_t &e1, int32_t &du) {
                      v = lv[j];
                      e0 = le0[j];
                      e1 = le1[j];
                      du = ldu[j];
                  });
        if (c.local) ss_drain(a, c, ph % 3, lds, qn, pre_s, e0_s, du_s);
    }
    ss_flush(a, ph % kSsSets, c.nf, c.fdeg, found ? -1 : beyond);
}

__device__ __forceinline__ int32_t ss_window_end(int32_t lo, int32_t delta) {
    return lo > INT_MAX - delta ? INT_MAX : lo + delta;
}

// The counts of step s - 1 from its stripes (set (s - 1) % 3); every lane of
// the calling wave gets the totals.
__device__ __forceinline__ void ss_counts(const SsArgs &a, int32_t set, int32_t *n, long long *fdeg, int32_t *mn) {
    static_assert(kSsStripes == 64, "one stripe per lane");
    const int32_t line = set * kSsStripes + lane_id();
    int32_t c = a.cnt[line * kSsLineI32];
    long long f = a.fdg[line * (kSsLineI32 / 2)];
    int32_t m = a.mn[line * kSsLineI32];
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        f += __shfl_xor(f, o);
        const int32_t y = __shfl_xor(m, o);
        m = y < m ? y : m;
    }
    *n = c;
    *fdeg = f;
    *mn = m;
}

// The decision of step s from step s - 1's record and counts (what the
// reference's queue does between two dequeues, gpu_priority_queue.h:115-185,
// with the light/heavy split): a relax, heavy or scan-and-relax step whose
// F' is not empty is followed by a relax step in the same window; a settled
// window with heavy edges waiting by the heavy step; otherwise by the
// scan-and-relax of the next window [hi, hi + delta). A scan that found no
// vertex is followed by a scan at the least pending distance it saw (the
// reference's queue starts a window at the least remaining priority too,
// gpu_priority_queue.h:126-146), or by the end when nothing is pending.
__device__ __forceinline__ SsRec ss_decide(const SsArgs &a, const SsRec &prev, int32_t n, long long fdeg,
                                           int32_t mn, int32_t delta, int32_t pull_min) {
    SsRec r = prev;
    if (prev.mode == SS_DONE) return r;
    auto relax_on = [&](int32_t steps_before) {
        r.mode = SS_RELAX;
        r.front = n;
        r.fdeg = fdeg;
        r.pull = n > pull_min;
        r.defer = a.lend != nullptr && !r.pull && fdeg >= a.defer_min;
        r.local = !r.pull && !r.defer && n <= kSsLocalMaxF && fdeg <= kSsLocalMaxE;
        r.steps = steps_before + 1;
        r.pulls = prev.pulls + r.pull;
    };
    auto scan_at = [&](int32_t lo) {
        r.mode = SS_ADVANCE;
        r.uq = prev.uq ^ 1;
        r.lo = lo;
        r.hi = ss_window_end(lo, delta);
        r.pull = 0;
        r.defer = 0;
        r.local = 1;  // (in the workgroups that find few vertices)
    };
    if (prev.mode == SS_RELAX || prev.mode == SS_HEAVY) {
        const bool deferred = prev.mode == SS_RELAX && prev.defer;
        const int32_t ws_any = prev.ws_any || deferred;
        r.ws_any = ws_any;
        r.ws_fdeg = ws_any ? prev.ws_fdeg + (deferred ? prev.fdeg : 0) : 0;
        if (n > 0) {
            relax_on(prev.steps);
        } else if (prev.mode == SS_RELAX && ws_any) {  // settled, heavy edges waiting
            r.mode = SS_HEAVY;
            // W's heavy edges pushed once each, or, when W holds a large part
            // of the graph's edges, pulled: every vertex beyond hi reads its
            // heavy in-edges in order (no random atomics on dist)
            r.pull = a.heavy_pull_min > 0 && r.ws_fdeg >= a.heavy_pull_min;
            r.ws_any = 0;
            r.ws_fdeg = 0;
            r.defer = 0;
            r.local = 0;
            r.steps = prev.steps + 1;
        } else {  // the window is settled
            scan_at(prev.hi);
        }
    } else {  // a scan-and-relax step; mn < 0: it found (and relaxed) vertices of its window
        const bool found = mn < 0;
        r.steps = prev.steps + (found ? 1 : 0);
        if (found) r.windows = prev.windows + 1;
        if (n > 0) {
            relax_on(r.steps);
        } else if (found) {  // the window settled in one step
            scan_at(prev.hi);
        } else if (mn == INT_MAX) {
            r.mode = SS_DONE;  // nothing pending
        } else {  // empty window: the next one starts at the least pending distance
            scan_at(mn);
        }
    }
    return r;
}

// One step, one launch: every workgroup decides the step from what the
// previous launch left (the same inputs and code, hence the same decision),
// the first writes the step's record and clears what the next launch adds
// into (its count set, the hub list of two steps before), each workgroup
// clears its flags of the set the next launch writes and its chunks of the
// bitmap two steps ahead, then the
// push, pull, heavy or scan body runs. Graph replays run launch k of a block
// as step base + k with phase (base + k) mod 12; eager launches pass their
// step with base 0.
__global__ void __launch_bounds__(kSsBlock) __attribute__((amdgpu_waves_per_eu(5))) ss_step_kernel(SsArgs a, int32_t k,
                                                                                           int32_t ph) {
    __shared__ int32_t lds[kBsChunkVerts];
    __shared__ int32_t pre_s[kSsBlock], e0_s[kSsBlock], du_s[kSsBlock];
    __shared__ int32_t q0[4 * kSsQCap], qn[2];
    static_assert(4 * kSsQCap <= kBsChunkVerts, "the second queue buffer fits the list space");
    __shared__ SsRec cur_s;
    __shared__ uint64_t mask_s[3];
    __shared__ int32_t um_s[64];
    __shared__ int32_t done_s, wsfree_s;
    __shared__ int32_t *dist_s;
    SsState *st = a.st;
    const int32_t p2 = ph & 1, p3 = ph % 3, p4 = ph & 3;
    // Wave 0 loads what the decision and the bodies need and decides; the
    // other waves take it from LDS (each of these loads touches 64 lines,
    // so one wave per workgroup loading them keeps the launch's requests to
    // the L2 down).
    if (threadIdx.x < 64) {
        const int32_t done = st->done;
        const SsRec prev = st->rec[p2 ^ 1];
        const int32_t delta = st->delta, pull_min = st->pull_min;
        int32_t *const dist = st->dist;
        int32_t n, mn;
        long long fdeg;
        ss_counts(a, (p3 + 2) % kSsSets, &n, &fdeg, &mn);
        // every chunk mask the step may use, before any wave of the
        // workgroup writes a flag
        const uint64_t fl_clear = ss_chunk_mask(ss_flag_set(a, (p4 + 2) & 3), a.chunks);
        const uint64_t fl_f = ss_chunk_mask(ss_flag_set(a, (p4 + 3) & 3), a.chunks);
        const uint64_t fl_ws = ss_chunk_mask(a.wsflag, a.chunks);
        // the bounds of this lane's first unit for a scan (a relaxation of
        // this launch may lower one, but only to a value at or above the
        // window's end, which leaves the scan's choice alone)
        int32_t um0 = INT_MAX;
        const int32_t ul = (int32_t)blockIdx.x + lane_id() * (int32_t)gridDim.x;
        if (ul < (a.words + kSsPullUnit - 1) / kSsPullUnit) {
            const int32_t up = ss_upos(ul, (int32_t)gridDim.x);
            const int32_t b0 = a.ubase[up], b1 = a.ucl[up], b2 = a.ucl[a.ustride + up];
            um0 = b0 < b1 ? b0 : b1;
            um0 = b2 < um0 ? b2 : um0;
        }
        um_s[lane_id()] = um0;
        if (threadIdx.x == 0) {
            cur_s = done ? prev : ss_decide(a, prev, n, fdeg, mn, delta, pull_min);
            done_s = done;
            wsfree_s = prev.mode == SS_HEAVY && prev.pull;
            dist_s = dist;
            mask_s[0] = fl_clear;
            mask_s[1] = fl_f;
            mask_s[2] = fl_ws;
            qn[0] = 0;
        }
    }
    __syncthreads();
    if (done_s) return;  // uniform: the run has ended
    const SsRec cur = cur_s;
    const uint64_t fl_clear = mask_s[0], fl_f = mask_s[1], fl_ws = mask_s[2];
    a.dist = dist_s;  // (the run's distances, from the state: one recording serves every run)
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            st->rec[p2] = cur;
            st->hub_n[(p3 + 2) % 3] = 0;
            if (cur.mode == SS_DONE) {
                st->slots = st->base + k;
                st->steps = cur.steps;
                st->windows = cur.windows;
                st->pulls = cur.pulls;
                st->done = 1;
                if (a.host_flag) {  // the host waits on the replay's event, then reads these
                    a.host_flag[1] = st->base + k;
                    a.host_flag[2] = cur.steps;
                    a.host_flag[3] = cur.windows;
                    a.host_flag[4] = cur.pulls;
                    a.host_flag[0] = 1;
                }
            }
        }
        const int32_t cs = (p3 + 1) % kSsSets;
        if (threadIdx.x < kSsStripes) {
            const int32_t line = cs * kSsStripes + (int32_t)threadIdx.x;
            a.cnt[line * kSsLineI32] = 0;
            a.fdg[line * (kSsLineI32 / 2)] = 0;
            a.mn[line * kSsLineI32] = INT_MAX;
        }
    }
    if (threadIdx.x < 4)  // this workgroup's 64 flags of the set the next launch writes
        reinterpret_cast<uint4 *>(ss_flag_set(a, (p4 + 1) & 3) + (size_t)blockIdx.x * 64)[threadIdx.x] =
            make_uint4(0u, 0u, 0u, 0u);
    if (cur.mode == SS_DONE) return;  // uniform
    // the bitmap two steps ahead (F_s-1, done with) is free: clear this
    // workgroup's non-empty chunks of it
    {
        uint64_t *bc = bs_pick3(a.bm[0], a.bm[1], a.bm[2], (p3 + 2) % 3);
        for (uint64_t m = fl_clear; m; m &= m - 1) {  // uniform
            const int32_t ch = (int32_t)blockIdx.x + (__ffsll((unsigned long long)m) - 1) * (int32_t)gridDim.x;
            if (threadIdx.x < kBsChunkWords) {
                const int32_t w = ch * kBsChunkWords + (int32_t)threadIdx.x;
                if (w < a.words) bc[w] = 0;
            }
        }
    }
    if (wsfree_s) {  // uniform: the step before pulled W in a heavy step, which is done with
        // (its bits were read by any workgroup, so the owners free it a step later)
        for (uint64_t m = fl_ws; m; m &= m - 1) {
            const int32_t ch = (int32_t)blockIdx.x + (__ffsll((unsigned long long)m) - 1) * (int32_t)gridDim.x;
            if (threadIdx.x < kBsChunkWords) {
                const int32_t w = ch * kBsChunkWords + (int32_t)threadIdx.x;
                if (w < a.words) a.ws[w] = 0;
            }
        }
        __syncthreads();  // (the flags were read at the start of the launch)
        if (threadIdx.x < 64 && ((fl_ws >> threadIdx.x) & 1ull)) a.wsflag[blockIdx.x * 64 + threadIdx.x] = 0;
    }
    if (cur.mode == SS_ADVANCE) {
        ss_scan_body(a, cur, ph, lds, pre_s, e0_s, du_s, q0, qn, um_s);
    } else if (cur.pull) {  // a pull relax step, or the heavy step pulled
        ss_pull_body(a, cur, ph, lds);
    } else {
        ss_push_body(a, cur, ph, cur.mode == SS_HEAVY ? fl_ws : fl_f, lds, pre_s, e0_s, du_s, q0, qn);
    }
}

// Grid-wide bin of step s (launched after it on graphs with hubs): the
// concatenated edges of the hubs its push step deferred, spread evenly.
__global__ void __launch_bounds__(kSsBlock) ss_hub_kernel(SsArgs a, int32_t k, int32_t ph) {
    (void)k;
    SsState *st = a.st;
    const SsRec cur = st->rec[ph & 1];
    if (cur.mode == SS_DONE || cur.pull) return;  // (a scan-and-relax lists hubs too)
    const int32_t hl = ph % 3;
    int32_t nh = st->hub_n[hl];
    if (st->done || nh <= 0) return;
    a.dist = st->dist;
    if (nh > kSsMaxHubs) nh = kSsMaxHubs;
    const bool heavy = cur.mode == SS_HEAVY, defer = !heavy && cur.defer;  // the step's edge range
    const int32_t *nbr = heavy || defer ? a.snbr : a.nbr;
    const int32_t *wt = heavy || defer ? a.swt : a.wt;
    SsCtx c = ss_ctx(a, cur, ph);
    __shared__ int32_t hscan[kSsMaxHubs + 1];
    __shared__ int32_t hs[kSsMaxHubs];  // first edge
    __shared__ int32_t hd[kSsMaxHubs];
    const int32_t per = (nh + kSsBlock - 1) / kSsBlock;
    const int32_t lo = threadIdx.x * per, hi = lo + per < nh ? lo + per : nh;
    int32_t part = 0;
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = st->hubs[hl][i];
        const int32_t e0 = heavy ? a.lend[s] : a.off[s], e1 = defer ? a.lend[s] : a.off[s + 1];
        hs[i] = e0;
        hd[i] = a.dist[s];
        part += e1 - e0;
    }
    int32_t total = 0;
    int32_t ex = block_exclusive_scan(part, &total);
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = st->hubs[hl][i];
        hscan[i] = ex;
        ex += (defer ? a.lend[s] : a.off[s + 1]) - hs[i];
    }
    __syncthreads();
    const int32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const int32_t k0 = blockIdx.x * chunk, k1 = k0 + chunk < total ? k0 + chunk : total;
    for (int32_t e = k0 + (int32_t)threadIdx.x; e < k1; e += kSsBlock) {
        const int32_t j = upper_bound_i32(hscan, 0, nh, e) - 1;
        ss_relax(a, c, nbr, wt, hd[j], hs[j] + (e - hscan[j]), true);
    }
    ss_flush(a, ph % kSsSets, c.nf, c.fdeg, INT_MAX);
}

struct SsSeed {
    int32_t source;
    int32_t delta;
    int32_t pull_min;
    int32_t grid;  // workgroups of the run's step launches (the flag layout)
};

// Clears every array of a run, seeds the source as F_0 (bitmap 0, its chunk
// flag in set 3, its count and out-degree in count set 2: what step 0 reads)
// and writes the record of step -1 (a scan of the window [0, delta) that
// found vertices and left the source in F', so step 0 relaxes it there).
__global__ void ss_reset_kernel(SsArgs a, SsSeed sd) {
    const int32_t lines = kSsSets * kSsStripes;
    const int32_t fwords = kSsFlagSets * a.flag_bytes / 8;
    const int32_t src = sd.source;
    // the source's chunk flag in set 3
    const int64_t fbyte = 3LL * a.flag_bytes + ss_fpos((src >> 6) / kBsChunkWords, sd.grid);
    int32_t n = a.V > fwords ? a.V : fwords;
    n = n > lines ? n : lines;
    n = n > a.ustride ? n : a.ustride;
    n = n > a.flag_bytes ? n : a.flag_bytes;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < a.V) a.dist[i] = i == src ? 0 : INT_MAX;
        if (i < a.ustride) {  // nothing pending
            a.ubase[i] = INT_MAX;
            a.ucl[i] = INT_MAX;
            a.ucl[a.ustride + i] = INT_MAX;
        }
        if (i < a.words) {
            a.ws[i] = 0;
            const uint64_t s = (i == (src >> 6)) ? (1ull << (src & 63)) : 0ull;
            a.bm[0][i] = s;
            a.bm[1][i] = 0;
            a.bm[2][i] = 0;
        }
        if (i < a.flag_bytes) a.wsflag[i] = 0;
        if (i < fwords)
            reinterpret_cast<uint64_t *>(a.flags)[i] = (i == (int32_t)(fbyte >> 3)) ? 1ull << (8 * (fbyte & 7)) : 0ull;
        if (i < lines) {
            const bool seed = i == 2 * kSsStripes;  // count set 2, first stripe
            a.cnt[i * kSsLineI32] = seed ? 1 : 0;
            a.fdg[i * (kSsLineI32 / 2)] = seed ? (long long)(a.off[src + 1] - a.off[src]) : 0;
            a.mn[i * kSsLineI32] = seed ? -1 : INT_MAX;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SsState *st = a.st;
        st->done = 0;
        st->slots = 0;
        st->steps = 0;
        st->windows = 0;
        st->pulls = 0;
        st->base = 0;
        st->delta = sd.delta;
        st->pull_min = sd.pull_min;
        st->dist = a.dist;
        for (int i = 0; i < 3; ++i) st->hub_n[i] = 0;
        SsRec r{};
        r.mode = SS_ADVANCE;
        r.lo = 0;
        r.hi = ss_window_end(0, sd.delta);
        r.steps = -1;    // the seeding scan is not counted as a step
        r.windows = -1;  // nor is the first window counted as entered
        st->rec[1] = r;
        st->rec[0] = r;
    }
}

// Light/heavy split of the out-lists for one delta: flag[e] = w < delta,
// scanned into L (L[e] = light edges before e); each list keeps its order
// within the light part and within the heavy part.
__global__ void ss_light_flag_kernel(const int32_t *wt, int64_t E, int32_t delta, int32_t *flag) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e <= E; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = e < E && wt[e] < delta ? 1 : 0;
}

__global__ void ss_split_kernel(const int32_t *esrc, const int32_t *off, const int32_t *nbr, const int32_t *wt,
                                const int32_t *L, int64_t E, int32_t *snbr, int32_t *swt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = esrc[e];
        const int32_t base = off[v];
        const int32_t lb = L[e] - L[base];               // light edges of v before e
        const int32_t nl = L[off[v + 1]] - L[base];      // light edges of v
        const bool light = L[e + 1] > L[e];
        const int32_t pos = light ? base + lb : base + nl + ((int32_t)e - base - lb);
        snbr[pos] = nbr[e];
        swt[pos] = wt[e];
    }
}

__global__ void ss_lend_kernel(const int32_t *off, const int32_t *L, int32_t V, int32_t *lend) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        lend[v] = off[v] + (L[off[v + 1]] - L[off[v]]);
}

// ---------------------------------------------------------------- host side
struct SsWork {
    SsState *st = nullptr;
    uint64_t *bm[3] = {nullptr, nullptr, nullptr};
    uint8_t *flags = nullptr;  // kSsFlagSets sets of chunk flags, flag_bytes apart
    int32_t flag_bytes = 0;
    int32_t chunks = 0;
    int32_t *ubase = nullptr, *ucl = nullptr;  // pending bounds
    int32_t ustride = 0;
    uint64_t *ws = nullptr;
    uint8_t *wsflag = nullptr;
    int32_t *cnt = nullptr, *mn = nullptr;  // count stripes
    long long *fdg = nullptr;
    int32_t *snbr = nullptr, *swt = nullptr, *lend = nullptr;  // light/heavy split
    int32_t split_delta = -1;  // delta of the split (-1: none built)
    bool split = false;        // the split has heavy edges and is used
    int32_t *host_flag = nullptr;  // pinned, coherent: {done, slots, steps, windows, pulls} written by the last step
    hipEvent_t flag_ev = nullptr;
    const int32_t *poff = nullptr, *pnbr = nullptr, *pwt = nullptr;
    int32_t pull_min = 0;
    bool hubs = false;  // some out-degree reaches kSsHubMin: the hub kernel follows each step
    int32_t wide_chunks = 0;
    int32_t *chunk_v = nullptr, *chunk_e0 = nullptr;
    std::map<int32_t, hipGraphExec_t> execs;  // by steps * 12 + first phase (the run's dist is read from the state)
    int32_t recent_steps[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // step launches of the last runs (ring)
    int32_t grid = 0;  // workgroups of the recorded step launches (the layout of flags and bounds)
    int32_t units = 0;
    int32_t recent_n = 0;
};

static PlanRegistry<SsWork> g_ss;

static void drop_ss_execs(SsWork *w) {
    for (auto &kv : w->execs) (void)hipGraphExecDestroy(kv.second);
    w->execs.clear();
}

static void free_ss(SsWork *w) {
    drop_ss_execs(w);
    dfree(w->st);
    for (auto p : w->bm) dfree(p);
    dfree(w->flags);
    dfree(w->ubase);
    dfree(w->ucl);
    dfree(w->ws);
    dfree(w->wsflag);
    dfree(w->cnt);  // (mn and fdg point into it)
    dfree(w->snbr);
    dfree(w->swt);
    dfree(w->lend);
    dfree(w->chunk_v);
    dfree(w->chunk_e0);
    if (w->flag_ev) (void)hipEventDestroy(w->flag_ev);
    if (w->host_flag) (void)hipHostFree(w->host_flag);
    delete w;
}

void release_sssp(const gi_graph *g) {
    if (SsWork *p = g_ss.take(g)) free_ss(p);
}

// Persistent grid: resident workgroups of the step kernel.
static int ss_grid() {
    static const int per_cu = resident_per_cu((const void *)ss_step_kernel, kSsBlock);
    int g = device_info().num_cus * per_cu;
    return g < kSsMaxGrid ? g : kSsMaxGrid;
}

static int ensure_ss(gi_graph *g, SsWork **out) {
    if ((*out = g_ss.find(g)) != nullptr) return GI_OK;
    const int64_t words = bitmap_words(g->V);
    const int64_t chunks = ceil_div(words, kBsChunkWords);
    // every workgroup owns at most 64 chunks (one per lane of a ballot)
    if (chunks > 64LL * ss_grid()) return fail(GI_ERR_RANGE, "fused SSSP: %d vertices exceed this device's grid", g->V);
    const gi_graph *pg = g;  // in-edges for pull steps
    if (g->directed) {
        gi_graph *t = nullptr;
        GI_TRY(gi_graph_transpose(g, &t));
        pg = t;
    }
    if (!pg->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    GI_TRY(ensure_max_degree(g, nullptr));
    SsWork *w = new SsWork();
    w->hubs = g->max_degree >= kSsHubMin;
    w->poff = pg->d_off;
    w->pnbr = pg->d_nbr;
    w->pwt = pg->d_w;
    w->pull_min = g->V / kSsPullDiv;
    w->chunks = (int32_t)chunks;
    w->units = (int32_t)ceil_div(words, kSsPullUnit);
    const size_t lines = (size_t)kSsSets * kSsStripes;
    int rc = dalloc(&w->st, 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->bm[i], (size_t)words + 1);
    if (rc == GI_OK) rc = dalloc(&w->ws, (size_t)words + 1);
    // a stripe's three counters share its line: |F'| at int 0, the least
    // pending distance at int 1, the out-edges at ints 2-3
    if (rc == GI_OK) rc = dalloc(&w->cnt, lines * kSsLineI32);
    if (rc == GI_OK) {
        w->mn = w->cnt + 1;
        w->fdg = reinterpret_cast<long long *>(w->cnt + 2);
    }
    if (rc == GI_OK &&
        hipHostMalloc((void **)&w->host_flag, 8 * sizeof(int32_t), hipHostMallocCoherent) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
    if (rc == GI_OK && hipEventCreateWithFlags(&w->flag_ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(GI_ERR_HIP, "event create");
    if (rc == GI_OK) {  // chunks of the wide in-lists, for pull steps
        std::vector<int32_t> poff((size_t)g->V + 1), cv, ce;
        if (hipMemcpy(poff.data(), pg->d_off, sizeof(int32_t) * (g->V + 1), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(GI_ERR_HIP, "offsets read");
        for (int32_t v = 0; rc == GI_OK && v < g->V; ++v) {
            if (poff[v + 1] - poff[v] < kSsPullWide) continue;
            for (int32_t e = poff[v]; e < poff[v + 1]; e += kSsPullChunk) {
                cv.push_back(v);
                ce.push_back(e);
            }
        }
        w->wide_chunks = (int32_t)cv.size();
        if (rc == GI_OK) rc = dalloc(&w->chunk_v, cv.size() + 1);
        if (rc == GI_OK) rc = dalloc(&w->chunk_e0, ce.size() + 1);
        if (rc == GI_OK && !cv.empty() &&
            (hipMemcpy(w->chunk_v, cv.data(), sizeof(int32_t) * cv.size(), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(w->chunk_e0, ce.data(), sizeof(int32_t) * ce.size(), hipMemcpyHostToDevice) != hipSuccess))
            rc = fail(GI_ERR_HIP, "plan upload");
    }
    if (rc != GI_OK) {
        free_ss(w);
        return rc;
    }
    g_ss.put(g, w);
    *out = w;
    return GI_OK;
}

static SsArgs ss_args(gi_graph *g, SsWork *w, int32_t *dist) {
    SsArgs a;
    a.st = w->st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.wt = g->d_w;
    a.snbr = w->split ? w->snbr : nullptr;
    a.swt = w->split ? w->swt : nullptr;
    a.lend = w->split ? w->lend : nullptr;
    a.defer_min = g->E / kSsDeferDiv > 1 ? g->E / kSsDeferDiv : 1;
    // symmetric graphs: the split out-lists are the split in-lists
    a.heavy_pull_min = w->split && !g->directed ? std::max<long long>(g->E / kSsHeavyPullDiv, 1) : 0;
    a.poff = w->poff;
    a.pnbr = w->pnbr;
    a.pwt = w->pwt;
    a.dist = dist;
    a.V = g->V;
    a.words = (int32_t)bitmap_words(g->V);
    for (int i = 0; i < 3; ++i) a.bm[i] = w->bm[i];
    a.flags = w->flags;
    a.flag_bytes = w->flag_bytes;
    a.ubase = w->ubase;
    a.ucl = w->ucl;
    a.ustride = w->ustride;
    a.ws = w->ws;
    a.wsflag = w->wsflag;
    a.chunks = w->chunks;
    a.cnt = w->cnt;
    a.fdg = w->fdg;
    a.mn = w->mn;
    a.host_flag = w->host_flag;
    a.wide_chunks = w->wide_chunks;
    a.chunk_v = w->chunk_v;
    a.chunk_e0 = w->chunk_e0;
    return a;
}

// The light/heavy split for `delta` (built on first use of a delta; the
// recorded steps are dropped when the arrays they name change).
static int ensure_split(gi_graph *g, SsWork *w, int32_t delta, hipStream_t s) {
    if (w->split_delta == delta) return GI_OK;
    const int64_t E = g->E;
    bool use = false;
    if (E > 0) {
        GI_TRY(ensure_coo(g, s));
        int32_t *flag = nullptr, *L = nullptr;
        int rc = dalloc(&flag, (size_t)E + 1);
        if (rc == GI_OK) rc = dalloc(&L, (size_t)E + 1);
        if (rc == GI_OK && !w->snbr) rc = dalloc(&w->snbr, (size_t)E);
        if (rc == GI_OK && !w->swt) rc = dalloc(&w->swt, (size_t)E);
        if (rc == GI_OK && !w->lend) rc = dalloc(&w->lend, (size_t)g->V + 1);
        int32_t light = 0;
        if (rc == GI_OK) {
            const int grid = persistent_grid(256);
            ss_light_flag_kernel<<<grid, 256, 0, s>>>(g->d_w, E, delta, flag);
            rc = exclusive_scan_i32(flag, L, E + 1, s);
            if (rc == GI_OK) rc = read_count(L + E, &light, s);
        }
        if (rc == GI_OK && light < E) {  // some edge is heavy: build and use the split
            const int grid = persistent_grid(256);
            ss_split_kernel<<<grid, 256, 0, s>>>(g->d_esrc, g->d_off, g->d_nbr, g->d_w, L, E, w->snbr, w->swt);
            ss_lend_kernel<<<grid, 256, 0, s>>>(g->d_off, L, g->V, w->lend);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
                rc = fail(GI_ERR_HIP, "light/heavy split");
            use = true;
        }
        dfree(flag);
        dfree(L);
        GI_TRY(rc);
    }
    if (use != w->split || use) drop_ss_execs(w);  // the arrays or their use changed
    w->split = use;
    w->split_delta = delta;
    for (int32_t &k : w->recent_steps) k = 0;  // step counts of other deltas do not carry over
    return GI_OK;
}

int sssp_light_split(gi_graph *g, int32_t delta, SsSplitView *out, hipStream_t s) {
    SsWork *w = nullptr;
    GI_TRY(ensure_ss(g, &w));
    GI_TRY(ensure_split(g, w, delta, s));
    *out = SsSplitView{};
    if (w->split) {
        out->nbr = w->snbr;
        out->wt = w->swt;
        out->lend = w->lend;
    }
    return GI_OK;
}

static int ss_reset(SsWork *w, const SsArgs &a, int32_t source, int32_t delta, int32_t pull_min, hipStream_t s) {
    w->host_flag[0] = 0;  // no run of this graph's engine is in flight: set again by the last step
    const int64_t n = std::max<int64_t>({(int64_t)a.V, (int64_t)kSsFlagSets * a.flag_bytes / 8,
                                         (int64_t)kSsSets * kSsStripes, (int64_t)a.ustride,
                                         (int64_t)a.flag_bytes});
    const int rg = (int)std::min<int64_t>(ceil_div(n, 256), 8192);
    ss_reset_kernel<<<rg, 256, 0, s>>>(a, SsSeed{source, delta, pull_min, w->grid});
    GI_LAUNCH_CHECK();
    return GI_OK;
}

// Launch k of a recorded block runs step base + k with phase (ph0 + k) mod
// 12, ph0 = base mod 12: one recording serves every run on this graph and
// delta and every block of a run that starts at the same phase. On a graph
// with hubs each step is followed by its grid-wide hub bin.
static int ss_step_graph(SsWork *w, const SsArgs &ca, int32_t steps, int32_t ph0, hipGraphExec_t *out) {
    const int32_t key = steps * 12 + ph0;
    auto it = w->execs.find(key);
    if (it != w->execs.end()) {
        *out = it->second;
        return GI_OK;
    }
    const int grid = w->grid;
    hipGraphExec_t exec = nullptr;
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t k = 0; k < steps; ++k) {
                const int32_t ph = (ph0 + k) % 12;
                L(ss_step_kernel, grid, kSsBlock, ca, k, ph);
                if (w->hubs) L(ss_hub_kernel, grid, kSsBlock, ca, k, ph);
            }
        },
        &exec));
    if (w->execs.size() >= 12) drop_ss_execs(w);
    w->execs[key] = exec;
    *out = exec;
    return GI_OK;
}

static int32_t ss_first_block(const SsWork *w) {
    int32_t b = 0;
    for (int32_t k : w->recent_steps) b = k + 1 > b ? k + 1 : b;
    b = b > 1 ? (b + 7) / 8 * 8 : 16;
    return b < kSsMaxBlock ? b : kSsMaxBlock;
}

// Workgroups of the step launches: the resident grid, or max_cta (the
// schedule's) when positive (at most kSsMaxGrid), but enough that every
// workgroup owns at most 64 chunks. The chunk flags and the pending bounds
// are laid out for it (workgroup b's 64 at b * 64), so they are allocated
// again, and the recorded steps dropped, when it changes.
static int ss_set_grid(SsWork *w, int32_t max_cta) {
    int32_t grid = ss_grid();
    if (max_cta > 0) grid = max_cta < kSsMaxGrid ? max_cta : kSsMaxGrid;
    const int32_t least = (int32_t)ceil_div(w->chunks, 64);
    if (grid < least) grid = least;
    if (grid == w->grid) return GI_OK;
    drop_ss_execs(w);
    dfree(w->flags);
    dfree(w->wsflag);
    dfree(w->ubase);
    dfree(w->ucl);
    w->flags = w->wsflag = nullptr;
    w->ubase = w->ucl = nullptr;
    w->grid = 0;
    w->flag_bytes = 64 * grid;  // a multiple of 16
    w->ustride = (int32_t)(ceil_div(w->units, 64LL * grid) * 64 * grid);
    int rc = dalloc(&w->flags, (size_t)kSsFlagSets * w->flag_bytes);
    if (rc == GI_OK) rc = dalloc(&w->wsflag, (size_t)w->flag_bytes);
    if (rc == GI_OK) rc = dalloc(&w->ubase, (size_t)w->ustride);
    if (rc == GI_OK) rc = dalloc(&w->ucl, 2 * (size_t)w->ustride);
    if (rc == GI_OK) w->grid = grid;
    return rc;
}

int sssp_fused_run(gi_graph *g, int32_t source, int32_t delta, int32_t *d_dist, int32_t *counts,
                   hipStream_t s, int32_t max_cta) {
    SsWork *w = nullptr;
    GI_TRY(ensure_ss(g, &w));
    GI_TRY(ensure_split(g, w, delta, s));
    GI_TRY(ss_set_grid(w, max_cta));
    SsArgs a = ss_args(g, w, d_dist);
    GI_TRY(ss_reset(w, a, source, delta, w->pull_min, s));
    SsArgs ca = a;
    ca.dist = nullptr;  // recorded kernels take the run's dist from the state
    // Replay a block of step launches as long as the longest of the last runs
    // (a launch past the end returns at once, a block too short costs a host
    // round trip), wait on its event and read what the last step wrote to
    // pinned memory; longer runs continue with blocks of 16, 32, ... 256
    // steps, each after copying its first step into the state.
    int32_t block = ss_first_block(w), next = 16;
    for (int64_t step = 0; step < (1ll << 26);) {
        hipGraphExec_t exec = nullptr;
        GI_TRY(ss_step_graph(w, ca, block, (int32_t)(step % 12), &exec));
        GI_HIP(hipGraphLaunch(exec, s));
        step += block;
        GI_HIP(hipEventRecord(w->flag_ev, s));
        GI_HIP(hipEventSynchronize(w->flag_ev));
        if (w->host_flag[0]) {
            const int32_t *h = w->host_flag;
            w->recent_steps[w->recent_n++ % 8] = h[1];
            if (counts) {  // relax steps, pull steps, windows entered, step launches
                counts[0] = h[2];
                counts[1] = h[4];
                counts[2] = h[3];
                counts[3] = h[1];
            }
            // a longer run than before: record the next run's first block
            // now, so that its launch does not wait on the recording
            hipGraphExec_t nx = nullptr;
            return ss_step_graph(w, ca, ss_first_block(w), 0, &nx);
        }
        w->host_flag[6] = (int32_t)step;  // the next block's first step
        GI_HIP(hipMemcpyAsync(&w->st->base, &w->host_flag[6], sizeof(int32_t), hipMemcpyHostToDevice, s));
        block = next;
        next = next * 2 < 256 ? next * 2 : 256;
    }
    return fail(GI_ERR_RANGE, "SSSP did not finish");
}

// Eager step loop for measurement: each step launch (and its hub bin) with
// HIP events around it; the step's record is read after it. A relax or heavy
// step is reported as a push or a pull step by the record's pull flag.
static int sssp_trace(gi_graph *g, int32_t source, int32_t delta, int32_t pull_min, int32_t *d_dist,
                      int32_t max_slots, float *kernel_ms, int64_t *info, int32_t *num_slots, hipStream_t s) {
    SsWork *w = nullptr;
    GI_TRY(ensure_ss(g, &w));
    GI_TRY(ensure_split(g, w, delta, s));
    GI_TRY(ss_set_grid(w, 0));
    SsArgs a = ss_args(g, w, d_dist);
    a.host_flag = nullptr;
    GI_TRY(ss_reset(w, a, source, delta, pull_min >= 0 ? pull_min : w->pull_min, s));
    hipEvent_t ev[3] = {};
    int rc = GI_OK;
    for (int i = 0; i < 3 && rc == GI_OK; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = fail(GI_ERR_HIP, "event create");
    const int grid = w->grid;
    SsRec r{};
    int32_t k = 0;
    for (; rc == GI_OK && k <= max_slots; ++k) {
        const int32_t ph = k % 12;
        (void)hipEventRecord(ev[0], s);
        ss_step_kernel<<<grid, kSsBlock, 0, s>>>(a, k, ph);
        (void)hipEventRecord(ev[1], s);
        if (w->hubs) ss_hub_kernel<<<grid, kSsBlock, 0, s>>>(a, k, ph);
        (void)hipEventRecord(ev[2], s);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(ev[2]) != hipSuccess ||
            hipMemcpy(&r, &w->st->rec[ph & 1], sizeof(SsRec), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = fail(GI_ERR_HIP, "step launch");
            break;
        }
        if (r.mode == SS_DONE) break;
        if (k == max_slots) break;
        float t[2];
        for (int i = 0; i < 2; ++i) (void)hipEventElapsedTime(&t[i], ev[i], ev[i + 1]);
        info[5 * k + 0] = r.mode;
        info[5 * k + 1] = r.pull;
        info[5 * k + 2] = r.front;
        info[5 * k + 3] = r.hi;
        info[5 * k + 4] = r.fdeg;
        float *row = kernel_ms + 6 * k;  // push, pull, hub, -, window scan, -
        for (int i = 0; i < 6; ++i) row[i] = 0.f;
        if (r.mode == SS_ADVANCE) row[4] = t[0];
        else if (r.pull) row[1] = t[0];
        else row[0] = t[0];
        row[2] = t[1];
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == GI_OK && k >= max_slots && r.mode != SS_DONE) rc = fail(GI_ERR_RANGE, "more than %d steps", max_slots);
    *num_slots = k;
    return rc;
}

}  // namespace gi

extern "C" int gi_sssp_trace(gi_graph *g, int32_t source, int32_t delta, int32_t pull_min, int32_t *d_dist,
                             int32_t max_slots, float *kernel_ms, int64_t *info, int32_t *num_slots,
                             void *stream) {
    if (!g || !d_dist || !kernel_ms || !info || !num_slots) return gi::fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return gi::fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (source < 0 || source >= g->V) return gi::fail(GI_ERR_RANGE, "source %d out of range", source);
    if (delta < 1) return gi::fail(GI_ERR_RANGE, "delta must be positive");
    if (!g->d_w) return gi::fail(GI_ERR_INVALID, "SSSP needs edge weights");
    if (max_slots < 1) return gi::fail(GI_ERR_RANGE, "max_slots must be positive");
    return gi::sssp_trace(g, source, delta, pull_min, d_dist, max_slots, kernel_ms, info, num_slots,
                          (hipStream_t)stream);
}
