// Fused delta-stepping SSSP: the loop of apps/sssp.cpp:34-57 (dequeue the
// current priority window, relax its out-edges with writeMin, re-enqueue
// improved vertices) with the window bookkeeping on the device and the
// phases recorded as hipGraphs. Shortest-path distances are unique, so the
// order of relaxations does not change the result.
//
// The window is [lo, hi). Every vertex below lo is settled; the near
// frontier F holds the vertices improved into the window since their last
// expansion (three rotating bitmaps with chunk summaries, as in the fused
// BFS). Vertices improved to a distance >= hi are pending: they are not kept
// in a set of their own (the reference's priority queue,
// gpu_priority_queue.h), because a vertex is pending exactly when
// hi <= dist < INT_MAX. Instead every unit of 1024 vertices keeps a lower
// bound of its pending distances (umin, lowered with atomicMin when a
// relaxation leaves a vertex pending; most relaxations find the bound
// already lower and skip the atomic). A window advance scans dist only in
// the units whose bound falls inside the new window, and rewrites their
// bounds exactly; the other units are skipped, so later windows read a small
// part of dist instead of all of it.
// Light and heavy edges (the delta-stepping split of Meyer and Sanders):
// each out-list is copied with its light edges (w < delta) first. An edge
// from the window [lo, hi) that is heavy lands at or beyond hi, so it can
// never put a vertex back into the window. A push step whose F has at least
// E / kSsDeferDiv out-edges relaxes light edges only and records F in the
// window set W; when the window is settled, one heavy step relaxes the heavy
// edges of W once, from final distances, instead of once per improvement of
// each vertex within the window. Steps with smaller F relax all edges (no
// extra step for windows that settle in one step).
// A recorded slot is five launches, each a no-op unless its mode is active:
//   ss_relax   push: relax the out-edges of F, spread evenly over each
//              workgroup (vertices of at least kSsHubMin edges go to the
//              grid-wide bin); an improvement below hi joins F';
//              or pull, for a large F: each unsettled vertex takes the
//              minimum over its in-neighbours in F (no atomics on distances);
//   ss_hub     the grid-wide bin: edges of very wide frontier vertices (not
//              launched for a graph without such vertices);
//   ss_decide  count F'; empty -> advance to [hi, hi + delta);
//   ss_farmove scan dist in the units whose bound is below hi: [lo, hi)
//              joins F; also the least pending distance >= hi;
//   ss_decide2 count the joins; none -> jump to the window that starts at
//              that least distance (gpu_priority_queue.h:126-146 starts a
//              window at the minimum remaining priority too), or done when
//              nothing is pending.
#include <climits>
#include <map>
#include <vector>

#include "gi_bitset.h"
#include "gi_lb.h"

namespace gi {

constexpr int kSsBlock = 256;
constexpr int kSsDecBlock = 1024;     // the single-workgroup decide kernels
constexpr int32_t kSsHubMin = 4096;   // at least this wide: grid-wide bin
constexpr int32_t kSsMaxHubs = 1024;
constexpr int32_t kSsMaxGrid = 8192;  // partial-sum slots per kernel (one per workgroup)
constexpr int32_t kSsPullDiv = 8;     // a step pulls when |F| > V / kSsPullDiv
constexpr int32_t kSsMinSlots = 9;    // slot row of the window scan's least pending distances

constexpr int kSsPullUnit = 4 * (kSsBlock / 64);  // words per unit (pull work, window scan, umin)

constexpr long long kSsDeferDiv = 16;  // F with E / kSsDeferDiv out-edges or more: heavy edges wait

enum SsMode : int32_t { SS_RELAX = 0, SS_ADVANCE = 1, SS_HEAVY = 2 };

struct SsState {
    int32_t done;
    int32_t mode;
    int32_t phase;       // F = bm[phase % 3]
    int32_t steps;       // relax steps run
    int32_t windows;     // windows entered after the first
    int32_t slots;       // recorded slots consumed
    int32_t pulls;       // relax steps that pulled
    int32_t hub_count;
    int32_t lo;          // window [lo, hi)
    int32_t hi;
    int32_t delta;
    int32_t front;       // |F|
    int32_t next_min;    // least pending distance >= hi seen by the last scan
    int32_t pull;        // the next relax step pulls (large frontier)
    int32_t pull_min;    // |F| above which a step pulls
    int32_t defer;       // the next push step relaxes light edges only and records F in W
    int32_t ws_any;      // W holds vertices whose heavy edges wait
    long long fdeg;      // out-edges of F
    int32_t *dist;       // the run's distances (recorded slots read it from here)
    int32_t hubs[kSsMaxHubs];
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
    const int32_t *poff;  // in-edges with their weights (the graph itself when symmetric)
    const int32_t *pnbr;
    const int32_t *pwt;
    int32_t *dist;
    int32_t V;
    int32_t words;
    int32_t sum_words;
    uint64_t *bm[3];
    uint64_t *sum[3];
    uint8_t *cflag;      // chunk flags of the bitmap being filled (bs_flag_chunk)
    int32_t *umin;       // per unit of kSsPullUnit words: lower bound of its pending distances
    uint64_t *ws;        // W: vertices of the window whose heavy edges wait
    uint64_t *wsum;      // its chunk summary (built when the heavy step starts)
    uint8_t *wsflag;     // its chunk flags
    int32_t chunks;
    int32_t *slots;      // [3 kernels][3 counters][kSsMaxGrid], then row kSsMinSlots
    int32_t wide_chunks;       // pull: chunks of the in-lists of wide vertices
    const int32_t *chunk_v;    // vertex of each chunk
    const int32_t *chunk_e0;   // first in-edge of each chunk
};

struct SsCtx {
    int32_t hi;
    uint64_t *bm_out;
    int32_t nf;    // per-lane: new members of F'
    int32_t fdeg;  // per-lane: their out-edges
};

// writeMin (apps/sssp.cpp:24-27) of dist[v] to nd, known to be below the
// value last read: a strict improvement below hi joins F'; one at or above
// hi leaves the vertex pending.
__device__ __forceinline__ void ss_claim(const SsArgs &a, SsCtx &c, int32_t v, int32_t nd) {
    const int32_t old = atomicMin(&a.dist[v], nd);
    if (nd >= old) return;
    if (nd >= c.hi) {  // pending: keep its unit's bound at or below nd
        int32_t *m = &a.umin[v / (64 * kSsPullUnit)];
        if (nd < *m) atomicMin(m, nd);
        return;
    }
    const uint64_t bit = 1ull << (v & 63);
    const unsigned long long o = atomicOr((unsigned long long *)&c.bm_out[v >> 6], bit);
    if (!(o & bit)) {
        c.nf += 1;
        c.fdeg += a.off[v + 1] - a.off[v];
        if (o == 0) bs_flag_chunk(a.cflag, v >> 6);
    }
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

// Push relax step. `pre` holds kBsMaxSumWords + 1 entries, `list`
// kBsChunkVerts (both LDS).
__device__ __forceinline__ void ss_push_body(const SsArgs &a, int32_t *pre, int32_t *list) {
    SsState *st = a.st;
    __shared__ int32_t pre_s[kSsBlock], e0_s[kSsBlock], du_s[kSsBlock];
    __shared__ int32_t n_s;
    const bool heavy = st->mode == SS_HEAVY;  // F is W, edges from lend
    const bool defer = !heavy && st->defer;   // edges up to lend; F joins W
    const int32_t *nbr = heavy || defer ? a.snbr : a.nbr;  // split lists only where the split is used
    const int32_t *wt = heavy || defer ? a.swt : a.wt;
    const int32_t phase = st->phase;
    const int32_t iin = phase % 3, iout = (phase + 1) % 3, iclr = (phase + 2) % 3;
    SsCtx c;
    c.hi = st->hi;
    c.bm_out = bs_pick3(a.bm[0], a.bm[1], a.bm[2], iout);
    c.nf = 0;
    c.fdeg = 0;
    // the bitmap two phases ahead is free (its summary is cleared by ss_decide)
    bs_clear_chunks(bs_pick3(a.bm[0], a.bm[1], a.bm[2], iclr), bs_pick3(a.sum[0], a.sum[1], a.sum[2], iclr),
                    a.words, a.sum_words, pre);
    const uint64_t *bm_in = heavy ? a.ws : bs_pick3(a.bm[0], a.bm[1], a.bm[2], iin);
    const uint64_t *si = heavy ? a.wsum : bs_pick3(a.sum[0], a.sum[1], a.sum[2], iin);
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    const int32_t nz = bs_summary_prefix(si, a.sum_words, pre);
    for (int32_t k = blockIdx.x; k < nz; k += gridDim.x) {
        const int32_t chunk = bs_kth_chunk(si, pre, a.sum_words, k);
        if (threadIdx.x < 64) {  // wave 0 lists the chunk's members
            const int32_t w = chunk * kBsChunkWords + (int32_t)threadIdx.x;
            uint64_t word = w < a.words ? bm_in[w] : 0ull;
            if (word) {  // this lane alone owns word w of W in this step
                if (heavy) {
                    a.ws[w] = 0;  // consumed
                } else if (defer) {
                    const uint64_t o = a.ws[w];
                    a.ws[w] = o | word;
                    if (!o) bs_flag_chunk(a.wsflag, w);
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
        // rounds of 256 members: their edges (hubs aside) are spread evenly
        // over the workgroup, each thread finding its edge's owner in LDS
        for (int32_t b0 = 0; b0 < n; b0 += kSsBlock) {
            const int32_t j = b0 + (int32_t)threadIdx.x;
            const bool has = j < n;
            const int32_t s = has ? list[j] : 0;
            int32_t e0 = has ? a.off[s] : 0, e1 = has ? a.off[s + 1] : 0;
            if (has && heavy) e0 = a.lend[s];
            if (has && defer) e1 = a.lend[s];
            const int32_t dg = e1 - e0;
            const bool hub = has && dg >= kSsHubMin;
            const int32_t hpos = wave_atomic_add(&st->hub_count, hub);
            if (hub && hpos < kSsMaxHubs) st->hubs[hpos] = s;
            const bool mine = has && (!hub || hpos >= kSsMaxHubs);
            int32_t total = 0;
            const int32_t ex = block_exclusive_scan(mine ? dg : 0, &total);
            pre_s[threadIdx.x] = ex;
            e0_s[threadIdx.x] = e0;
            du_s[threadIdx.x] = mine ? a.dist[s] : 0;
            __syncthreads();
            for (int32_t k0 = 0; k0 < total; k0 += kSsBlock * kSsBatch) {
                int32_t v[kSsBatch], nd[kSsBatch], dv[kSsBatch];
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p) {
                    const int32_t k = k0 + p * kSsBlock + (int32_t)threadIdx.x;
                    v[p] = -1;
                    nd[p] = 0;
                    if (k < total) {  // owner: the last member whose prefix is <= k
                        const int32_t o = upper_bound_i32(pre_s, 0, kSsBlock, k) - 1;
                        const int32_t e = e0_s[o] + (k - pre_s[o]);
                        v[p] = nbr[e];
                        nd[p] = du_s[o] + wt[e];
                    }
                }
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p) dv[p] = v[p] >= 0 ? a.dist[v[p]] : INT_MIN;
#pragma unroll
                for (int p = 0; p < kSsBatch; ++p)
                    if (nd[p] < dv[p]) ss_claim(a, c, v[p], nd[p]);
            }
            __syncthreads();
        }
        __syncthreads();
    }
    (void)wid;
    bs_flush3<kSsBlock>(a.slots, kSsMaxGrid, c.nf, c.fdeg, 0);
}

// Pull relax step for a large frontier: every vertex not yet settled below
// the window takes the minimum of dist[u] + w over its in-neighbours u in F
// (no atomics on dist: a vertex's word of 64 vertices belongs to one wave).
// Work is in units of 16 words per workgroup (4 per wave), so a unit's
// summary bit costs at most one atomic.
constexpr int32_t kSsPullWide = 4096;            // in-degree at least this: chunked pull
constexpr int32_t kSsPullChunk = 16 * 1024;      // in-edges per chunk (a workgroup's item)
constexpr int kSsPasses = 4;                     // cooperative passes in flight per wave

// Pull relax step; `best_s` holds kSsBlock entries of LDS.
__device__ __forceinline__ void ss_pull_body(const SsArgs &a, int32_t *pre, int32_t *best_s) {
    SsState *st = a.st;
    const int32_t phase = st->phase, hi = st->hi, lo = st->lo;
    const int32_t iin = phase % 3, iout = (phase + 1) % 3, iclr = (phase + 2) % 3;
    bs_clear_chunks(bs_pick3(a.bm[0], a.bm[1], a.bm[2], iclr), bs_pick3(a.sum[0], a.sum[1], a.sum[2], iclr),
                    a.words, a.sum_words, pre);
    const uint64_t *bm_in = bs_pick3(a.bm[0], a.bm[1], a.bm[2], iin);
    uint64_t *bm_out = bs_pick3(a.bm[0], a.bm[1], a.bm[2], iout);
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
            bool open = v < a.V && dv > lo;  // settled vertices cannot improve
            const int32_t e0 = open ? a.poff[v] : 0;
            int32_t len = open ? a.poff[v + 1] - e0 : 0;
            if (len >= kSsPullWide) {  // wide: the chunk items below handle it
                open = false;
                len = 0;
            }
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
                    uu[p] = ee[p] >= 0 ? a.pnbr[ee[p]] : -1;
                    ww[p] = ee[p] >= 0 ? a.pwt[ee[p]] : 0;
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
            if (lane == 0 && pend < a.umin[un]) atomicMin(&a.umin[un], pend);
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
        if (__syncthreads_or(any) && threadIdx.x == 0) bs_flag_chunk(a.cflag, un * kSsPullUnit);
    }
    // wide vertices: each chunk of an in-list is a workgroup item; its minimum
    // joins dist[v] by atomicMin, and an improvement below hi joins F'
    const int32_t per_wave = kSsPullChunk / (kSsBlock / 64);
    for (int32_t k = blockIdx.x; k < a.wide_chunks; k += gridDim.x) {
        const int32_t v = a.chunk_v[k];
        const int32_t dv = a.dist[v];
        if (dv <= lo) continue;  // settled (block-uniform)
        const int32_t end = a.poff[v + 1];
        const int32_t e0 = a.chunk_e0[k] + wid * per_wave;
        const int32_t e1 = e0 + per_wave < end ? e0 + per_wave : end;
        int32_t m = INT_MAX;
#pragma unroll 4
        for (int32_t e = e0 + lane; e < e1; e += 64) {
            const int32_t u = a.pnbr[e];
            if (bs_test(bm_in, u)) {
                const int32_t c = a.dist[u] + a.pwt[e];
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
                int32_t *um = &a.umin[v / (64 * kSsPullUnit)];
                if (m < *um) atomicMin(um, m);
            }
            if (m < old && m < hi) {
                const uint64_t bit = 1ull << (v & 63);
                const unsigned long long o = atomicOr((unsigned long long *)&bm_out[v >> 6], bit);
                if (!(o & bit)) {
                    nf += 1;
                    fdeg += a.off[v + 1] - a.off[v];
                    if (o == 0) bs_flag_chunk(a.cflag, v >> 6);
                }
            }
        }
    }
    bs_flush3<kSsBlock>(a.slots, kSsMaxGrid, nf, fdeg, 0);
}

// One relax step, push or pull as the decide kernels set `pull` (one launch
// either way; the body not taken costs nothing).
__global__ void __launch_bounds__(kSsBlock) ss_relax_kernel(SsArgs a) {
    SsState *st = a.st;
    if (st->done || st->mode == SS_ADVANCE) return;  // uniform
    a.dist = st->dist;
    __shared__ int32_t pre[kBsMaxSumWords + 1];
    __shared__ int32_t lds[kBsChunkVerts];
    if (st->pull)
        ss_pull_body(a, pre, lds);
    else
        ss_push_body(a, pre, lds);
}

// Grid-wide bin: the concatenated edges of the deferred hubs, spread evenly.
__global__ void __launch_bounds__(kSsBlock) ss_hub_kernel(SsArgs a) {
    SsState *st = a.st;
    if (st->done || st->mode == SS_ADVANCE || st->pull) return;
    int32_t nh = st->hub_count;
    if (nh <= 0) return;  // slots stay zero
    a.dist = st->dist;
    if (nh > kSsMaxHubs) nh = kSsMaxHubs;
    const bool heavy = st->mode == SS_HEAVY, defer = !heavy && st->defer;  // the step's edge range
    const int32_t *nbr = heavy || defer ? a.snbr : a.nbr;
    const int32_t *wt = heavy || defer ? a.swt : a.wt;
    const int32_t phase = st->phase;
    SsCtx c;
    c.hi = st->hi;
    c.bm_out = bs_pick3(a.bm[0], a.bm[1], a.bm[2], (phase + 1) % 3);
    c.nf = 0;
    c.fdeg = 0;
    __shared__ int32_t hscan[kSsMaxHubs + 1];
    __shared__ int32_t hs[kSsMaxHubs];  // first edge
    __shared__ int32_t hd[kSsMaxHubs];
    const int32_t per = (nh + kSsBlock - 1) / kSsBlock;
    const int32_t lo = threadIdx.x * per, hi = lo + per < nh ? lo + per : nh;
    int32_t part = 0;
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = st->hubs[i];
        const int32_t e0 = heavy ? a.lend[s] : a.off[s], e1 = defer ? a.lend[s] : a.off[s + 1];
        hs[i] = e0;
        hd[i] = a.dist[s];
        part += e1 - e0;
    }
    int32_t total = 0;
    int32_t ex = block_exclusive_scan(part, &total);
    for (int32_t i = lo; i < hi; ++i) {
        const int32_t s = st->hubs[i];
        hscan[i] = ex;
        ex += (defer ? a.lend[s] : a.off[s + 1]) - hs[i];
    }
    __syncthreads();
    const int32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const int32_t k0 = blockIdx.x * chunk, k1 = k0 + chunk < total ? k0 + chunk : total;
    for (int32_t k = k0 + (int32_t)threadIdx.x; k < k1; k += kSsBlock) {
        const int32_t j = upper_bound_i32(hscan, 0, nh, k) - 1;
        ss_relax(a, c, nbr, wt, hd[j], hs[j] + (k - hscan[j]), true);
    }
    bs_flush3<kSsBlock>(a.slots + 3 * kSsMaxGrid, kSsMaxGrid, c.nf, c.fdeg, 0);
}

// Sums rows r[0..N) of the slot array (row = region * 3 + counter) over
// nslots workgroups in one pass, zeroing them for the next kernel.
template <int N>
__device__ __forceinline__ void ss_collect(int32_t *slot_base, const int32_t (&r)[N], int32_t nslots,
                                          long long (&out)[N]) {
    __shared__ long long red[N][kSsDecBlock / 64];
    long long t[N];
#pragma unroll
    for (int j = 0; j < N; ++j) t[j] = 0;
#pragma unroll 1
    for (int32_t i = threadIdx.x; i < nslots; i += kSsDecBlock) {
        int32_t x[N];
#pragma unroll
        for (int j = 0; j < N; ++j) x[j] = slot_base[r[j] * kSsMaxGrid + i];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            t[j] += x[j];
            slot_base[r[j] * kSsMaxGrid + i] = 0;
        }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        for (int o = 32; o > 0; o >>= 1) t[j] += __shfl_xor(t[j], o);
        if (lane_id() == 0) red[j][threadIdx.x >> 6] = t[j];
    }
    __syncthreads();
    const int32_t lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {  // every wave reduces the per-wave sums
        long long v = lane < kSsDecBlock / 64 ? red[j][lane] : 0;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        out[j] = v;
    }
    __syncthreads();
}

__device__ __forceinline__ int32_t ss_window_end(int32_t lo, int32_t delta) {
    return lo > INT_MAX - delta ? INT_MAX : lo + delta;
}

__global__ void __launch_bounds__(kSsDecBlock) ss_decide_kernel(SsArgs a, int32_t nslots) {
    SsState *st = a.st;
    // every thread reads the state before the barriers in ss_collect; thread 0
    // rewrites it after them
    const int32_t mode = st->mode;
    if (st->done || mode == SS_ADVANCE) return;  // uniform
    const int32_t phase = st->phase;
    const bool ws_any = st->ws_any || (mode == SS_RELAX && st->defer);
    const int32_t rows[4] = {0, 3, 1, 4};  // |F'| and its out-edges from the relax and hub kernels
    long long c[4];
    ss_collect(a.slots, rows, nslots, c);
    const long long nf = c[0] + c[1], fdeg = c[2] + c[3];
    uint64_t *sclr = bs_pick3(a.sum[0], a.sum[1], a.sum[2], (phase + 2) % 3);
    for (int32_t i = threadIdx.x; i <= a.sum_words; i += kSsDecBlock) sclr[i] = 0;
    // the summary of F', from the flags its chunks set
    bs_build_summary<kSsDecBlock>(a.cflag, a.chunks, bs_pick3(a.sum[0], a.sum[1], a.sum[2], (phase + 1) % 3));
    const bool to_heavy = nf == 0 && mode == SS_RELAX && ws_any;  // settled, heavy edges waiting
    if (to_heavy)
        bs_build_summary<kSsDecBlock>(a.wsflag, a.chunks, a.wsum);
    else if (mode == SS_HEAVY)  // the heavy step consumed W
        for (int32_t i = threadIdx.x; i <= a.sum_words; i += kSsDecBlock) a.wsum[i] = 0;
    if (threadIdx.x == 0) {
        st->phase = phase + 1;
        st->steps += 1;
        st->pulls += st->pull;
        st->front = (int32_t)nf;
        st->fdeg = fdeg;
        st->hub_count = 0;
        st->pull = nf > st->pull_min;
        st->defer = a.lend != nullptr && !st->pull && fdeg >= a.defer_min;
        st->ws_any = ws_any;
        if (to_heavy) {
            st->mode = SS_HEAVY;
            st->ws_any = 0;
            st->defer = 0;
        } else if (nf == 0) {  // the window is settled: scan for [hi, hi + delta)
            st->mode = SS_ADVANCE;
            st->lo = st->hi;
            st->hi = ss_window_end(st->hi, st->delta);
            st->next_min = INT_MAX;
        } else if (mode == SS_HEAVY) {
            st->mode = SS_RELAX;  // not reached: heavy edges land at or beyond hi
        }
    }
}

// Window scan over the units whose pending bound is below hi: every vertex
// with lo <= dist < hi joins F (each wave owns its words: plain stores into
// the empty bitmap), the unit's bound becomes its least distance >= hi, and
// the out-edges of the joined vertices are counted. A unit whose bound is at
// least hi holds nothing below hi and is skipped. The least bound over all
// units (the least pending distance) goes to the workgroup's min slot. Units
// of 16 words per workgroup (4 per wave, loads in flight together); a unit's
// chunk flag costs one store.
__global__ void __launch_bounds__(kSsBlock) ss_farmove_kernel(SsArgs a) {
    SsState *st = a.st;
    if (st->done || st->mode != SS_ADVANCE) return;
    a.dist = st->dist;
    __shared__ int32_t ured[2][kSsBlock / 64];
    const int32_t lo = st->lo, hi = st->hi, phase = st->phase;
    uint64_t *f = bs_pick3(a.bm[0], a.bm[1], a.bm[2], phase % 3);  // empty: the last step found nothing
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    int32_t moved = 0, fdeg = 0;
    int32_t beyond = INT_MAX;  // thread 0: least bound over this workgroup's units
    int32_t scanned = 0;       // block-uniform: LDS buffer of the unit reduction
    const int32_t units = (a.words + kSsPullUnit - 1) / kSsPullUnit;
    for (int32_t un = blockIdx.x; un < units; un += gridDim.x) {
        const int32_t um = a.umin[un];  // block-uniform
        if (um >= hi) {
            beyond = um < beyond ? um : beyond;
            continue;
        }
        const int32_t w0 = un * kSsPullUnit + wid * 4;
        int32_t dv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int32_t v = (w0 + i) * 64 + lane;
            dv[i] = v < a.V ? a.dist[v] : INT_MAX;
        }
        bool any = false;
        int32_t ub = INT_MAX;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int32_t v = (w0 + i) * 64 + lane;
            const bool in = dv[i] >= lo && dv[i] < hi;  // dv is INT_MAX past V
            const uint64_t m = __ballot(in);
            if (in) fdeg += a.off[v + 1] - a.off[v];
            if (dv[i] >= hi && dv[i] < ub) ub = dv[i];
            if (m) {
                any = true;
                if (lane == 0) {
                    f[w0 + i] = m;
                    moved += __popcll(m);
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int32_t y = __shfl_xor(ub, o);
            ub = y < ub ? y : ub;
        }
        int32_t *r = ured[scanned & 1];
        if (lane == 0) r[wid] = ub;
        ++scanned;
        if (__syncthreads_or(any) && threadIdx.x == 0) bs_flag_chunk(a.cflag, un * kSsPullUnit);
        if (threadIdx.x == 0) {  // the unit's exact bound
            int32_t m = r[0];
            for (int i = 1; i < kSsBlock / 64; ++i) m = r[i] < m ? r[i] : m;
            a.umin[un] = m;
            beyond = m < beyond ? m : beyond;
        }
    }
    if (threadIdx.x == 0)  // one slot per workgroup: no atomics on a single word
        a.slots[kSsMinSlots * kSsMaxGrid + blockIdx.x] = beyond;
    bs_flush3<kSsBlock>(a.slots + 6 * kSsMaxGrid, kSsMaxGrid, moved, fdeg, 0);
}

__global__ void __launch_bounds__(kSsDecBlock) ss_decide2_kernel(SsArgs a, int32_t nslots) {
    SsState *st = a.st;
    if (st->done) return;
    if (threadIdx.x == 0) st->slots += 1;
    if (st->mode != SS_ADVANCE) return;  // uniform
    const int32_t rows[2] = {6, 7};  // joined vertices and their out-edges
    long long c[2];
    ss_collect(a.slots, rows, nslots, c);
    const long long moved = c[0], fdeg = c[1];
    // the summary of the window's F
    bs_build_summary<kSsDecBlock>(a.cflag, a.chunks, bs_pick3(a.sum[0], a.sum[1], a.sum[2], st->phase % 3));
    // least pending distance over the window scan's workgroups
    __shared__ int32_t red[kSsDecBlock / 64];
    int32_t m = INT_MAX;
    for (int32_t i = threadIdx.x; i < nslots; i += kSsDecBlock) {
        const int32_t x = a.slots[kSsMinSlots * kSsMaxGrid + i];
        m = x < m ? x : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t y = __shfl_xor(m, o);
        m = y < m ? y : m;
    }
    if (lane_id() == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < kSsDecBlock / 64; ++i) m = red[i] < m ? red[i] : m;
        st->next_min = m;
        if (moved > 0) {
            st->mode = SS_RELAX;
            st->front = (int32_t)moved;
            st->fdeg = fdeg;
            st->pull = moved > st->pull_min;
            st->defer = a.lend != nullptr && !st->pull && fdeg >= a.defer_min;
            st->windows += 1;
        } else if (st->next_min == INT_MAX) {
            st->done = 1;  // nothing pending
        } else {  // empty window: the next one starts at the least pending distance
            st->lo = st->next_min;
            st->hi = ss_window_end(st->next_min, st->delta);
            st->next_min = INT_MAX;
        }
    }
}

struct SsSeed {
    int32_t source;
    int32_t delta;
    int32_t pull_min;
};

__global__ void ss_reset_kernel(SsArgs a, SsSeed sd) {
    const int32_t n = a.V > a.sum_words + 1 ? a.V : a.sum_words + 1;
    const int32_t units = (a.words + kSsPullUnit - 1) / kSsPullUnit;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < a.V) a.dist[i] = i == sd.source ? 0 : INT_MAX;
        if (i < units) a.umin[i] = INT_MAX;  // nothing pending
        if (i < a.words) a.ws[i] = 0;
        if (i <= a.sum_words) a.wsum[i] = 0;
        if (i < a.chunks) a.wsflag[i] = 0;
        if (i < a.words) {
            const uint64_t s = (i == (sd.source >> 6)) ? (1ull << (sd.source & 63)) : 0ull;
            a.bm[0][i] = s;
            a.bm[1][i] = 0;
            a.bm[2][i] = 0;
        }
        if (i <= a.sum_words) {
            const int32_t ch = (sd.source >> 6) / kBsChunkWords;
            a.sum[0][i] = (i == (ch >> 6)) ? (1ull << (ch & 63)) : 0ull;
            a.sum[1][i] = 0;
            a.sum[2][i] = 0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SsState *st = a.st;
        st->done = 0;
        st->mode = SS_RELAX;
        st->phase = 0;
        st->steps = 0;
        st->windows = 0;
        st->slots = 0;
        st->pulls = 0;
        st->hub_count = 0;
        st->lo = 0;
        st->hi = ss_window_end(0, sd.delta);
        st->delta = sd.delta;
        st->front = 1;
        st->next_min = INT_MAX;
        st->pull = 0;
        st->defer = 0;
        st->ws_any = 0;
        st->fdeg = 0;
        st->pull_min = sd.pull_min;
        st->dist = a.dist;
    }
}

__global__ void ss_zero_slots_kernel(int32_t *slots, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) slots[i] = 0;
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
    uint64_t *sum[3] = {nullptr, nullptr, nullptr};
    uint8_t *cflag = nullptr;
    int32_t chunks = 0;
    int32_t *umin = nullptr;
    uint64_t *ws = nullptr, *wsum = nullptr;
    uint8_t *wsflag = nullptr;
    int32_t *snbr = nullptr, *swt = nullptr, *lend = nullptr;  // light/heavy split
    int32_t split_delta = -1;  // delta of the split (-1: none built)
    bool split = false;        // the split has heavy edges and is used
    int32_t *slots = nullptr;
    int32_t sum_words = 0;
    int32_t *host = nullptr;  // pinned: done, mode, phase, steps, windows, slots, pulls
    const int32_t *poff = nullptr, *pnbr = nullptr, *pwt = nullptr;
    int32_t pull_min = 0;
    bool hubs = false;  // some out-degree reaches kSsHubMin: the hub kernel is launched
    int32_t wide_chunks = 0;
    int32_t *chunk_v = nullptr, *chunk_e0 = nullptr;
    std::map<int32_t, hipGraphExec_t> execs;  // by slot count (the run's dist is read from the state)
    int32_t recent_slots[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // slots of the last runs (ring)
    int32_t recent_n = 0;
};

static PlanRegistry<SsWork> g_ss;

static void free_ss(SsWork *w) {
    for (auto &kv : w->execs) (void)hipGraphExecDestroy(kv.second);
    dfree(w->st);
    for (auto p : w->bm) dfree(p);
    for (auto p : w->sum) dfree(p);
    dfree(w->slots);
    dfree(w->cflag);
    dfree(w->umin);
    dfree(w->ws);
    dfree(w->wsum);
    dfree(w->wsflag);
    dfree(w->snbr);
    dfree(w->swt);
    dfree(w->lend);
    dfree(w->chunk_v);
    dfree(w->chunk_e0);
    if (w->host) (void)hipHostFree(w->host);
    delete w;
}

void release_sssp(const gi_graph *g) {
    if (SsWork *p = g_ss.take(g)) free_ss(p);
}

static int ensure_ss(gi_graph *g, SsWork **out) {
    if ((*out = g_ss.find(g)) != nullptr) return GI_OK;
    const int64_t words = bitmap_words(g->V);
    const int64_t sum_words = ceil_div(ceil_div(words, kBsChunkWords), 64);
    if (sum_words > kBsMaxSumWords) return fail(GI_ERR_RANGE, "fused SSSP supports up to 2^26 vertices");
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
    w->sum_words = (int32_t)sum_words;
    w->poff = pg->d_off;
    w->pnbr = pg->d_nbr;
    w->pwt = pg->d_w;
    w->pull_min = g->V / kSsPullDiv;
    int rc = dalloc(&w->st, 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->bm[i], (size_t)words + 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->sum[i], (size_t)sum_words + 1);
    if (rc == GI_OK) rc = dalloc(&w->slots, (kSsMinSlots + 1) * (size_t)kSsMaxGrid);
    if (rc == GI_OK) rc = dalloc(&w->umin, (size_t)ceil_div(words, kSsPullUnit) + 1);
    if (rc == GI_OK) rc = dalloc(&w->ws, (size_t)words + 1);
    if (rc == GI_OK) rc = dalloc(&w->wsum, (size_t)sum_words + 1);
    w->chunks = (int32_t)ceil_div(words, kBsChunkWords);
    const size_t flag_bytes = (size_t)bs_flag_bytes(w->chunks);
    if (rc == GI_OK) rc = dalloc(&w->cflag, flag_bytes);
    if (rc == GI_OK && hipMemset(w->cflag, 0, flag_bytes) != hipSuccess) rc = fail(GI_ERR_HIP, "flag init");
    if (rc == GI_OK) rc = dalloc(&w->wsflag, flag_bytes);
    if (rc == GI_OK && hipMemset(w->wsflag, 0, flag_bytes) != hipSuccess) rc = fail(GI_ERR_HIP, "flag init");
    if (rc == GI_OK && hipHostMalloc((void **)&w->host, 8 * sizeof(int32_t)) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
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
    if (rc == GI_OK) {
        ss_zero_slots_kernel<<<64, 256>>>(w->slots, (kSsMinSlots + 1) * kSsMaxGrid);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(GI_ERR_HIP, "slot init");
    }
    if (rc != GI_OK) {
        free_ss(w);
        return rc;
    }
    g_ss.put(g, w);
    *out = w;
    return GI_OK;
}

// Persistent grid: resident workgroups of the step kernel.
static int ss_grid() {
    static const int per_cu = resident_per_cu((const void *)ss_relax_kernel, kSsBlock);
    int g = device_info().num_cus * per_cu;
    return g < kSsMaxGrid ? g : kSsMaxGrid;
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
    a.poff = w->poff;
    a.pnbr = w->pnbr;
    a.pwt = w->pwt;
    a.dist = dist;
    a.V = g->V;
    a.words = (int32_t)bitmap_words(g->V);
    a.sum_words = w->sum_words;
    for (int i = 0; i < 3; ++i) {
        a.bm[i] = w->bm[i];
        a.sum[i] = w->sum[i];
    }
    a.slots = w->slots;
    a.cflag = w->cflag;
    a.umin = w->umin;
    a.ws = w->ws;
    a.wsum = w->wsum;
    a.wsflag = w->wsflag;
    a.chunks = w->chunks;
    a.wide_chunks = w->wide_chunks;
    a.chunk_v = w->chunk_v;
    a.chunk_e0 = w->chunk_e0;
    return a;
}

// The light/heavy split for `delta` (built on first use of a delta; the
// recorded slots are dropped when the arrays they name change).
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
    if (use != w->split || use) {  // the arrays or their use changed
        for (auto &kv : w->execs) (void)hipGraphExecDestroy(kv.second);
        w->execs.clear();
    }
    w->split = use;
    w->split_delta = delta;
    for (int32_t &k : w->recent_slots) k = 0;  // slot counts of other deltas do not carry over
    return GI_OK;
}

static void launch_slot(const SsArgs &a, bool hubs, Launcher &L) {
    const int grid = ss_grid();
    L(ss_relax_kernel, grid, kSsBlock, a);
    if (hubs) L(ss_hub_kernel, grid, kSsBlock, a);
    L(ss_decide_kernel, 1, kSsDecBlock, a, grid);
    L(ss_farmove_kernel, grid, kSsBlock, a);
    L(ss_decide2_kernel, 1, kSsDecBlock, a, grid);
}

int sssp_fused_run(gi_graph *g, int32_t source, int32_t delta, int32_t *d_dist, int32_t *counts,
                   hipStream_t s) {
    SsWork *w = nullptr;
    GI_TRY(ensure_ss(g, &w));
    GI_TRY(ensure_split(g, w, delta, s));
    SsArgs a = ss_args(g, w, d_dist);
    const int32_t n = g->V > w->sum_words + 1 ? g->V : w->sum_words + 1;
    int rg = (int)ceil_div(n, 256);
    ss_reset_kernel<<<rg < 8192 ? rg : 8192, 256, 0, s>>>(a, SsSeed{source, delta, w->pull_min});
    GI_LAUNCH_CHECK();
    // the first block holds as many slots as the longest of the last runs
    // (rounded up to 8; a slot past the end is five no-op launches, a block
    // too short a host round trip), later ones 16
    auto first_block = [&]() {
        int32_t b = 0;
        for (int32_t k : w->recent_slots) b = k > b ? k : b;
        b = b > 0 ? (b + 7) / 8 * 8 : 16;
        return b < 1024 ? b : 1024;
    };
    SsArgs ca = a;
    ca.dist = nullptr;  // recorded kernels take the run's dist from the state
    auto exec_of = [&](int32_t block, hipGraphExec_t *out) -> int {
        auto it = w->execs.find(block);
        if (it == w->execs.end()) {
            if (w->execs.size() >= 8) {
                for (auto &kv : w->execs) (void)hipGraphExecDestroy(kv.second);
                w->execs.clear();
            }
            hipGraphExec_t exec = nullptr;
            GI_TRY(build_graph(
                [&](Launcher &L) {
                    for (int32_t r = 0; r < block; ++r) launch_slot(ca, w->hubs, L);
                },
                &exec));
            it = w->execs.emplace(block, exec).first;
        }
        *out = it->second;
        return GI_OK;
    };
    int32_t block = first_block();
    for (int64_t used = 0; used < (1ll << 24);) {
        hipGraphExec_t exec = nullptr;
        GI_TRY(exec_of(block, &exec));
        GI_HIP(hipGraphLaunch(exec, s));
        GI_HIP(hipMemcpyAsync(w->host, w->st, 7 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        used += block;
        if (w->host[0]) {  // done
            w->recent_slots[w->recent_n++ % 8] = w->host[5];
            if (counts) {  // steps, pull steps, windows entered, slots
                counts[0] = w->host[3];
                counts[1] = w->host[6];
                counts[2] = w->host[4];
                counts[3] = w->host[5];
            }
            // a longer run than before: record the next run's first block
            // now, so that its launch does not wait on the recording
            hipGraphExec_t next = nullptr;
            return exec_of(first_block(), &next);
        }
        block = 16;
    }
    return fail(GI_ERR_RANGE, "SSSP did not finish");
}

// Eager slot loop for measurement: the launches of every slot bracketed by
// HIP events; the state is read before each slot. The relax launch is
// reported as a push or a pull step by the state's pull flag.
static int sssp_trace(gi_graph *g, int32_t source, int32_t delta, int32_t pull_min, int32_t *d_dist,
                      int32_t max_slots, float *kernel_ms, int64_t *info, int32_t *num_slots, hipStream_t s) {
    SsWork *w = nullptr;
    GI_TRY(ensure_ss(g, &w));
    GI_TRY(ensure_split(g, w, delta, s));
    SsArgs a = ss_args(g, w, d_dist);
    const int32_t n = g->V > w->sum_words + 1 ? g->V : w->sum_words + 1;
    int rg = (int)ceil_div(n, 256);
    ss_reset_kernel<<<rg < 8192 ? rg : 8192, 256, 0, s>>>(a, SsSeed{source, delta, pull_min >= 0 ? pull_min : w->pull_min});
    GI_LAUNCH_CHECK();
    hipEvent_t ev[6] = {};
    int rc = GI_OK;
    for (int i = 0; i < 6 && rc == GI_OK; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = fail(GI_ERR_HIP, "event create");
    const int grid = ss_grid();
    SsState h{};
    int32_t k = 0;
    for (; rc == GI_OK && k < max_slots; ++k) {
        if (hipMemcpyAsync(&h, w->st, offsetof(SsState, hubs), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(GI_ERR_HIP, "state read");
            break;
        }
        if (h.done) break;
        info[5 * k + 0] = h.mode;
        info[5 * k + 1] = h.pull;
        info[5 * k + 2] = h.front;
        info[5 * k + 3] = h.hi;
        info[5 * k + 4] = h.fdeg;
        (void)hipEventRecord(ev[0], s);
        ss_relax_kernel<<<grid, kSsBlock, 0, s>>>(a);
        (void)hipEventRecord(ev[1], s);
        if (w->hubs) ss_hub_kernel<<<grid, kSsBlock, 0, s>>>(a);
        (void)hipEventRecord(ev[2], s);
        ss_decide_kernel<<<1, kSsDecBlock, 0, s>>>(a, grid);
        (void)hipEventRecord(ev[3], s);
        ss_farmove_kernel<<<grid, kSsBlock, 0, s>>>(a);
        (void)hipEventRecord(ev[4], s);
        ss_decide2_kernel<<<1, kSsDecBlock, 0, s>>>(a, grid);
        (void)hipEventRecord(ev[5], s);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(ev[5]) != hipSuccess) {
            rc = fail(GI_ERR_HIP, "slot launch");
            break;
        }
        float t[5];
        for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&t[i], ev[i], ev[i + 1]);
        const bool pulled = h.mode == SS_RELAX && h.pull;
        float *row = kernel_ms + 6 * k;
        row[0] = pulled ? 0.f : t[0];
        row[1] = pulled ? t[0] : 0.f;
        for (int i = 1; i < 5; ++i) row[i + 1] = t[i];
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == GI_OK && k == max_slots && !h.done) rc = fail(GI_ERR_RANGE, "more than %d slots", max_slots);
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
    return gi::sssp_trace(g, source, delta, pull_min, d_dist, max_slots, kernel_ms, info, num_slots,
                          (hipStream_t)stream);
}
