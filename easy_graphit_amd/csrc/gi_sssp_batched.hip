// Batched delta-stepping SSSP: the distances from each of many sources
// (dist[source][vertex]). The SSSP app (apps/sssp.cpp:34-57) runs one source
// per call: dequeue the window [lo, lo + delta) of the priority queue, relax
// its out-edges with writeMin, re-enqueue what improved
// (gpu_priority_queue.h:115-185). Here a pass serves up to 64 sources, one
// wave lane each, laid out vertex-major:
//
//   D[v][64]  the pass's distances of v, one 256-byte row: a relaxation of
//             edge (u, v) reads one row and serves every source at once
//   F[v]      frontier mask (bit b: v is in source b's frontier), three sets
//             in rotation (relaxed, being built, being cleared)
//   W[v]      the window's members (bit b: v was in source b's frontier in
//             this window), two sets, one per window parity
//   P[v]      pending mask (bit b: v improved to a distance beyond the window)
//   Q, QW     the out-lists of F's and W's vertices as items of at most 256
//             edges: the light part (w < delta) for F, the heavy part for W
//
// All sources share one window [lo, lo + delta). Light steps relax F's light
// edges until no source's F' gains a member; then one heavy step relaxes the
// heavy edges of W once, from the settled distances (a heavy edge from the
// window lands at or beyond hi, so it can only add pending members). The
// next window starts at the least pending distance over all sources (the
// reference's queue starts a window at the least remaining priority too,
// gpu_priority_queue.h:126-146) and the pending members inside it move to F.
// A source with no member in a window idles through it. A step pushes (a
// wave per item: writeMin of D[v][b] for the sources in the item's mask) or,
// when its lists hold more than an eighth of the edges of their kind, pulls
// (a wave per 64 vertices: each vertex takes the minimum over its
// in-neighbours u of D[u][b] + w for the b in u's mask and stores it; no
// atomics on D). Shortest-path distances are unique, so neither the order of
// relaxations nor the window schedule changes them. A step is four launches
// recorded in hipGraph blocks; the host reads the done flag once per block.
#include <climits>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kVbLanes = 64;         // sources per pass
constexpr int kVbBlock = 256;
constexpr int kVbSeg = 256;          // out-edges per list item
constexpr int kVbPullWide = 1024;    // in-lists this long are pulled in chunks
constexpr int kVbPullChunk = 1024;   // in-edges per chunk item
constexpr int kVbBlockSteps = 16;
constexpr int kVbPullDiv = 8;        // a step pulls when its lists hold more than 1 / kVbPullDiv of the edges
constexpr int kVbBatch = 4;          // rows in flight per wave

enum VbMode : int32_t { VB_LIGHT = 0, VB_HEAVY = 1 };

struct VbState {
    int32_t done;
    int32_t step;     // F = f[step % 3]
    int32_t adv;      // the window is settled: open the next one
    int32_t mode;     // the coming relax step: light edges of F, or heavy edges of W
    int32_t lo, hi, delta;
    int32_t pmin;     // least pending distance (advance)
    int32_t steps, pulls, windows;
    int32_t wi;       // window index: W = w[wi & 1]
    int32_t wclr;     // the coming relax step clears the other W set
    int32_t qn[3];    // F items per set
    int32_t wn[2];    // W items per set
    int32_t pad;
    long long qe[3];  // edges of the F items per set
    long long we[2];  // edges of the W items per set
};

// An item list and its counters.
struct VbList {
    int4 *q;  // (vertex, first edge, end edge)
    int32_t *n;
    long long *e;
};

struct VbArgs {
    VbState *st;
    const int32_t *off, *nbr, *wt;     // out-lists (light part first when split)
    const int32_t *lend;               // end of each light part (null: every edge is light)
    const int32_t *poff, *pnbr, *pwt;  // in-lists
    const int32_t *plend;              // end of each light in-part (null: select by weight)
    int32_t V;
    int32_t qcap;
    long long pull_light, pull_heavy;  // lists with more edges than these pull
    int32_t *D;  // [V][64]
    uint64_t *f[3];
    uint64_t *w[2];
    uint64_t *p;
    int4 *q[3];
    int4 *qw[2];
    const int4 *wide;  // chunks of the long in-lists (vertex, first edge, end edge)
    int32_t nwide;
};

__device__ __forceinline__ int32_t vb_wave() { return (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int32_t vb_waves() { return (int32_t)((gridDim.x * blockDim.x) >> 6); }
__device__ __forceinline__ int32_t *vb_row(const VbArgs &a, int32_t v) { return a.D + (int64_t)v * kVbLanes; }
__device__ __forceinline__ uint64_t vb_readlane64(uint64_t x, int32_t j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)x, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(x >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ VbList vb_flist(const VbArgs &a, int32_t k) { return VbList{a.q[k], &a.st->qn[k], &a.st->qe[k]}; }
__device__ __forceinline__ VbList vb_wlist(const VbArgs &a, int32_t k) { return VbList{a.qw[k], &a.st->wn[k], &a.st->we[k]}; }

// The light or heavy part of v's out-list.
__device__ __forceinline__ void vb_part(const VbArgs &a, int32_t v, bool heavy, int32_t &e0, int32_t &e1) {
    const int32_t o0 = a.off[v], o1 = a.off[v + 1];
    const int32_t l = a.lend ? a.lend[v] : o1;
    e0 = heavy ? l : o0;
    e1 = heavy ? o1 : l;
}

// v joined the mask set of list L: its part of the out-list joins L as items
// of at most kVbSeg edges (one empty item for an empty part, so that the
// clearing pass finds every vertex with a mask). One lane calls.
__device__ void vb_enqueue(const VbArgs &a, const VbList &L, bool heavy, int32_t v) {
    int32_t e0, e1;
    vb_part(a, v, heavy, e0, e1);
    int32_t n = (e1 - e0 + kVbSeg - 1) / kVbSeg;
    if (n == 0) n = 1;
    const int32_t p = atomicAdd(L.n, n);
    for (int32_t j = 0; j < n && p + j < a.qcap; ++j) {
        const int32_t s0 = e0 + j * kVbSeg;
        L.q[p + j] = make_int4(v, s0, s0 + kVbSeg < e1 ? s0 + kVbSeg : e1, 0);
    }
    if (e1 > e0) atomicAdd((unsigned long long *)L.e, (unsigned long long)(e1 - e0));
}

// The same for the vertices held by the lanes in `mask` (lane j holds vj),
// with one pair of counter atomics for the wave. The whole wave calls.
__device__ void vb_enqueue_wave(const VbArgs &a, const VbList &L, bool heavy, uint64_t mask, int32_t vj) {
    if (!mask) return;  // uniform
    const int32_t lane = lane_id();
    const bool on = (mask >> lane) & 1ull;
    int32_t e0 = 0, e1 = 0;
    if (on) vb_part(a, vj, heavy, e0, e1);
    int32_t n = on ? (e1 - e0 + kVbSeg - 1) / kVbSeg : 0;
    if (on && n == 0) n = 1;
    int32_t incl = n;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    long long deg = e1 - e0;
    for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o);
    const int32_t total = __builtin_amdgcn_readlane(incl, 63);  // read with the whole wave active
    int32_t base = 0;
    if (lane == 0) {
        base = atomicAdd(L.n, total);
        if (deg) atomicAdd((unsigned long long *)L.e, (unsigned long long)deg);
    }
    base = __shfl(base, 0);
    for (int32_t j = 0, p = base + incl - n; j < n && p + j < a.qcap; ++j) {
        const int32_t s0 = e0 + j * kVbSeg;
        L.q[p + j] = make_int4(vj, s0, s0 + kVbSeg < e1 ? s0 + kVbSeg : e1, 0);
    }
}

// Where the improvements of a relax step go.
struct VbOut {
    int32_t nxt;  // F' = f[nxt]
    int32_t cw;   // W = w[cw]
    int32_t hi;
};

// Improvements of v found by a writeMin (one lane calls): below hi they join
// F' and W, beyond it P.
__device__ void vb_mark_atomic(const VbArgs &a, const VbOut &o, int32_t v, uint64_t near, uint64_t far) {
    if (near) {
        if (!atomicOr((unsigned long long *)&a.f[o.nxt][v], near)) vb_enqueue(a, vb_flist(a, o.nxt), false, v);
        if (a.lend && (a.w[o.cw][v] & near) != near &&
            !atomicOr((unsigned long long *)&a.w[o.cw][v], near))
            vb_enqueue(a, vb_wlist(a, o.cw), true, v);
    }
    if (far && (a.p[v] & far) != far) atomicOr((unsigned long long *)&a.p[v], far);
}

// Push over the items of list L with masks mk: a wave per item, lane b
// relaxes for source b when b is in mk[u]; kVbBatch target rows are read and
// written at a time. The vertices that join F' or W are listed 64 at a time.
__device__ void vb_push(const VbArgs &a, const VbList &L, const uint64_t *mk, const VbOut &o) {
    const int32_t n = *L.n < a.qcap ? *L.n : a.qcap;
    const int32_t lane = lane_id();
    int32_t fj = 0, nf = 0, wj = 0, nw = 0;  // pending listings (lane j holds the j-th)
    const VbList FL = vb_flist(a, o.nxt), WL = vb_wlist(a, o.cw);
    for (int32_t i = vb_wave(); i < n; i += vb_waves()) {
        const int4 it = L.q[i];
        const int32_t u = it.x;
        const uint64_t m = mk[u];
        const bool mine = (m >> lane) & 1ull;
        const int32_t du = mine ? vb_row(a, u)[lane] : 0;
        for (int32_t e0 = it.y; e0 < it.z; e0 += 64) {
            const int32_t ne = it.z - e0 < 64 ? it.z - e0 : 64;
            const int32_t vl = lane < ne ? a.nbr[e0 + lane] : 0;
            const int32_t wl = lane < ne ? a.wt[e0 + lane] : 0;
            for (int32_t j = 0; j < ne; j += kVbBatch) {
                int32_t v[kVbBatch], nd[kVbBatch], dv[kVbBatch], old[kVbBatch];
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    const int32_t jj = j + k < ne ? j + k : ne - 1;
                    v[k] = j + k < ne ? __builtin_amdgcn_readlane(vl, jj) : -1;
                    nd[k] = du + __builtin_amdgcn_readlane(wl, jj);
                }
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) dv[k] = mine && v[k] >= 0 ? vb_row(a, v[k])[lane] : INT_MIN;
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    old[k] = INT_MIN;
                    if (nd[k] < dv[k]) old[k] = atomicMin(&vb_row(a, v[k])[lane], nd[k]);
                }
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    const bool imp = nd[k] < old[k];
                    const uint64_t near = __ballot(imp && nd[k] < o.hi), far = __ballot(imp && nd[k] >= o.hi);
                    if (!(near | far)) continue;  // uniform
                    unsigned long long fo = 1, wo = 1;
                    if (lane == 0) {
                        if (near) {
                            fo = atomicOr((unsigned long long *)&a.f[o.nxt][v[k]], near);
                            if (a.lend && (a.w[o.cw][v[k]] & near) != near)
                                wo = atomicOr((unsigned long long *)&a.w[o.cw][v[k]], near);
                        }
                        if (far && (a.p[v[k]] & far) != far) atomicOr((unsigned long long *)&a.p[v[k]], far);
                    }
                    if (!vb_readlane64(fo, 0)) {  // uniform: v joined F' here
                        if (lane == nf) fj = v[k];
                        if (++nf == 64) {
                            vb_enqueue_wave(a, FL, false, ~0ull, fj);
                            nf = 0;
                        }
                    }
                    if (!vb_readlane64(wo, 0)) {  // uniform: v joined W here
                        if (lane == nw) wj = v[k];
                        if (++nw == 64) {
                            vb_enqueue_wave(a, WL, true, ~0ull, wj);
                            nw = 0;
                        }
                    }
                }
            }
        }
    }
    vb_enqueue_wave(a, FL, false, nf == 64 ? ~0ull : (1ull << nf) - 1, fj);
    vb_enqueue_wave(a, WL, true, nw == 64 ? ~0ull : (1ull << nw) - 1, wj);
}

// The in-edges a step reads: the light or heavy part through plend, or the
// whole list (then the weight test in the pull picks the kind).
__device__ __forceinline__ void vb_in_part(const VbArgs &a, int32_t v, bool heavy, int32_t &p0, int32_t &p1) {
    p0 = a.poff[v];
    p1 = a.poff[v + 1];
    if (a.plend) {
        const int32_t l = a.plend[v];
        if (heavy) p0 = l;
        else p1 = l;
    }
}

// The minimum over in-edges [e, e + cnt) of D[u][lane] + w for the sources
// in mk[u], folded into cand (sel: 0 every edge, 1 light only, 2 heavy
// only); used by the chunk items of long lists.
__device__ __forceinline__ int32_t vb_pull_range(const VbArgs &a, const uint64_t *mk, int32_t sel, int32_t delta,
                                                 int32_t e, int32_t cnt, int32_t cand) {
    const int32_t lane = lane_id();
    const int32_t uj = lane < cnt ? a.pnbr[e + lane] : 0;
    const int32_t wj = lane < cnt ? a.pwt[e + lane] : 0;
    const bool kind = sel == 0 || (sel == 1 ? wj < delta : wj >= delta);
    const uint64_t mj = lane < cnt && kind ? mk[uj] : 0ull;
    uint64_t act = __ballot(mj != 0);
    while (act) {
        int32_t j[kVbBatch], du[kVbBatch], w[kVbBatch];
        bool on[kVbBatch];
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k) {
            j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
            act &= act - 1;
        }
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k) {
            const int32_t jj = j[k] >= 0 ? j[k] : 0;
            const int32_t u = __builtin_amdgcn_readlane(uj, jj);
            const uint64_t m = vb_readlane64(mj, jj);
            w[k] = __builtin_amdgcn_readlane(wj, jj);
            on[k] = j[k] >= 0 && ((m >> lane) & 1ull);
            du[k] = on[k] ? vb_row(a, u)[lane] : 0;
        }
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k)
            if (on[k] && du[k] + w[k] < cand) cand = du[k] + w[k];
    }
    return cand;
}

// The pull step's result for vertex v (owned by this wave; no other wave
// writes its row or masks in a pull step): store the improved distances and
// set its F', W and pending bits. Bit 0 of the result: v joined F'; bit 1: W.
__device__ __forceinline__ int32_t vb_pull_store(const VbArgs &a, const VbOut &o, int32_t v, int32_t cand) {
    const int32_t lane = lane_id();
    int32_t *row = vb_row(a, v);
    const bool imp = cand < row[lane];
    if (imp) row[lane] = cand;
    const uint64_t near = __ballot(imp && cand < o.hi), far = __ballot(imp && cand >= o.hi);
    uint64_t w0 = 1;
    if (lane == 0) {
        if (near) {
            a.f[o.nxt][v] = near;  // F' was clear
            if (a.lend) {
                w0 = a.w[o.cw][v];
                a.w[o.cw][v] = w0 | near;
            }
        }
        if (far) a.p[v] |= far;
    }
    return (near ? 1 : 0) | (near && !vb_readlane64(w0, 0) ? 2 : 0);
}

// Pull over every vertex: a wave takes 64 consecutive vertices, reads their
// in-lists 64 edges at a time (lane j takes edge j of the concatenation),
// and for each edge whose source u has a nonzero mask reads D[u] once for
// all sources. In-lists of kVbPullWide edges or more are chunk items below.
__device__ void vb_pull(const VbArgs &a, const uint64_t *mk, bool heavy, const VbOut &o) {
    const int32_t lane = lane_id();
    const int32_t delta = a.st->delta;
    const int32_t sel = a.plend || !a.lend ? 0 : (heavy ? 2 : 1);
    const VbList FL = vb_flist(a, o.nxt), WL = vb_wlist(a, o.cw);
    const int32_t groups = (a.V + 63) / 64;
    for (int32_t gi = vb_wave(); gi < groups; gi += vb_waves()) {
        const int32_t v = gi * 64 + lane;
        int32_t p0 = 0, p1 = 0;
        if (v < a.V && a.poff[v + 1] - a.poff[v] < kVbPullWide) vb_in_part(a, v, heavy, p0, p1);
        const int32_t len = p1 - p0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        int32_t cv = -1, cand = INT_MAX;  // the vertex being gathered (lane index) and its minimum
        uint64_t jf = 0, jw = 0;          // the group's vertices that joined F' and W
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;  // owner: the last lane whose exclusive prefix is <= g
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(p0, ow) + (g - __shfl(pre, ow));
            const bool has = g < total;
            const int32_t uj = has ? a.pnbr[e] : 0;
            const int32_t wj = has ? a.pwt[e] : 0;
            const bool kind = sel == 0 || (sel == 1 ? wj < delta : wj >= delta);
            const uint64_t mj = has && kind ? mk[uj] : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                int32_t j[kVbBatch], du[kVbBatch], w[kVbBatch], own[kVbBatch];
                bool on[kVbBatch];
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                }
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    const int32_t jj = j[k] >= 0 ? j[k] : 0;
                    const int32_t u = __builtin_amdgcn_readlane(uj, jj);
                    const uint64_t m = vb_readlane64(mj, jj);
                    w[k] = __builtin_amdgcn_readlane(wj, jj);
                    own[k] = j[k] >= 0 ? __builtin_amdgcn_readlane(ow, jj) : -1;
                    on[k] = j[k] >= 0 && ((m >> lane) & 1ull);
                    du[k] = on[k] ? vb_row(a, u)[lane] : 0;
                }
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    if (own[k] < 0) continue;  // uniform
                    if (own[k] != cv) {        // uniform: edges come in vertex order
                        if (cv >= 0) {
                            const int32_t r = vb_pull_store(a, o, gi * 64 + cv, cand);
                            jf |= (uint64_t)(r & 1) << cv;
                            jw |= (uint64_t)(r >> 1) << cv;
                        }
                        cv = own[k];
                        cand = INT_MAX;
                    }
                    if (on[k] && du[k] + w[k] < cand) cand = du[k] + w[k];
                }
            }
        }
        if (cv >= 0) {
            const int32_t r = vb_pull_store(a, o, gi * 64 + cv, cand);
            jf |= (uint64_t)(r & 1) << cv;
            jw |= (uint64_t)(r >> 1) << cv;
        }
        vb_enqueue_wave(a, FL, false, jf, v);
        vb_enqueue_wave(a, WL, true, jw, v);
    }
    // chunks of the long in-lists: several waves share a vertex, so the
    // result joins D by writeMin and the masks by atomicOr
    for (int32_t i = vb_wave(); i < a.nwide; i += vb_waves()) {
        const int4 c = a.wide[i];
        int32_t p0, p1;
        vb_in_part(a, c.x, heavy, p0, p1);
        const int32_t e0 = c.y > p0 ? c.y : p0, e1 = c.z < p1 ? c.z : p1;
        int32_t cand = INT_MAX;
        for (int32_t e = e0; e < e1; e += 64) cand = vb_pull_range(a, mk, sel, delta, e, e1 - e < 64 ? e1 - e : 64, cand);
        bool imp = false;
        if (cand != INT_MAX) imp = cand < atomicMin(&vb_row(a, c.x)[lane], cand);
        const uint64_t near = __ballot(imp && cand < o.hi), far = __ballot(imp && cand >= o.hi);
        if (lane == 0 && (near | far)) vb_mark_atomic(a, o, c.x, near, far);
    }
}

// One relax step: clear the F set two steps back (and, at a window's first
// step, the last window's W set; their vertices are their items), then push
// or pull the light edges of F or the heavy edges of W. Every workgroup
// reads the same counts, so all take the same branch.
__global__ void __launch_bounds__(kVbBlock) vb_relax_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done) return;  // uniform
    const int32_t s = st->step, cur = s % 3, nxt = (s + 1) % 3, prv = (s + 2) % 3;
    const int32_t cw = st->wi & 1;
    const int32_t t0 = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x), nt = (int32_t)(gridDim.x * blockDim.x);
    const int32_t nprev = st->qn[prv] < a.qcap ? st->qn[prv] : a.qcap;
    for (int32_t i = t0; i < nprev; i += nt) a.f[prv][a.q[prv][i].x] = 0;
    if (st->wclr) {
        const int32_t ow = cw ^ 1, nw = st->wn[ow] < a.qcap ? st->wn[ow] : a.qcap;
        for (int32_t i = t0; i < nw; i += nt) a.w[ow][a.qw[ow][i].x] = 0;
    }
    const VbOut o{nxt, cw, st->hi};
    if (st->mode == VB_HEAVY) {
        if (st->we[cw] > a.pull_heavy) vb_pull(a, a.w[cw], true, o);
        else vb_push(a, vb_wlist(a, cw), a.w[cw], o);
    } else {
        if (st->qe[cur] > a.pull_light) vb_pull(a, a.f[cur], false, o);
        else vb_push(a, vb_flist(a, cur), a.f[cur], o);
    }
}

// After the relax step: while F' gains members the window's light steps go
// on; then (with a split) one heavy step; then the window is settled.
__global__ void vb_decide_kernel(VbArgs a) {
    VbState *st = a.st;
    if (threadIdx.x != 0 || st->done) return;
    const int32_t s = st->step, cur = s % 3, nxt = (s + 1) % 3, prv = (s + 2) % 3;
    const int32_t cw = st->wi & 1;
    st->steps += 1;
    st->pulls += st->mode == VB_HEAVY ? st->we[cw] > a.pull_heavy : st->qe[cur] > a.pull_light;
    if (st->wclr) {  // the relax step cleared the other W set
        st->wn[cw ^ 1] = 0;
        st->we[cw ^ 1] = 0;
        st->wclr = 0;
    }
    st->adv = 0;
    if (st->qn[nxt] > 0) {
        st->mode = VB_LIGHT;
    } else if (st->mode == VB_LIGHT && a.lend) {
        st->mode = VB_HEAVY;
    } else {
        st->mode = VB_LIGHT;
        st->adv = 1;
        st->pmin = INT_MAX;
        st->windows += 1;
        st->wi += 1;
        st->wclr = a.lend != nullptr;
    }
    st->qn[prv] = 0;
    st->qe[prv] = 0;
    st->step = s + 1;
}

// Advance, first pass: the least pending distance over all sources; pending
// bits whose distance fell below hi are stale (those vertices were settled
// through F) and are dropped.
__global__ void __launch_bounds__(kVbBlock) vb_pmin_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done || !st->adv) return;  // uniform
    const int32_t lane = lane_id(), hi = st->hi;
    const int32_t groups = (a.V + 63) / 64;
    int32_t mn = INT_MAX;
    for (int32_t gi = vb_wave(); gi < groups; gi += vb_waves()) {
        const int32_t v0 = gi * 64;
        const uint64_t pj = v0 + lane < a.V ? a.p[v0 + lane] : 0ull;
        uint64_t act = __ballot(pj != 0);
        while (act) {
            int32_t j[kVbBatch], d[kVbBatch];
            uint64_t m[kVbBatch];
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
                act &= act - 1;
                m[k] = j[k] >= 0 ? vb_readlane64(pj, j[k]) : 0ull;
                d[k] = (m[k] >> lane) & 1ull ? vb_row(a, v0 + j[k])[lane] : INT_MAX;
            }
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                if (j[k] < 0) continue;  // uniform
                const bool keep = d[k] != INT_MAX && d[k] >= hi;
                mn = keep && d[k] < mn ? d[k] : mn;
                const uint64_t km = __ballot(keep);
                if (lane == 0 && km != m[k]) a.p[v0 + j[k]] = km;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t y = __shfl_xor(mn, o);
        mn = y < mn ? y : mn;
    }
    if (lane == 0 && mn != INT_MAX) atomicMin(&st->pmin, mn);
}

// Advance, second pass: the window opens at the least pending distance (none:
// done) and the pending members inside it move to F and W.
__global__ void __launch_bounds__(kVbBlock) vb_move_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done || !st->adv) return;  // uniform
    const int32_t m0 = st->pmin;
    if (m0 == INT_MAX) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
        return;
    }
    const int32_t hi = m0 > INT_MAX - st->delta ? INT_MAX : m0 + st->delta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // read by the next relax step only
        st->lo = m0;
        st->hi = hi;
    }
    const int32_t cur = st->step % 3, cw = st->wi & 1, lane = lane_id();
    const VbList FL = vb_flist(a, cur), WL = vb_wlist(a, cw);
    const int32_t groups = (a.V + 63) / 64;
    for (int32_t gi = vb_wave(); gi < groups; gi += vb_waves()) {
        const int32_t v0 = gi * 64;
        const uint64_t pj = v0 + lane < a.V ? a.p[v0 + lane] : 0ull;
        uint64_t act = __ballot(pj != 0), joined = 0;
        while (act) {
            int32_t j[kVbBatch], d[kVbBatch];
            uint64_t m[kVbBatch];
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
                act &= act - 1;
                m[k] = j[k] >= 0 ? vb_readlane64(pj, j[k]) : 0ull;
                d[k] = (m[k] >> lane) & 1ull ? vb_row(a, v0 + j[k])[lane] : INT_MAX;
            }
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                if (j[k] < 0) continue;  // uniform
                const uint64_t in = __ballot(d[k] < hi);
                if (!in) continue;  // uniform
                joined |= 1ull << j[k];
                if (lane == 0) {  // F and this window's W are clear at an advance: this wave sets v's masks
                    const int32_t v = v0 + j[k];
                    a.p[v] = m[k] & ~in;
                    a.f[cur][v] = in;
                    if (a.lend) a.w[cw][v] = in;
                }
            }
        }
        vb_enqueue_wave(a, FL, false, joined, v0 + lane);
        if (a.lend) vb_enqueue_wave(a, WL, true, joined, v0 + lane);
    }
}

__global__ void vb_reset_kernel(VbArgs a) {
    const int64_t nd = (int64_t)a.V * kVbLanes;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += (int64_t)gridDim.x * blockDim.x) {
        a.D[i] = INT_MAX;
        if (i < a.V) {
            a.f[0][i] = 0;
            a.f[1][i] = 0;
            a.f[2][i] = 0;
            a.w[0][i] = 0;
            a.w[1][i] = 0;
            a.p[i] = 0;
        }
    }
}

// One thread per source of the pass (sources may repeat: each has its lane).
__global__ void vb_seed_kernel(VbArgs a, const int32_t *src, int32_t B, int32_t delta) {
    VbState *st = a.st;
    const int32_t b = threadIdx.x;
    if (b == 0) {
        st->done = 0;
        st->step = 0;
        st->adv = 0;
        st->mode = VB_LIGHT;
        st->lo = 0;
        st->hi = delta;
        st->delta = delta;
        st->pmin = INT_MAX;
        st->steps = st->pulls = st->windows = 0;
        st->wi = 0;
        st->wclr = 0;
        for (int k = 0; k < 3; ++k) {
            st->qn[k] = 0;
            st->qe[k] = 0;
        }
        for (int k = 0; k < 2; ++k) {
            st->wn[k] = 0;
            st->we[k] = 0;
        }
    }
    __syncthreads();
    if (b < B) {
        const int32_t s = src[b];
        vb_row(a, s)[b] = 0;
        if (!atomicOr((unsigned long long *)&a.f[0][s], 1ull << b)) vb_enqueue(a, vb_flist(a, 0), false, s);
        if (a.lend && !atomicOr((unsigned long long *)&a.w[0][s], 1ull << b)) vb_enqueue(a, vb_wlist(a, 0), true, s);
    }
}

// dist[b][v] = D[v][b] for the pass's B sources, through a 64 x 64 LDS tile.
__global__ void __launch_bounds__(kVbBlock) vb_out_kernel(VbArgs a, int32_t *dist, int32_t B) {
    __shared__ int32_t tile[64][65];
    const int32_t tiles = (a.V + 63) / 64;
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    for (int32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int32_t v0 = t * 64;
        for (int32_t r = wid; r < 64; r += kVbBlock / 64)
            tile[r][lane] = v0 + r < a.V ? vb_row(a, v0 + r)[lane] : 0;
        __syncthreads();
        for (int32_t b = wid; b < B; b += kVbBlock / 64)
            if (v0 + lane < a.V) dist[(int64_t)b * a.V + v0 + lane] = tile[lane][b];
        __syncthreads();
    }
}

// Chunk items of the in-lists of kVbPullWide edges or more.
__global__ void vb_wide_kernel(const int32_t *poff, int32_t V, int4 *out, int32_t cap, unsigned long long *n) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int32_t e0 = poff[v], e1 = poff[v + 1];
        if (e1 - e0 < kVbPullWide) continue;
        const int32_t k = (e1 - e0 + kVbPullChunk - 1) / kVbPullChunk;
        const int32_t p = (int32_t)atomicAdd(n, (unsigned long long)k);
        for (int32_t j = 0; j < k && p + j < cap; ++j) {
            const int32_t s0 = e0 + j * kVbPullChunk;
            out[p + j] = make_int4(v, s0, s0 + kVbPullChunk < e1 ? s0 + kVbPullChunk : e1, 0);
        }
    }
}

// Light edges of the out-lists (the sum of lend[v] - off[v]).
__global__ void vb_light_count_kernel(const int32_t *off, const int32_t *lend, int32_t V, unsigned long long *n) {
    unsigned long long c = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) c += lend[v] - off[v];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0 && c) atomicAdd(n, c);
}

static void vb_launch_step(const VbArgs &a, Launcher &L) {
    const int grid = persistent_grid(kVbBlock, 8);
    L(vb_relax_kernel, grid, kVbBlock, a);
    L(vb_decide_kernel, 1, 64, a);
    L(vb_pmin_kernel, grid, kVbBlock, a);
    L(vb_move_kernel, grid, kVbBlock, a);
}

struct VbRun {
    int32_t *D = nullptr;
    uint64_t *f[3] = {nullptr, nullptr, nullptr};
    uint64_t *w[2] = {nullptr, nullptr};
    uint64_t *p = nullptr;
    int4 *q[3] = {nullptr, nullptr, nullptr};
    int4 *qw[2] = {nullptr, nullptr};
    int4 *wide = nullptr;
    unsigned long long *cnt = nullptr;  // wide chunks, light edges
    VbState *st = nullptr;
    VbState *host = nullptr;  // pinned
    int32_t *d_src = nullptr;
    hipGraphExec_t exec = nullptr;
    ~VbRun() {
        if (exec) (void)hipGraphExecDestroy(exec);
        dfree(D);
        for (auto x : f) dfree(x);
        for (auto x : w) dfree(x);
        for (auto x : q) dfree(x);
        for (auto x : qw) dfree(x);
        dfree(p);
        dfree(wide);
        dfree(cnt);
        dfree(st);
        dfree(d_src);
        if (host) (void)hipHostFree(host);
    }
};

static int sssp_batched(gi_graph *g, const int32_t *h_src, int32_t n, int32_t delta, int32_t batch,
                        int32_t *d_dist, gi_run_stats *stats, hipStream_t s) {
    const gi_graph *pg = g;  // in-lists for pull steps
    if (g->directed) {
        gi_graph *t = nullptr;
        GI_TRY(gi_graph_transpose(g, &t));
        pg = t;
    }
    if (!pg->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    SsSplitView sv;
    GI_TRY(sssp_light_split(g, delta, &sv, s));
    VbRun r;
    const int64_t V = g->V;
    // a set lists each vertex once: max(1, ceil(part / kVbSeg)) items
    const int64_t qcap = 2 * V + g->E / kVbSeg + 1;
    if (qcap > INT32_MAX) return fail(GI_ERR_RANGE, "batched SSSP: graph too large");
    GI_TRY(dalloc(&r.D, (size_t)V * kVbLanes + 1));
    for (int k = 0; k < 3; ++k) {
        GI_TRY(dalloc(&r.f[k], (size_t)V + 1));
        GI_TRY(dalloc(&r.q[k], (size_t)qcap));
    }
    for (int k = 0; k < 2; ++k) {
        GI_TRY(dalloc(&r.w[k], (size_t)V + 1));
        GI_TRY(dalloc(&r.qw[k], (size_t)qcap));
    }
    GI_TRY(dalloc(&r.p, (size_t)V + 1));
    GI_TRY(dalloc(&r.st, 1));
    GI_TRY(dalloc(&r.d_src, (size_t)kVbLanes));
    GI_TRY(dalloc(&r.cnt, 2));
    const int64_t wcap64 = 2 * g->E / kVbPullWide + 1;
    const int32_t wcap = (int32_t)(wcap64 < INT32_MAX ? wcap64 : INT32_MAX);
    GI_TRY(dalloc(&r.wide, (size_t)wcap));
    if (hipHostMalloc((void **)&r.host, sizeof(VbState)) != hipSuccess) return fail(GI_ERR_OOM, "pinned allocation failed");
    GI_HIP(hipMemsetAsync(r.cnt, 0, 2 * sizeof(unsigned long long), s));
    const int vgrid = (int)(ceil_div(V, 256) < 4096 ? ceil_div(V, 256) : 4096);
    if (V) vb_wide_kernel<<<vgrid, 256, 0, s>>>(pg->d_off, g->V, r.wide, wcap, r.cnt);
    if (V && sv.lend) vb_light_count_kernel<<<vgrid, 256, 0, s>>>(g->d_off, sv.lend, g->V, r.cnt + 1);
    GI_LAUNCH_CHECK();
    unsigned long long hc[2] = {0, 0};
    GI_HIP(hipMemcpyAsync(hc, r.cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    const int64_t nwide = (int64_t)hc[0];
    const long long light = sv.lend ? (long long)hc[1] : g->E;
    VbArgs a{};
    a.st = r.st;
    a.off = g->d_off;
    a.nbr = sv.lend ? sv.nbr : g->d_nbr;
    a.wt = sv.lend ? sv.wt : g->d_w;
    a.lend = sv.lend;
    a.poff = pg->d_off;
    if (g->directed || !sv.lend) {  // whole in-lists, the kind picked by weight
        a.pnbr = pg->d_nbr;
        a.pwt = pg->d_w;
        a.plend = nullptr;
    } else {  // symmetric: the split out-lists are the split in-lists
        a.pnbr = sv.nbr;
        a.pwt = sv.wt;
        a.plend = sv.lend;
    }
    a.V = g->V;
    a.qcap = (int32_t)qcap;
    const bool whole = a.plend == nullptr;  // a pull reads every in-edge
    a.pull_light = (whole ? g->E : light) / kVbPullDiv;
    a.pull_heavy = (whole ? g->E : g->E - light) / kVbPullDiv;
    a.D = r.D;
    for (int k = 0; k < 3; ++k) {
        a.f[k] = r.f[k];
        a.q[k] = r.q[k];
    }
    for (int k = 0; k < 2; ++k) {
        a.w[k] = r.w[k];
        a.qw[k] = r.qw[k];
    }
    a.p = r.p;
    a.wide = r.wide;
    a.nwide = (int32_t)(nwide < wcap ? nwide : wcap);
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t i = 0; i < kVbBlockSteps; ++i) vb_launch_step(a, L);
        },
        &r.exec));
    const int rgrid = persistent_grid(256);
    EventTimer timer(s);
    long long steps = 0, pulls = 0;
    for (int32_t g0 = 0; g0 < n; g0 += batch) {
        const int32_t B = n - g0 < batch ? n - g0 : batch;
        GI_HIP(hipMemcpyAsync(r.d_src, h_src + g0, sizeof(int32_t) * B, hipMemcpyHostToDevice, s));
        vb_reset_kernel<<<rgrid, 256, 0, s>>>(a);
        vb_seed_kernel<<<1, kVbLanes, 0, s>>>(a, r.d_src, B, delta);
        GI_LAUNCH_CHECK();
        for (int64_t used = 0;; used += kVbBlockSteps) {
            // a safety net: every window settles a pending member of some source
            if (used > 4 * (V + 64) * kVbLanes) return fail(GI_ERR_HIP, "batched SSSP never reported done");
            GI_HIP(hipGraphLaunch(r.exec, s));
            GI_HIP(hipMemcpyAsync(r.host, r.st, sizeof(VbState), hipMemcpyDeviceToHost, s));
            GI_HIP(hipStreamSynchronize(s));
            if (r.host->done) break;
        }
        steps += r.host->steps;
        pulls += r.host->pulls;
        vb_out_kernel<<<rgrid, kVbBlock, 0, s>>>(a, d_dist + (int64_t)g0 * V, B);
        GI_LAUNCH_CHECK();
    }
    const double ms = timer.stop();
    if (stats) {
        *stats = gi_run_stats{};
        stats->iterations = (int32_t)(steps < INT32_MAX ? steps : INT32_MAX);
        stats->pull_iterations = (int32_t)pulls;
        stats->push_iterations = stats->iterations - stats->pull_iterations;
        stats->ms = ms;
    }
    return GI_OK;
}

}  // namespace gi

using namespace gi;

extern "C" int gi_sssp_batched(gi_graph *g, const int32_t *sources, int32_t num_sources, int32_t delta,
                               int32_t batch, int32_t *d_dist, gi_run_stats *stats, void *stream) {
    if (!g || (num_sources > 0 && (!sources || !d_dist))) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (!g->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    if (delta < 1) return fail(GI_ERR_RANGE, "delta must be positive");
    if (num_sources < 0) return fail(GI_ERR_RANGE, "negative source count");
    for (int32_t i = 0; i < num_sources; ++i)
        if (sources[i] < 0 || sources[i] >= g->V)
            return fail(GI_ERR_RANGE, "source %d out of range [0,%d)", sources[i], g->V);
    if (num_sources == 0) {
        if (stats) *stats = gi_run_stats{};
        return GI_OK;
    }
    if (batch < 1 || batch > kVbLanes) batch = kVbLanes;
    if (batch > num_sources) batch = num_sources;
    return sssp_batched(g, sources, num_sources, delta, batch, d_dist, stats, (hipStream_t)stream);
}
