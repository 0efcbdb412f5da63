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
//
// All sources share one window [lo, lo + delta). Light steps relax F's light
// edges (w < delta) until no source's F' gains a member; then one heavy step
// relaxes the heavy edges of W once, from the settled distances (a heavy edge
// from the window lands at or beyond hi, so it can only add pending
// members). The next window starts at the least pending distance over all
// sources (the reference's queue starts a window at the least remaining
// priority too, gpu_priority_queue.h:126-146) and the pending members inside
// it move to F. A source with no member in a window idles through it.
//
// A step pushes or, when its vertices hold more than an eighth of the edges
// of their kind, pulls. A push sweeps the mask array 64 vertices per wave
// and relaxes the edges of the vertices found (writeMin of D[v][b] for the
// sources in the mask); edge lists longer than 256 are listed as segments
// that any wave takes. A pull gives a wave 16 consecutive vertices, each
// taking the minimum over its in-neighbours u of D[u][b] + w for the b in
// u's mask, and stores it (no atomics on D). No vertex lists are kept:
// counts of the joining vertices and their edges go to per-workgroup sums
// that the one-workgroup decision reads. Shortest-path distances are unique,
// so neither the order of relaxations nor the window schedule changes them.
// A step is five launches recorded in hipGraph blocks; the host reads the
// done flag once per block.
#include <climits>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kVbLanes = 64;         // sources per pass
constexpr int kVbBlock = 256;
constexpr int kVbWaves = kVbBlock / 64;
constexpr int kVbSeg = 256;          // edge lists longer than this are listed as segments of this many edges
constexpr int kVbPullWide = 1024;    // in-lists this long are pulled in chunks
constexpr int kVbPullChunk = 1024;   // in-edges per chunk item
constexpr int kVbBlockSteps = 16;
constexpr int kVbPullDiv = 8;        // a step pulls when its vertices hold more than 1 / kVbPullDiv of the edges
constexpr int kVbBatch = 4;          // rows in flight per wave
constexpr int kVbGroup = 16;         // vertices per wave in a pull

enum VbMode : int32_t { VB_LIGHT = 0, VB_HEAVY = 1 };

// What a relax or move launch's workgroup adds up for the next decision:
// vertices that joined F (or F') and their light out-edges, vertices that
// joined W and their heavy out-edges.
struct VbPart {
    long long fn, fe, wn, we;
};

struct alignas(128) VbCount {  // on a line of its own
    int32_t n;
};

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
    int32_t pad;
    long long fe;     // light out-edges of the F the coming relax step reads
    long long wn, we; // this window's W: vertices, heavy out-edges
    VbCount seg[5];   // segments listed for the F sets (0-2) and the W sets (3-4)
};

struct VbArgs {
    VbState *st;
    const int32_t *off, *nbr, *wt;     // out-lists (light part first when split)
    const int32_t *lend;               // end of each light part (null: every edge is light)
    const int32_t *poff, *pnbr, *pwt;  // in-lists
    const int32_t *plend;              // end of each light in-part (null: select by weight)
    int32_t V;
    int32_t scap;                      // segment list capacity
    long long pull_light, pull_heavy;  // vertices with more edges than these pull
    int32_t *D;  // [V][64]
    uint64_t *f[3];
    uint64_t *w[2];
    uint64_t *p;
    int4 *seg[5];   // (vertex, first edge, end edge) of the long lists of the F sets and the W sets
    VbPart *part;   // [grid]
    int32_t grid;
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

// The light or heavy part of v's out-list.
__device__ __forceinline__ void vb_part(const VbArgs &a, int32_t v, bool heavy, int32_t &e0, int32_t &e1) {
    const int32_t o0 = a.off[v], o1 = a.off[v + 1];
    const int32_t l = a.lend ? a.lend[v] : o1;
    e0 = heavy ? l : o0;
    e1 = heavy ? o1 : l;
}

// A lane's running counts of the vertices that joined F' / W and their edges.
struct VbAcc {
    long long fn = 0, fe = 0, wn = 0, we = 0;
};

// The part [e0, e1) of v's list joins segment list k (0-2: F sets, 3-4: W
// sets) when it is longer than kVbSeg.
__device__ void vb_list_segments(const VbArgs &a, int32_t k, int32_t v, int32_t e0, int32_t e1) {
    const int32_t n = (e1 - e0 + kVbSeg - 1) / kVbSeg;
    const int32_t p = atomicAdd(&a.st->seg[k].n, n);
    for (int32_t j = 0; j < n && p + j < a.scap; ++j) {
        const int32_t s0 = e0 + j * kVbSeg;
        a.seg[k][p + j] = make_int4(v, s0, s0 + kVbSeg < e1 ? s0 + kVbSeg : e1, 0);
    }
}

// Lane call: v joined F set k (heavy false) or W set k (heavy true).
__device__ __forceinline__ void vb_note(const VbArgs &a, int32_t k, bool heavy, int32_t v, VbAcc &acc) {
    int32_t e0, e1;
    vb_part(a, v, heavy, e0, e1);
    if (heavy) {
        acc.wn += 1;
        acc.we += e1 - e0;
    } else {
        acc.fn += 1;
        acc.fe += e1 - e0;
    }
    if (e1 - e0 > kVbSeg) vb_list_segments(a, heavy ? 3 + k : k, v, e0, e1);
}

// The same for the vertices held by the lanes in `mask` (lane j holds vj),
// their lists read in parallel. The whole wave calls.
__device__ __forceinline__ void vb_note_wave(const VbArgs &a, int32_t k, bool heavy, uint64_t mask, int32_t vj,
                                             VbAcc &acc) {
    if ((mask >> lane_id()) & 1ull) vb_note(a, k, heavy, vj, acc);
}

// Where the improvements of a relax step go.
struct VbOut {
    int32_t nxt;  // F' = f[nxt]
    int32_t cw;   // W = w[cw]
    int32_t hi;
};

// Improvements of v found by a writeMin (lane 0 calls): below hi they join
// F' and W, beyond it P. Bit 0 of the result: v joined F' here; bit 1: W.
__device__ int32_t vb_mark_atomic(const VbArgs &a, const VbOut &o, int32_t v, uint64_t near, uint64_t far) {
    int32_t r = 0;
    if (near) {
        if (!atomicOr((unsigned long long *)&a.f[o.nxt][v], near)) r |= 1;
        if (a.lend && (a.w[o.cw][v] & near) != near && !atomicOr((unsigned long long *)&a.w[o.cw][v], near)) r |= 2;
    }
    if (far && (a.p[v] & far) != far) atomicOr((unsigned long long *)&a.p[v], far);
    return r;
}

// Relaxes the edges held by lanes [0, cnt) (lane j: edge uj -> vj of weight
// wj, sources mj), kVbBatch at a time: lane b reads D[u][b] and D[v][b] for
// the sources b in the mask and writeMins D[v][b].
__device__ void vb_relax_lanes(const VbArgs &a, const VbOut &o, int32_t uj, int32_t vj, int32_t wj, uint64_t mj,
                               int32_t cnt, VbAcc &acc) {
    const int32_t lane = lane_id();
    int32_t fv = 0, wv = 0, nf = 0, nw = 0;  // vertices that joined F' / W (lane i holds the i-th): at most cnt
    for (int32_t j = 0; j < cnt; j += kVbBatch) {
        int32_t v[kVbBatch], nd[kVbBatch], dv[kVbBatch], old[kVbBatch];
        bool on[kVbBatch];
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k) {
            const int32_t jj = j + k < cnt ? j + k : cnt - 1;
            const int32_t u = __builtin_amdgcn_readlane(uj, jj);
            const uint64_t m = vb_readlane64(mj, jj);
            v[k] = j + k < cnt ? __builtin_amdgcn_readlane(vj, jj) : -1;
            on[k] = v[k] >= 0 && ((m >> lane) & 1ull);
            nd[k] = on[k] ? vb_row(a, u)[lane] : 0;
            dv[k] = on[k] ? vb_row(a, v[k])[lane] : INT_MIN;
        }
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k) {
            nd[k] += __builtin_amdgcn_readlane(wj, j + k < cnt ? j + k : cnt - 1);
            old[k] = INT_MIN;
            if (on[k] && nd[k] < dv[k]) old[k] = atomicMin(&vb_row(a, v[k])[lane], nd[k]);
        }
#pragma unroll
        for (int k = 0; k < kVbBatch; ++k) {
            const bool imp = nd[k] < old[k];
            const uint64_t near = __ballot(imp && nd[k] < o.hi), far = __ballot(imp && nd[k] >= o.hi);
            if (!(near | far)) continue;  // uniform
            int32_t r = 0;
            if (lane == 0) r = vb_mark_atomic(a, o, v[k], near, far);
            r = __builtin_amdgcn_readfirstlane(r);
            if (r & 1) {
                if (lane == nf) fv = v[k];
                ++nf;
            }
            if (r & 2) {
                if (lane == nw) wv = v[k];
                ++nw;
            }
        }
    }
    vb_note_wave(a, o.nxt, false, nf == 64 ? ~0ull : (1ull << nf) - 1, fv, acc);
    vb_note_wave(a, o.cw, true, nw == 64 ? ~0ull : (1ull << nw) - 1, wv, acc);
}

// Push: sweep the masks mk 64 vertices per wave; the lists of the vertices
// found (their light or heavy part, up to kVbSeg edges) are concatenated and
// relaxed 64 edges at a time. Longer lists are segment list k's items.
__device__ void vb_push(const VbArgs &a, const uint64_t *mk, bool heavy, int32_t k, const VbOut &o, VbAcc &acc) {
    const int32_t lane = lane_id();
    const int32_t groups = (a.V + 63) / 64;
    for (int32_t gi = vb_wave(); gi < groups; gi += vb_waves()) {
        const int32_t v = gi * 64 + lane;
        const uint64_t mv = v < a.V ? mk[v] : 0ull;
        if (!__ballot(mv != 0)) continue;  // uniform
        int32_t e0 = 0, e1 = 0;
        if (mv) vb_part(a, v, heavy, e0, e1);
        const int32_t len = e1 - e0 <= kVbSeg ? e1 - e0 : 0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;  // owner: the last lane whose exclusive prefix is <= g
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(e0, ow) + (g - __shfl(pre, ow));
            const bool has = g < total;
            const int32_t vj = has ? a.nbr[e] : 0;
            const int32_t wj = has ? a.wt[e] : 0;
            const uint64_t mj = __shfl(mv, ow);
            const int32_t cnt = total - base < 64 ? total - base : 64;
            vb_relax_lanes(a, o, gi * 64 + ow, vj, wj, mj, cnt, acc);
        }
    }
    const int32_t ns = a.st->seg[k].n < a.scap ? a.st->seg[k].n : a.scap;
    for (int32_t i = vb_wave(); i < ns; i += vb_waves()) {
        const int4 it = a.seg[k][i];
        const uint64_t m = mk[it.x];
        for (int32_t e = it.y; e < it.z; e += 64) {
            const int32_t cnt = it.z - e < 64 ? it.z - e : 64;
            const int32_t vj = lane < cnt ? a.nbr[e + lane] : 0;
            const int32_t wj = lane < cnt ? a.wt[e + lane] : 0;
            vb_relax_lanes(a, o, it.x, vj, wj, m, cnt, acc);
        }
    }
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

// Pull over every vertex: a wave takes kVbGroup consecutive vertices, reads
// their in-lists 64 edges at a time (lane j takes edge j of the
// concatenation), and for each edge whose source u has a nonzero mask reads
// D[u] once for all sources, folding D[u][b] + w into the wave's LDS table
// of minima (row: vertex of the group, column: source; a lane only touches
// its own column). Then the rows of the vertices that were reached are read
// kVbBatch at a time and the improvements stored; each vertex is owned by
// this wave, so its row, F' and W masks and pending bits take plain stores.
// In-lists of kVbPullWide edges or more are chunk items below.
__device__ void vb_pull(const VbArgs &a, const uint64_t *mk, bool heavy, const VbOut &o, int32_t (*cand)[64],
                        VbAcc &acc) {
    const int32_t lane = lane_id();
    const int32_t delta = a.st->delta;
    const int32_t sel = a.plend || !a.lend ? 0 : (heavy ? 2 : 1);
    const int32_t groups = (a.V + kVbGroup - 1) / kVbGroup;
    for (int32_t gi = vb_wave(); gi < groups; gi += vb_waves()) {
        const int32_t v0 = gi * kVbGroup;
        const int32_t v = v0 + lane;
        int32_t p0 = 0, p1 = 0;
        if (lane < kVbGroup && v < a.V && a.poff[v + 1] - a.poff[v] < kVbPullWide) vb_in_part(a, v, heavy, p0, p1);
        const int32_t len = p1 - p0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t r = 0; r < kVbGroup; ++r) cand[r][lane] = INT_MAX;
        uint64_t reached = 0;  // the group's vertices with an in-neighbour in the mask
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
                    own[k] = __builtin_amdgcn_readlane(ow, jj);
                    on[k] = j[k] >= 0 && ((m >> lane) & 1ull);
                    du[k] = on[k] ? vb_row(a, u)[lane] : 0;
                }
#pragma unroll
                for (int k = 0; k < kVbBatch; ++k) {
                    if (j[k] < 0) continue;  // uniform
                    reached |= 1ull << own[k];
                    if (on[k] && du[k] + w[k] < cand[own[k]][lane]) cand[own[k]][lane] = du[k] + w[k];
                }
            }
        }
        uint64_t jf = 0, jw = 0;  // the group's vertices that joined F' and W
        while (reached) {
            int32_t r[kVbBatch], dv[kVbBatch];
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                r[k] = reached ? __ffsll((unsigned long long)reached) - 1 : -1;
                reached &= reached - 1;
                dv[k] = r[k] >= 0 ? vb_row(a, v0 + r[k])[lane] : INT_MIN;
            }
#pragma unroll
            for (int k = 0; k < kVbBatch; ++k) {
                if (r[k] < 0) continue;  // uniform
                const int32_t c = cand[r[k]][lane];
                const bool imp = c < dv[k];
                const int32_t vv = v0 + r[k];
                if (imp) vb_row(a, vv)[lane] = c;
                const uint64_t near = __ballot(imp && c < o.hi), far = __ballot(imp && c >= o.hi);
                uint64_t w0 = 1;
                if (lane == 0) {
                    if (near) {
                        a.f[o.nxt][vv] = near;  // F' was clear
                        if (a.lend) {
                            w0 = a.w[o.cw][vv];
                            a.w[o.cw][vv] = w0 | near;
                        }
                    }
                    if (far) a.p[vv] |= far;
                }
                if (near) jf |= 1ull << r[k];
                if (near && !vb_readlane64(w0, 0)) jw |= 1ull << r[k];
            }
        }
        vb_note_wave(a, o.nxt, false, jf, v, acc);
        vb_note_wave(a, o.cw, true, jw, v, acc);
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
        if (lane == 0 && (near | far)) {
            const int32_t r = vb_mark_atomic(a, o, c.x, near, far);
            if (r & 1) vb_note(a, o.nxt, false, c.x, acc);
            if (r & 2) vb_note(a, o.cw, true, c.x, acc);
        }
    }
}

// The workgroup's sums of its lanes' counts, stored as its row of a.part.
// The whole workgroup calls.
__device__ void vb_store_part(const VbArgs &a, VbAcc acc) {
    __shared__ long long sum_s[kVbWaves][4];
    for (int o = 32; o > 0; o >>= 1) {
        acc.fn += __shfl_xor(acc.fn, o);
        acc.fe += __shfl_xor(acc.fe, o);
        acc.wn += __shfl_xor(acc.wn, o);
        acc.we += __shfl_xor(acc.we, o);
    }
    const int32_t wid = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sum_s[wid][0] = acc.fn;
        sum_s[wid][1] = acc.fe;
        sum_s[wid][2] = acc.wn;
        sum_s[wid][3] = acc.we;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        long long t = 0;
        for (int k = 0; k < kVbWaves; ++k) t += sum_s[k][threadIdx.x];
        (&a.part[blockIdx.x].fn)[threadIdx.x] = t;
    }
}

// The sums over the grid's rows of a.part (whole workgroup calls; thread 0
// gets them).
__device__ VbPart vb_sum_parts(const VbArgs &a) {
    __shared__ long long sum_s[kVbBlock / 64][4];
    VbPart t{0, 0, 0, 0};
    for (int32_t i = threadIdx.x; i < a.grid; i += blockDim.x) {
        const VbPart q = a.part[i];
        t.fn += q.fn;
        t.fe += q.fe;
        t.wn += q.wn;
        t.we += q.we;
    }
    for (int o = 32; o > 0; o >>= 1) {
        t.fn += __shfl_xor(t.fn, o);
        t.fe += __shfl_xor(t.fe, o);
        t.wn += __shfl_xor(t.wn, o);
        t.we += __shfl_xor(t.we, o);
    }
    if (lane_id() == 0) {
        sum_s[threadIdx.x >> 6][0] = t.fn;
        sum_s[threadIdx.x >> 6][1] = t.fe;
        sum_s[threadIdx.x >> 6][2] = t.wn;
        sum_s[threadIdx.x >> 6][3] = t.we;
    }
    __syncthreads();
    VbPart r{0, 0, 0, 0};
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) {
        r.fn += sum_s[k][0];
        r.fe += sum_s[k][1];
        r.wn += sum_s[k][2];
        r.we += sum_s[k][3];
    }
    return r;
}

// One relax step: clear the F set two steps back (and, at a window's first
// step, the last window's W set), then push or pull the light edges of F or
// the heavy edges of W. Every workgroup reads the same counts, so all take
// the same branch.
__global__ void __launch_bounds__(kVbBlock) vb_relax_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done) return;  // uniform
    const int32_t s = st->step, cur = s % 3, nxt = (s + 1) % 3, prv = (s + 2) % 3;
    const int32_t cw = st->wi & 1;
    const bool wclr = st->wclr != 0;
    for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < a.V; i += (int32_t)(gridDim.x * blockDim.x)) {
        a.f[prv][i] = 0;
        if (wclr) a.w[cw ^ 1][i] = 0;
    }
    const VbOut o{nxt, cw, st->hi};
    __shared__ int32_t cand[kVbWaves][kVbGroup][64];  // a pull's minima, one table per wave
    int32_t(*mine)[64] = cand[threadIdx.x >> 6];
    VbAcc acc;
    if (st->mode == VB_HEAVY) {
        if (st->we > a.pull_heavy) vb_pull(a, a.w[cw], true, o, mine, acc);
        else vb_push(a, a.w[cw], true, 3 + cw, o, acc);
    } else {
        if (st->fe > a.pull_light) vb_pull(a, a.f[cur], false, o, mine, acc);
        else vb_push(a, a.f[cur], false, cur, o, acc);
    }
    vb_store_part(a, acc);
}

// After the relax step: while F' gains members the window's light steps go
// on; then (with a split) one heavy step; then the window is settled.
__global__ void __launch_bounds__(kVbBlock) vb_decide_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done) return;  // uniform
    const VbPart t = vb_sum_parts(a);
    if (threadIdx.x != 0) return;
    const int32_t s = st->step, prv = (s + 2) % 3;
    st->steps += 1;
    st->pulls += st->mode == VB_HEAVY ? st->we > a.pull_heavy : st->fe > a.pull_light;
    st->wn += t.wn;
    st->we += t.we;
    st->wclr = 0;  // the relax step cleared the other W set
    st->adv = 0;
    if (t.fn > 0) {
        st->mode = VB_LIGHT;
        st->fe = t.fe;
    } else if (st->mode == VB_LIGHT && a.lend) {
        st->mode = VB_HEAVY;
    } else {
        st->mode = VB_LIGHT;
        st->adv = 1;
        st->pmin = INT_MAX;
        st->windows += 1;
        st->wi += 1;
        st->wclr = a.lend != nullptr;
        st->wn = 0;
        st->we = 0;
        st->seg[3 + (st->wi & 1)].n = 0;
    }
    st->seg[prv].n = 0;  // the set the next relax step builds
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
    const int32_t groups = (a.V + 63) / 64;
    VbAcc acc;
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
        vb_note_wave(a, cur, false, joined, v0 + lane, acc);
        if (a.lend) vb_note_wave(a, cw, true, joined, v0 + lane, acc);
    }
    vb_store_part(a, acc);
}

// After an advance: the counts of the F and W the move pass filled.
__global__ void __launch_bounds__(kVbBlock) vb_count_kernel(VbArgs a) {
    VbState *st = a.st;
    if (st->done || !st->adv) return;  // uniform
    const VbPart t = vb_sum_parts(a);
    if (threadIdx.x != 0) return;
    st->fe = t.fe;
    st->wn += t.wn;
    st->we += t.we;
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
        st->fe = 0;
        st->wn = 0;
        st->we = 0;
        for (int k = 0; k < 5; ++k) st->seg[k].n = 0;
    }
    __syncthreads();
    VbAcc acc;
    if (b < B) {
        const int32_t s = src[b];
        vb_row(a, s)[b] = 0;
        if (!atomicOr((unsigned long long *)&a.f[0][s], 1ull << b)) vb_note(a, 0, false, s, acc);
        if (a.lend && !atomicOr((unsigned long long *)&a.w[0][s], 1ull << b)) vb_note(a, 0, true, s, acc);
    }
    __syncthreads();
    if (b < B) {  // a few sources: atomics on the totals are fine here
        atomicAdd((unsigned long long *)&st->fe, (unsigned long long)acc.fe);
        atomicAdd((unsigned long long *)&st->wn, (unsigned long long)acc.wn);
        atomicAdd((unsigned long long *)&st->we, (unsigned long long)acc.we);
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
    L(vb_relax_kernel, a.grid, kVbBlock, a);
    L(vb_decide_kernel, 1, kVbBlock, a);
    L(vb_pmin_kernel, a.grid, kVbBlock, a);
    L(vb_move_kernel, a.grid, kVbBlock, a);
    L(vb_count_kernel, 1, kVbBlock, a);
}

struct VbRun {
    int32_t *D = nullptr;
    uint64_t *f[3] = {nullptr, nullptr, nullptr};
    uint64_t *w[2] = {nullptr, nullptr};
    uint64_t *p = nullptr;
    int4 *seg[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    VbPart *part = nullptr;
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
        for (auto x : seg) dfree(x);
        dfree(p);
        dfree(part);
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
    // a set lists at most E / kVbSeg + (lists longer than kVbSeg) <= 2 E / kVbSeg segments
    const int64_t scap64 = 2 * g->E / kVbSeg + 1;
    if (scap64 > INT32_MAX) return fail(GI_ERR_RANGE, "batched SSSP: graph too large");
    const int32_t grid = persistent_grid(kVbBlock, 8);
    GI_TRY(dalloc(&r.D, (size_t)V * kVbLanes + 1));
    for (int k = 0; k < 3; ++k) GI_TRY(dalloc(&r.f[k], (size_t)V + 1));
    for (int k = 0; k < 2; ++k) GI_TRY(dalloc(&r.w[k], (size_t)V + 1));
    for (int k = 0; k < 5; ++k) GI_TRY(dalloc(&r.seg[k], (size_t)scap64));
    GI_TRY(dalloc(&r.p, (size_t)V + 1));
    GI_TRY(dalloc(&r.part, (size_t)grid));
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
    a.scap = (int32_t)scap64;
    const bool whole = a.plend == nullptr;  // a pull reads every in-edge
    a.pull_light = (whole ? g->E : light) / kVbPullDiv;
    a.pull_heavy = (whole ? g->E : g->E - light) / kVbPullDiv;
    a.D = r.D;
    for (int k = 0; k < 3; ++k) a.f[k] = r.f[k];
    for (int k = 0; k < 2; ++k) a.w[k] = r.w[k];
    for (int k = 0; k < 5; ++k) a.seg[k] = r.seg[k];
    a.p = r.p;
    a.part = r.part;
    a.grid = grid;
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
