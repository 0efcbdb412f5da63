// Batched Brandes betweenness, vertex-major: a pass serves up to 64 sources,
// one wave lane each. The BC app (apps/bc.cpp:64-125) runs one source at a
// time: a forward BFS that counts shortest paths (sigma, apps/bc.cpp:33-35),
// then a backward sweep from the deepest level that accumulates the
// dependencies (apps/bc.cpp:49-54), then the per-vertex score
// (apps/bc.cpp:58-63). Here the pass's values of vertex v sit in one row per
// array, so reading an edge (u, v) serves every source whose bit is set:
//
//   depth[v][64]  int32   level of v from source b (-1 unreached)
//   sigma[v][64]  f64     shortest-path counts (apps/bc.cpp:20 uses double)
//   dep[v][64]    f32     D(v) of the backward recurrence (apps/bc.cpp:21)
//   F[v]          u64     frontier masks, three sets in rotation
//   S[v]          u64     sources that reached v at earlier levels
//   snd, rcv      u64     backward: the masks of level l + 1 and of the
//                         receivers (level l on a symmetric graph, levels
//                         0..l on a directed one, apps/bc.cpp:108-114)
//
// Forward level l: a push sweeps F 64 vertices per wave and, for each edge
// (u, v) and source b of F[u] that has not reached v, adds sigma[u][b] into
// sigma[v][b] (atomic) and sets depth and F'; a pull (symmetric graphs, a
// frontier with more than E / 16 out-edges) gives a wave 8 vertices, each
// summing sigma[u][b] over its in-neighbours u in F for the sources it has
// not been reached by, in an LDS table, then storing it (no atomics). The
// backward pass mirrors it: a level launch adds 1/sigma over level l + 1 and
// builds snd and rcv; a step pushes dep[w][b] from the senders into the
// receivers (atomic) or pulls it into the receivers (LDS sums, no atomics).
// The arithmetic is the source-major engine's (gi_bc_batched.hip); the score
// of a vertex sums (dep - 1/sigma) * sigma over the pass's sources in double.
// Out-lists longer than 256 edges are listed as segments when their vertex
// joins F. Per-workgroup sums of the joins and their edges feed one-workgroup
// decision launches; the state lives on the device and the slots are replayed
// from recorded hipGraphs.
#include <climits>
#include <map>
#include <tuple>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kBvLanes = 64;
constexpr int kBvBlock = 256;
constexpr int kBvWaves = kBvBlock / 64;
constexpr int kBvGroup = 8;    // vertices per wave in a pull (16: 44.4 ms, 8: 42.6 on RMAT-20, 256 sources)
constexpr int kBvBatch = 4;    // rows in flight per wave
constexpr int kBvSeg = 256;    // out-lists longer than this are segments
constexpr int kBvSlots = 8;    // slots per recorded block after the first
constexpr int kBvWide = 1024;  // lists this long are pulled in chunks of this many edges

enum BvPhase : int32_t { BV_FWD = 0, BV_BWD = 1, BV_DONE = 2 };

struct BvPart {  // one workgroup's sums
    long long n, e, rn, re;  // forward: joins of F' and their out-edges; backward: senders, receivers
};

struct alignas(128) BvCount {
    int32_t n;
};

struct BvState {
    int32_t phase;
    int32_t level;   // forward: the level expanded; backward: l (senders at l + 1)
    int32_t levels;  // L, set when the forward pass ends
    int32_t pull;    // the coming step pulls
    long long fe;    // forward: out-edges of F[level]
    BvCount seg[3];  // segments listed per F set
};

struct BvArgs {
    BvState *st;
    const int32_t *off, *nbr;
    int32_t V;
    int32_t B;
    int32_t can_pull;  // symmetric graph: in-neighbours are the neighbours
    int32_t scap;
    long long pull_fwd;  // a forward frontier with more out-edges pulls
    int32_t *depth;
    double *sigma;
    float *dep;
    uint64_t *S;
    uint64_t *f[3];
    uint64_t *snd, *rcv;
    int4 *seg[3];
    BvPart *part;
    int32_t grid;
    const int4 *wide;  // chunks of the long lists (vertex, first edge, end edge)
    int32_t nwide;
};

__device__ __forceinline__ int32_t bv_wave() { return (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int32_t bv_waves() { return (int32_t)((gridDim.x * blockDim.x) >> 6); }
__device__ __forceinline__ int64_t bv_at(int32_t v) { return (int64_t)v * kBvLanes + lane_id(); }
__device__ __forceinline__ uint64_t bv_readlane64(uint64_t x, int32_t j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)x, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(x >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

struct BvAcc {
    long long n = 0, e = 0, rn = 0, re = 0;
};

// Lane call: v joined F set k; long out-lists are listed as segments.
__device__ void bv_note(const BvArgs &a, int32_t k, int32_t v, BvAcc &acc) {
    const int32_t e0 = a.off[v], e1 = a.off[v + 1];
    acc.n += 1;
    acc.e += e1 - e0;
    if (e1 - e0 > kBvSeg) {
        const int32_t n = (e1 - e0 + kBvSeg - 1) / kBvSeg;
        const int32_t p = atomicAdd(&a.st->seg[k].n, n);
        for (int32_t j = 0; j < n && p + j < a.scap; ++j) {
            const int32_t s0 = e0 + j * kBvSeg;
            a.seg[k][p + j] = make_int4(v, s0, s0 + kBvSeg < e1 ? s0 + kBvSeg : e1, 0);
        }
    }
}

// The workgroup's sums of its lanes' counts as its row of a.part.
__device__ void bv_store_part(const BvArgs &a, BvAcc acc) {
    __shared__ long long sum_s[kBvWaves][4];
    for (int o = 32; o > 0; o >>= 1) {
        acc.n += __shfl_xor(acc.n, o);
        acc.e += __shfl_xor(acc.e, o);
        acc.rn += __shfl_xor(acc.rn, o);
        acc.re += __shfl_xor(acc.re, o);
    }
    const int32_t wid = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sum_s[wid][0] = acc.n;
        sum_s[wid][1] = acc.e;
        sum_s[wid][2] = acc.rn;
        sum_s[wid][3] = acc.re;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        long long t = 0;
        for (int k = 0; k < kBvWaves; ++k) t += sum_s[k][threadIdx.x];
        (&a.part[blockIdx.x].n)[threadIdx.x] = t;
    }
}

// The sums over the grid's rows (whole workgroup calls).
__device__ BvPart bv_sum_parts(const BvArgs &a) {
    __shared__ long long sum_s[kBvWaves][4];
    BvPart t{0, 0, 0, 0};
    for (int32_t i = threadIdx.x; i < a.grid; i += blockDim.x) {
        const BvPart q = a.part[i];
        t.n += q.n;
        t.e += q.e;
        t.rn += q.rn;
        t.re += q.re;
    }
    for (int o = 32; o > 0; o >>= 1) {
        t.n += __shfl_xor(t.n, o);
        t.e += __shfl_xor(t.e, o);
        t.rn += __shfl_xor(t.rn, o);
        t.re += __shfl_xor(t.re, o);
    }
    if (lane_id() == 0) {
        sum_s[threadIdx.x >> 6][0] = t.n;
        sum_s[threadIdx.x >> 6][1] = t.e;
        sum_s[threadIdx.x >> 6][2] = t.rn;
        sum_s[threadIdx.x >> 6][3] = t.re;
    }
    __syncthreads();
    BvPart r{0, 0, 0, 0};
    for (int k = 0; k < kBvWaves; ++k) {
        r.n += sum_s[k][0];
        r.e += sum_s[k][1];
        r.rn += sum_s[k][2];
        r.re += sum_s[k][3];
    }
    return r;
}

// ------------------------------------------------------------------ forward

// Edges held by lanes [0, cnt) (lane j: uj -> vj, sources mj of uj at level
// l), kBvBatch at a time: the sources of the mask that have not reached v
// add sigma[u][b] into sigma[v][b] and set depth and F'. Vertices whose F'
// mask this sets first are noted.
__device__ void bv_fwd_lanes(const BvArgs &a, int32_t l, const uint64_t *fc, uint64_t *fo, int32_t nxt,
                             int32_t uj, int32_t vj, uint64_t mj, int32_t cnt, BvAcc &acc) {
    const int32_t lane = lane_id();
    int32_t jv = 0, nj = 0;  // vertices that joined F' (lane i holds the i-th)
    for (int32_t j = 0; j < cnt; j += kBvBatch) {
        int32_t v[kBvBatch];
        uint64_t nw[kBvBatch];
        double su[kBvBatch];
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) {
            const int32_t jj = j + k < cnt ? j + k : cnt - 1;
            v[k] = j + k < cnt ? __builtin_amdgcn_readlane(vj, jj) : -1;
            const uint64_t m = bv_readlane64(mj, jj);
            nw[k] = v[k] >= 0 ? m & ~(a.S[v[k]] | fc[v[k]]) : 0ull;  // sources that have not reached v
            const int32_t u = __builtin_amdgcn_readlane(uj, jj);
            su[k] = (nw[k] >> lane) & 1ull ? a.sigma[bv_at(u)] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) {
            if (!nw[k]) continue;  // uniform
            if ((nw[k] >> lane) & 1ull) {
                atomicAdd(&a.sigma[bv_at(v[k])], su[k]);
                a.depth[bv_at(v[k])] = l + 1;
            }
            unsigned long long was = 1;
            if (lane == 0 && (fo[v[k]] & nw[k]) != nw[k]) was = atomicOr((unsigned long long *)&fo[v[k]], nw[k]);
            if (!bv_readlane64(was, 0)) {  // uniform: v joined F' here
                if (lane == nj) jv = v[k];
                ++nj;
            }
        }
    }
    if (lane < nj) bv_note(a, nxt, jv, acc);
}

// Forward push: sweep F 64 vertices per wave, relax the concatenated out-lists
// of the vertices found (up to kBvSeg edges each), then the segments.
__device__ void bv_fwd_push(const BvArgs &a, int32_t l, int32_t cur, int32_t nxt, BvAcc &acc) {
    const int32_t lane = lane_id();
    const uint64_t *fc = a.f[cur];
    uint64_t *fo = a.f[nxt];
    const int32_t groups = (a.V + 63) / 64;
    for (int32_t gi = bv_wave(); gi < groups; gi += bv_waves()) {
        const int32_t v = gi * 64 + lane;
        const uint64_t mv = v < a.V ? fc[v] : 0ull;
        if (!__ballot(mv != 0)) continue;  // uniform
        const int32_t e0 = mv ? a.off[v] : 0, e1 = mv ? a.off[v + 1] : 0;
        const int32_t len = e1 - e0 <= kBvSeg ? e1 - e0 : 0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(e0, ow) + (g - __shfl(pre, ow));
            const int32_t vj = g < total ? a.nbr[e] : 0;
            const uint64_t mj = __shfl(mv, ow);
            const int32_t cnt = total - base < 64 ? total - base : 64;
            bv_fwd_lanes(a, l, fc, fo, nxt, gi * 64 + ow, vj, mj, cnt, acc);
        }
    }
    const int32_t ns = a.st->seg[cur].n < a.scap ? a.st->seg[cur].n : a.scap;
    for (int32_t i = bv_wave(); i < ns; i += bv_waves()) {
        const int4 it = a.seg[cur][i];
        const uint64_t m = fc[it.x];
        for (int32_t e = it.y; e < it.z; e += 64) {
            const int32_t cnt = it.z - e < 64 ? it.z - e : 64;
            const int32_t vj = lane < cnt ? a.nbr[e + lane] : 0;
            bv_fwd_lanes(a, l, fc, fo, nxt, it.x, vj, m, cnt, acc);
        }
    }
}

// Forward pull (symmetric graphs): a wave takes kBvGroup vertices; for the
// sources that have not reached v, the sum of sigma[u][b] over the
// neighbours u in F is gathered in the wave's LDS table (a lane owns its
// column), then stored by v's owner: sigma, depth, F'.
__device__ void bv_fwd_pull(const BvArgs &a, int32_t l, int32_t cur, int32_t nxt, double (*sum)[64], BvAcc &acc) {
    const int32_t lane = lane_id();
    const uint64_t *fc = a.f[cur];
    uint64_t *fo = a.f[nxt];
    const uint64_t all = a.B >= 64 ? ~0ull : (1ull << a.B) - 1;
    const int32_t groups = (a.V + kBvGroup - 1) / kBvGroup;
    for (int32_t gi = bv_wave(); gi < groups; gi += bv_waves()) {
        const int32_t v0 = gi * kBvGroup;
        const int32_t v = v0 + lane;
        const bool mine = lane < kBvGroup && v < a.V;
        const uint64_t open = mine ? all & ~(a.S[v] | fc[v]) : 0ull;  // sources that have not reached v
        int32_t e0 = open ? a.off[v] : 0, e1 = open ? a.off[v + 1] : 0;
        if (e1 - e0 >= kBvWide) e1 = e0;  // a chunk item below
        const int32_t len = e1 - e0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t r = 0; r < kBvGroup; ++r) sum[r][lane] = 0.0;
        uint64_t reached = 0;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(e0, ow) + (g - __shfl(pre, ow));
            const uint64_t op = __shfl(open, ow);
            const int32_t uj = g < total ? a.nbr[e] : 0;
            const uint64_t mj = g < total ? fc[uj] & op : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                int32_t j[kBvBatch], own[kBvBatch];
                double su[kBvBatch];
                bool on[kBvBatch];
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    const int32_t jj = j[k] >= 0 ? j[k] : 0;
                    const int32_t u = __builtin_amdgcn_readlane(uj, jj);
                    const uint64_t m = bv_readlane64(mj, jj);
                    own[k] = __builtin_amdgcn_readlane(ow, jj);
                    on[k] = j[k] >= 0 && ((m >> lane) & 1ull);
                    su[k] = on[k] ? a.sigma[bv_at(u)] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    if (j[k] < 0) continue;  // uniform
                    reached |= 1ull << own[k];
                    if (on[k]) sum[own[k]][lane] += su[k];
                }
            }
        }
        uint64_t joined = 0;
        while (reached) {
            const int32_t r = __ffsll((unsigned long long)reached) - 1;
            reached &= reached - 1;
            const double s = sum[r][lane];
            const bool found = s > 0.0;  // path counts are positive
            const int32_t vv = v0 + r;
            if (found) {
                a.sigma[bv_at(vv)] = s;
                a.depth[bv_at(vv)] = l + 1;
            }
            const uint64_t fm = __ballot(found);
            if (fm) {
                if (lane == 0) fo[vv] = fm;  // F' was clear; v is this wave's
                joined |= 1ull << r;
            }
        }
        if ((joined >> lane) & 1ull) bv_note(a, nxt, v, acc);
    }
    // chunks of the long lists: several waves share a vertex, so the sums
    // join sigma by atomicAdd and F' by atomicOr
    for (int32_t i = bv_wave(); i < a.nwide; i += bv_waves()) {
        const int4 c = a.wide[i];
        const uint64_t open = all & ~(a.S[c.x] | fc[c.x]);
        if (!open) continue;  // uniform
        double t = 0.0;
        for (int32_t e = c.y; e < c.z; e += 64) {
            const int32_t cnt = c.z - e < 64 ? c.z - e : 64;
            const int32_t uj = lane < cnt ? a.nbr[e + lane] : 0;
            const uint64_t mj = lane < cnt ? fc[uj] & open : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                double su[kBvBatch];
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    const int32_t j = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                    const int32_t jj = j >= 0 ? j : 0;
                    const uint64_t m = bv_readlane64(mj, jj);
                    su[k] = j >= 0 && ((m >> lane) & 1ull) ? a.sigma[bv_at(__builtin_amdgcn_readlane(uj, jj))] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) t += su[k];
            }
        }
        const bool found = t > 0.0;
        if (found) {
            atomicAdd(&a.sigma[bv_at(c.x)], t);
            a.depth[bv_at(c.x)] = l + 1;
        }
        const uint64_t fm = __ballot(found);
        if (lane == 0 && fm && !atomicOr((unsigned long long *)&fo[c.x], fm)) bv_note(a, nxt, c.x, acc);
    }
}

// Forward step of the state's level l: fold F[l] into S (the tests read S |
// F[l], so the fold may overlap them), clear F[l - 1] for reuse as F[l + 2],
// then push or pull.
__global__ void __launch_bounds__(kBvBlock) bv_fwd_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_FWD) return;  // uniform
    const int32_t l = st->level, cur = l % 3, nxt = (l + 1) % 3, prv = (l + 2) % 3;
    for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < a.V; i += (int32_t)(gridDim.x * blockDim.x)) {
        const uint64_t m = a.f[cur][i];
        if (m) a.S[i] |= m;
        a.f[prv][i] = 0;
    }
    __shared__ double sum[kBvWaves][kBvGroup][64];
    BvAcc acc;
    if (st->pull) bv_fwd_pull(a, l, cur, nxt, sum[threadIdx.x >> 6], acc);
    else bv_fwd_push(a, l, cur, nxt, acc);
    bv_store_part(a, acc);
}

// After a forward step: an empty F' ends the pass with L = l + 1 levels and
// starts the backward pass at l = L - 2; otherwise the next level and its
// direction.
__global__ void __launch_bounds__(kBvBlock) bv_fwd_decide_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_FWD) return;  // uniform
    const BvPart t = bv_sum_parts(a);
    if (threadIdx.x != 0) return;
    const int32_t l = st->level;
    st->seg[(l + 2) % 3].n = 0;  // the set the next step builds
    if (t.n == 0) {
        st->levels = l + 1;
        st->level = l - 1;
        st->phase = BV_BWD;
        st->pull = 0;
    } else {
        st->level = l + 1;
        st->fe = t.e;
        st->pull = a.can_pull && t.e > a.pull_fwd;
    }
}

// ----------------------------------------------------------------- backward

// Backward level l: dep += 1/sigma over level l + 1 (backward_vertex_f,
// apps/bc.cpp:49-52) and the masks of the senders (level l + 1) and the
// receivers, with their union counts and out-edges for the step's
// direction. kBvBatch rows at a time.
__global__ void __launch_bounds__(kBvBlock) bv_level_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_BWD) return;  // uniform
    const int32_t l = st->level, lane = lane_id();
    BvAcc acc;
    for (int32_t v0 = bv_wave() * kBvBatch; v0 < a.V; v0 += bv_waves() * kBvBatch) {
        int32_t d[kBvBatch];
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) d[k] = v0 + k < a.V && lane < a.B ? a.depth[bv_at(v0 + k)] : -1;
        const int32_t dg = lane < kBvBatch && v0 + lane < a.V ? a.off[v0 + lane + 1] - a.off[v0 + lane] : 0;
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) {
            const int32_t v = v0 + k;
            if (v >= a.V) break;  // uniform
            const bool s = d[k] == l + 1;
            if (s) {
                const int64_t i = bv_at(v);
                a.dep[i] = (float)((double)a.dep[i] + 1.0 / a.sigma[i]);
            }
            const uint64_t sm = __ballot(s);
            const uint64_t rm = __ballot(a.can_pull ? d[k] == l : (d[k] >= 0 && d[k] <= l));
            if (lane == 0) {
                a.snd[v] = sm;
                a.rcv[v] = rm;
            }
            const int32_t deg = __builtin_amdgcn_readlane(dg, k);
            if (lane == 0 && (sm || rm)) {
                if (sm) {
                    acc.n += 1;
                    acc.e += deg;
                }
                if (rm) {
                    acc.rn += 1;
                    acc.re += deg;
                }
            }
        }
    }
    bv_store_part(a, acc);
}

// The direction of the backward step: pull (symmetric graphs) unless the
// receivers' lists are far longer than the senders' (a push adds with one
// atomic per edge and source: RMAT-20, 64 sources, the two densest levels
// took 8.6 and 13.3 ms pushed).
__global__ void __launch_bounds__(kBvBlock) bv_bwd_decide1_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_BWD || st->level < 0) return;  // uniform
    const BvPart t = bv_sum_parts(a);
    if (threadIdx.x != 0) return;
    st->pull = a.can_pull && t.re <= 64 * t.e;
}

// Backward push: sweep the senders 64 vertices per wave; for each edge (w,
// v) the sources that have w as a sender and v as a receiver add dep[w][b]
// into dep[v][b] (atomic).
__device__ void bv_bwd_push(const BvArgs &a) {
    const int32_t lane = lane_id();
    const int32_t groups = (a.V + 63) / 64;
    for (int32_t gi = bv_wave(); gi < groups; gi += bv_waves()) {
        const int32_t w = gi * 64 + lane;
        const uint64_t mw = w < a.V ? a.snd[w] : 0ull;
        if (!__ballot(mw != 0)) continue;  // uniform
        const int32_t e0 = mw ? a.off[w] : 0, e1 = mw ? a.off[w + 1] : 0;
        const int32_t len = e1 - e0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(e0, ow) + (g - __shfl(pre, ow));
            const uint64_t mo = __shfl(mw, ow);  // every lane takes part
            const int32_t vj = g < total ? a.nbr[e] : 0;
            const uint64_t mj = g < total ? mo & a.rcv[vj] : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                int32_t v[kBvBatch];
                float dw[kBvBatch];
                bool on[kBvBatch];
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    const int32_t j = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                    const int32_t jj = j >= 0 ? j : 0;
                    const uint64_t m = bv_readlane64(mj, jj);
                    v[k] = __builtin_amdgcn_readlane(vj, jj);
                    on[k] = j >= 0 && ((m >> lane) & 1ull);
                    dw[k] = on[k] ? a.dep[bv_at(gi * 64 + __builtin_amdgcn_readlane(ow, jj))] : 0.f;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k)
                    if (on[k]) atomicAdd(&a.dep[bv_at(v[k])], dw[k]);
            }
        }
    }
}

// Backward pull (symmetric graphs): a wave takes kBvGroup vertices; each
// receiver sums dep[w][b] over its neighbours w that are senders of the same
// source in the wave's LDS table, then adds it to its own dep.
__device__ void bv_bwd_pull(const BvArgs &a, float (*sum)[64]) {
    const int32_t lane = lane_id();
    const int32_t groups = (a.V + kBvGroup - 1) / kBvGroup;
    for (int32_t gi = bv_wave(); gi < groups; gi += bv_waves()) {
        const int32_t v0 = gi * kBvGroup;
        const int32_t v = v0 + lane;
        const uint64_t rm = lane < kBvGroup && v < a.V ? a.rcv[v] : 0ull;
        int32_t e0 = rm ? a.off[v] : 0, e1 = rm ? a.off[v + 1] : 0;
        if (e1 - e0 >= kBvWide) e1 = e0;  // a chunk item below
        const int32_t len = e1 - e0;
        int32_t incl = len;
        for (int s = 1; s < 64; s <<= 1) {
            const int32_t y = __shfl_up(incl, s);
            if (lane >= s) incl += y;
        }
        const int32_t total = __builtin_amdgcn_readlane(incl, 63);
        const int32_t pre = incl - len;
        for (int32_t r = 0; r < kBvGroup; ++r) sum[r][lane] = 0.f;
        uint64_t reached = 0;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t g = base + lane;
            int32_t ow = 0;
            for (int32_t step = 32; step > 0; step >>= 1) {
                const int32_t t = ow + step;
                const int32_t pt = __shfl(pre, t & 63);  // every lane takes part
                if (t < 64 && pt <= g) ow = t;
            }
            const int32_t e = __shfl(e0, ow) + (g - __shfl(pre, ow));
            const uint64_t rv = __shfl(rm, ow);
            const int32_t wj = g < total ? a.nbr[e] : 0;
            const uint64_t mj = g < total ? a.snd[wj] & rv : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                int32_t j[kBvBatch], own[kBvBatch];
                float dw[kBvBatch];
                bool on[kBvBatch];
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    j[k] = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    const int32_t jj = j[k] >= 0 ? j[k] : 0;
                    const int32_t w = __builtin_amdgcn_readlane(wj, jj);
                    const uint64_t m = bv_readlane64(mj, jj);
                    own[k] = __builtin_amdgcn_readlane(ow, jj);
                    on[k] = j[k] >= 0 && ((m >> lane) & 1ull);
                    dw[k] = on[k] ? a.dep[bv_at(w)] : 0.f;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    if (j[k] < 0) continue;  // uniform
                    reached |= 1ull << own[k];
                    if (on[k]) sum[own[k]][lane] += dw[k];
                }
            }
        }
        while (reached) {
            const int32_t r = __ffsll((unsigned long long)reached) - 1;
            reached &= reached - 1;
            const float s = sum[r][lane];
            if (s != 0.f) a.dep[bv_at(v0 + r)] += s;  // v is this wave's
        }
    }
    // chunks of the long lists: their sums join dep by atomicAdd
    for (int32_t i = bv_wave(); i < a.nwide; i += bv_waves()) {
        const int4 c = a.wide[i];
        const uint64_t rv = a.rcv[c.x];
        if (!rv) continue;  // uniform
        float t = 0.f;
        for (int32_t e = c.y; e < c.z; e += 64) {
            const int32_t cnt = c.z - e < 64 ? c.z - e : 64;
            const int32_t wj = lane < cnt ? a.nbr[e + lane] : 0;
            const uint64_t mj = lane < cnt ? a.snd[wj] & rv : 0ull;
            uint64_t act = __ballot(mj != 0);
            while (act) {
                float dw[kBvBatch];
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) {
                    const int32_t j = act ? __ffsll((unsigned long long)act) - 1 : -1;
                    act &= act - 1;
                    const int32_t jj = j >= 0 ? j : 0;
                    const uint64_t m = bv_readlane64(mj, jj);
                    dw[k] = j >= 0 && ((m >> lane) & 1ull) ? a.dep[bv_at(__builtin_amdgcn_readlane(wj, jj))] : 0.f;
                }
#pragma unroll
                for (int k = 0; k < kBvBatch; ++k) t += dw[k];
            }
        }
        if (t != 0.f) atomicAdd(&a.dep[bv_at(c.x)], t);
    }
}

__global__ void __launch_bounds__(kBvBlock) bv_bwd_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_BWD || st->level < 0) return;  // uniform
    __shared__ float sum[kBvWaves][kBvGroup][64];
    if (st->pull) bv_bwd_pull(a, sum[threadIdx.x >> 6]);
    else bv_bwd_push(a);
}

__global__ void bv_bwd_decide2_kernel(BvArgs a) {
    BvState *st = a.st;
    if (st->phase != BV_BWD || threadIdx.x != 0) return;
    st->level -= 1;
    if (st->level < -1) st->phase = BV_DONE;
}

// scores[v] += sum over the pass's sources of (dep - 1/sigma) * sigma where
// sigma != 0 (apps/bc.cpp:58-63), in double; kBvBatch rows at a time.
__global__ void __launch_bounds__(kBvBlock) bv_final_kernel(BvArgs a, float *scores) {
    const int32_t lane = lane_id();
    for (int32_t v0 = bv_wave() * kBvBatch; v0 < a.V; v0 += bv_waves() * kBvBatch) {
        double t[kBvBatch];
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) {
            t[k] = 0.0;
            if (v0 + k < a.V && lane < a.B) {
                const double s = a.sigma[bv_at(v0 + k)];
                if (s != 0.0) t[k] = ((double)a.dep[bv_at(v0 + k)] - 1.0 / s) * s;
            }
        }
#pragma unroll
        for (int k = 0; k < kBvBatch; ++k) {
            for (int o = 32; o > 0; o >>= 1) t[k] += __shfl_xor(t[k], o);
            if (lane == 0 && v0 + k < a.V) scores[v0 + k] += (float)t[k];
        }
    }
}

__global__ void bv_reset_kernel(BvArgs a) {
    const int64_t n = (int64_t)a.V * kBvLanes;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a.depth[i] = -1;
        a.sigma[i] = 0.0;
        a.dep[i] = 0.f;
        if (i < a.V) {
            a.S[i] = 0;
            a.f[0][i] = 0;
            a.f[1][i] = 0;
            a.f[2][i] = 0;
        }
    }
}

// Level 0: source b in lane b (sources may repeat: each has its lane).
__global__ void bv_seed_kernel(BvArgs a, const int32_t *src) {
    BvState *st = a.st;
    const int32_t b = threadIdx.x;
    if (b == 0) {
        st->phase = BV_FWD;
        st->level = 0;
        st->levels = 0;
        st->pull = 0;
        st->fe = 0;
        for (int k = 0; k < 3; ++k) st->seg[k].n = 0;
    }
    __syncthreads();
    BvAcc acc;
    if (b < a.B) {
        const int32_t s = src[b];
        a.depth[(int64_t)s * kBvLanes + b] = 0;
        a.sigma[(int64_t)s * kBvLanes + b] = 1.0;
        if (!atomicOr((unsigned long long *)&a.f[0][s], 1ull << b)) bv_note(a, 0, s, acc);
    }
    long long e = acc.e;  // one wave: the sum over the sources' lists
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
    if (b == 0) {
        st->fe = e;
        st->pull = a.can_pull && e > a.pull_fwd;
    }
}

// Chunk items of the lists of kBvWide edges or more.
__global__ void bv_wide_kernel(const int32_t *off, int32_t V, int4 *out, int32_t cap, int32_t *n) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int32_t e0 = off[v], e1 = off[v + 1];
        if (e1 - e0 < kBvWide) continue;
        const int32_t k = (e1 - e0 + kBvWide - 1) / kBvWide;
        const int32_t p = atomicAdd(n, k);
        for (int32_t j = 0; j < k && p + j < cap; ++j) {
            const int32_t s0 = e0 + j * kBvWide;
            out[p + j] = make_int4(v, s0, s0 + kBvWide < e1 ? s0 + kBvWide : e1, 0);
        }
    }
}

static void bv_launch_slot(const BvArgs &a, int32_t phase, Launcher &L) {
    if (phase == BV_FWD) {
        L(bv_fwd_kernel, a.grid, kBvBlock, a);
        L(bv_fwd_decide_kernel, 1, kBvBlock, a);
    } else {
        L(bv_level_kernel, a.grid, kBvBlock, a);
        L(bv_bwd_decide1_kernel, 1, kBvBlock, a);
        L(bv_bwd_kernel, a.grid, kBvBlock, a);
        L(bv_bwd_decide2_kernel, 1, 64, a);
    }
}

struct BvRun {
    int32_t *depth = nullptr, *d_src = nullptr;
    double *sigma = nullptr;
    float *dep = nullptr;
    uint64_t *S = nullptr, *snd = nullptr, *rcv = nullptr;
    uint64_t *f[3] = {nullptr, nullptr, nullptr};
    int4 *seg[3] = {nullptr, nullptr, nullptr};
    int4 *wide = nullptr;
    int32_t *nwide = nullptr;
    BvPart *part = nullptr;
    BvState *st = nullptr;
    BvState *host = nullptr;  // pinned
    std::map<std::pair<int32_t, int32_t>, hipGraphExec_t> execs;  // (phase, slots)
    ~BvRun() {
        for (auto &kv : execs) (void)hipGraphExecDestroy(kv.second);
        dfree(depth);
        dfree(d_src);
        dfree(sigma);
        dfree(dep);
        dfree(S);
        dfree(snd);
        dfree(rcv);
        for (auto p : f) dfree(p);
        for (auto p : seg) dfree(p);
        dfree(wide);
        dfree(nwide);
        dfree(part);
        dfree(st);
        if (host) (void)hipHostFree(host);
    }
};

// Replays recorded slots of one pass until the state leaves it: `first`
// slots, then kBvSlots at a time.
static int bv_pass(BvRun &r, const BvArgs &a, int32_t phase, int32_t first, hipStream_t s) {
    int32_t block = first < 1 ? 1 : (first > 512 ? 512 : first);
    for (int64_t used = 0; used < (1ll << 24);) {
        const auto key = std::make_pair(phase, block);
        auto it = r.execs.find(key);
        if (it == r.execs.end()) {
            hipGraphExec_t exec = nullptr;
            GI_TRY(build_graph(
                [&](Launcher &L) {
                    for (int32_t i = 0; i < block; ++i) bv_launch_slot(a, phase, L);
                },
                &exec));
            it = r.execs.emplace(key, exec).first;
        }
        GI_HIP(hipGraphLaunch(it->second, s));
        GI_HIP(hipMemcpyAsync(r.host, r.st, sizeof(BvState), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        used += block;
        if (r.host->phase != phase) return GI_OK;
        block = kBvSlots;
    }
    return fail(GI_ERR_RANGE, "BC did not finish");
}

int bc_vertex_major(gi_graph *g, const int32_t *h_src, int32_t num_sources, int32_t batch, float *d_scores,
                    hipStream_t s, int64_t *levels) {
    if (levels) *levels = 0;
    if (num_sources <= 0) return GI_OK;
    if (batch > kBvLanes) batch = kBvLanes;
    if (batch > num_sources) batch = num_sources;
    const int64_t V = g->V;
    const int64_t scap64 = 2 * g->E / kBvSeg + 1;
    if (scap64 > INT32_MAX) return fail(GI_ERR_RANGE, "BC: graph too large");
    BvRun r;
    const int32_t grid = persistent_grid(kBvBlock, 8);
    GI_TRY(dalloc(&r.depth, (size_t)V * kBvLanes + 1));
    GI_TRY(dalloc(&r.sigma, (size_t)V * kBvLanes + 1));
    GI_TRY(dalloc(&r.dep, (size_t)V * kBvLanes + 1));
    GI_TRY(dalloc(&r.S, (size_t)V + 1));
    GI_TRY(dalloc(&r.snd, (size_t)V + 1));
    GI_TRY(dalloc(&r.rcv, (size_t)V + 1));
    for (auto &p : r.f) GI_TRY(dalloc(&p, (size_t)V + 1));
    for (auto &p : r.seg) GI_TRY(dalloc(&p, (size_t)scap64));
    GI_TRY(dalloc(&r.part, (size_t)grid));
    const int32_t wcap = (int32_t)(g->E / kBvWide + 1);  // a chunk holds kBvWide edges but the last of a list
    GI_TRY(dalloc(&r.wide, (size_t)wcap * 2));
    GI_TRY(dalloc(&r.nwide, 1));
    GI_HIP(hipMemsetAsync(r.nwide, 0, sizeof(int32_t), s));
    if (V) bv_wide_kernel<<<(int)(ceil_div(V, 256) < 4096 ? ceil_div(V, 256) : 4096), 256, 0, s>>>(
        g->d_off, g->V, r.wide, 2 * wcap, r.nwide);
    GI_LAUNCH_CHECK();
    int32_t nwide = 0;
    GI_HIP(hipMemcpyAsync(&nwide, r.nwide, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    GI_TRY(dalloc(&r.d_src, (size_t)kBvLanes));
    GI_TRY(dalloc(&r.st, 1));
    if (hipHostMalloc((void **)&r.host, sizeof(BvState)) != hipSuccess) return fail(GI_ERR_OOM, "pinned allocation failed");
    BvArgs a{};
    a.st = r.st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.V = g->V;
    a.can_pull = !g->directed;
    a.scap = (int32_t)scap64;
    a.pull_fwd = g->E / 16;
    a.depth = r.depth;
    a.sigma = r.sigma;
    a.dep = r.dep;
    a.S = r.S;
    for (int k = 0; k < 3; ++k) {
        a.f[k] = r.f[k];
        a.seg[k] = r.seg[k];
    }
    a.snd = r.snd;
    a.rcv = r.rcv;
    a.part = r.part;
    a.grid = grid;
    a.wide = r.wide;
    a.nwide = nwide < 2 * wcap ? nwide : 2 * wcap;
    const int rgrid = persistent_grid(256);
    int32_t hint = 8;  // levels of the last pass
    for (int32_t i = 0; i < num_sources; i += batch) {
        a.B = num_sources - i < batch ? num_sources - i : batch;
        GI_HIP(hipMemcpyAsync(r.d_src, h_src + i, sizeof(int32_t) * a.B, hipMemcpyHostToDevice, s));
        bv_reset_kernel<<<rgrid, 256, 0, s>>>(a);
        bv_seed_kernel<<<1, kBvLanes, 0, s>>>(a, r.d_src);
        GI_LAUNCH_CHECK();
        // the recorded slots carry a.B: a pass of another size records its own
        if (i > 0 && a.B != batch) {
            for (auto &kv : r.execs) (void)hipGraphExecDestroy(kv.second);
            r.execs.clear();
        }
        GI_TRY(bv_pass(r, a, BV_FWD, hint, s));
        const int32_t L = r.host->levels;
        GI_TRY(bv_pass(r, a, BV_BWD, L, s));  // l = L-2 .. -1: L slots
        bv_final_kernel<<<rgrid, kBvBlock, 0, s>>>(a, d_scores);
        GI_LAUNCH_CHECK();
        hint = L;
        if (levels) *levels += L;
    }
    GI_HIP(hipStreamSynchronize(s));
    return GI_OK;
}

}  // namespace gi
