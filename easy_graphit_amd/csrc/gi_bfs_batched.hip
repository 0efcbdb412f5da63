// Batched BFS: the depth of every vertex from each of many sources
// (depth[source][vertex]), 64 sources per pass over the edges. The BFS app
// (apps/bfs.cpp:44-60) runs one source per call; here a group of up to 64
// sources advances together (multi-source BFS, Then et al., VLDB 2015):
// every vertex keeps a 64-bit mask of the group's sources that have reached
// it (seen) and a mask of those that reached it at the current level (its
// frontier mask), so one sweep of a vertex's edges serves every source whose
// bit is set.
//
// Level l (F = fr[l % 3], F' = fr[(l + 1) % 3]):
//   push  every u with a nonzero mask ORs mask(u) & ~seen(v) into F'(v) of
//         its out-neighbours (atomicOr, skipped when the bits are present);
//         out-lists of kMsChunk edges or more are chunk items that any wave
//         takes (a hub's list walked by one wave had made RMAT-20's second
//         level take 1.1 ms of a 64-source pass's 2.1);
//   pull  every v missing some of the group's sources ORs F(u) over its
//         in-neighbours u, restricted to the missing bits, and stops once
//         every missing bit is found (a plain store: a vertex has one owner);
//   finalize  F'(v) &= ~seen(v); seen(v) |= F'(v); depth[b][v] = l + 1 for
//         each bit b of F'(v); fr[(l + 2) % 3] is cleared for the next level;
//         the new frontier's size and out-degree go to per-workgroup slots;
//   decide  (one workgroup) sums the slots: done when F' is empty, else the
//         next level pulls when F' holds at least E / kMsPullDiv out-edges.
// A level is three launches recorded in hipGraph blocks (every kernel reads
// the level, the group's depth rows and its size from the device state, so
// one recording serves every group); the host reads the state once per
// replayed block.
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kMsBlock = 256;
constexpr int kMsDecBlock = 1024;
constexpr int32_t kMsWide = 64;       // degree from which a whole wave walks the list
constexpr int32_t kMsMaxGrid = 8192;  // partial-sum slots (one per workgroup)
// pull when F' has >= E / kMsPullDiv out-edges: a vertex's pull stops only
// once every source it misses is found, so an early level pulls nearly every
// in-edge (RMAT-20, 256 sources: E/16 8.1 ms, E/4 4.9, E/2 5.1, never 8.5)
constexpr long long kMsPullDiv = 4;
constexpr int32_t kMsBlockLevels = 8;
constexpr int32_t kMsChunk = 1024;    // out-lists this long are pushed as chunk items of this many edges

struct MsState {
    int32_t done;
    int32_t level;    // level of the frontier the next step expands
    int32_t pull;     // the next step pulls
    int32_t levels;   // levels of the group's deepest BFS + 1 (when done)
    int32_t nb;       // sources in the group (<= 64)
    int32_t pad;
    int32_t *depth;   // the group's depth rows, [nb][V]
    long long front;  // vertices with a nonzero frontier mask
    long long fdeg;   // their out-edges
};

struct MsArgs {
    MsState *st;
    const int32_t *off, *nbr;    // out-edges (push)
    const int32_t *poff, *pnbr;  // in-edges (pull; the graph itself when symmetric)
    int32_t V;
    int64_t E;
    uint64_t *seen;
    uint64_t *fr[3];
    long long *slots;  // [2][kMsMaxGrid]: front, fdeg per workgroup
    const int4 *wide;  // chunk items of the long out-lists (vertex, first edge, end edge)
    int32_t nwide;
};

__device__ __forceinline__ uint64_t *ms_pick3(uint64_t *a, uint64_t *b, uint64_t *c, int32_t i) {
    return i == 0 ? a : (i == 1 ? b : c);
}

__device__ __forceinline__ uint64_t ms_full(int32_t nb) { return nb >= 64 ? ~0ull : ((1ull << nb) - 1ull); }

__device__ __forceinline__ uint64_t ms_wave_or(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ void ms_push_edge(const MsArgs &a, uint64_t *nxt, uint64_t fu, int32_t v) {
    const uint64_t m = fu & ~a.seen[v];
    if (m && (nxt[v] & m) != m) atomicOr((unsigned long long *)&nxt[v], m);
}

__global__ void __launch_bounds__(kMsBlock) ms_step_kernel(MsArgs a) {
    MsState *st = a.st;
    if (st->done) return;  // uniform
    const int32_t l = st->level;
    const uint64_t *cur = ms_pick3(a.fr[0], a.fr[1], a.fr[2], l % 3);
    uint64_t *nxt = ms_pick3(a.fr[0], a.fr[1], a.fr[2], (l + 1) % 3);
    const uint64_t full = ms_full(st->nb);
    const int32_t lane = lane_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t nwords = ((int64_t)a.V + 63) >> 6;
    if (st->pull) {
        for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nwords; w += nwaves) {
            const int32_t v = (int32_t)(w * 64) + lane;
            const bool in = v < a.V;
            const uint64_t miss = in ? full & ~a.seen[v] : 0ull;
            int32_t e0 = 0, e1 = 0;
            if (miss) {
                e0 = a.poff[v];
                e1 = a.poff[v + 1];
            }
            const bool wide = e1 - e0 >= kMsWide;
            uint64_t acc = 0;
            if (!wide) {
                for (int32_t e = e0; e < e1; ++e) {
                    acc |= cur[a.pnbr[e]] & miss;
                    if (acc == miss) break;
                }
            }
            uint64_t wm = __ballot(wide);
            while (wm) {  // wide lists: the whole wave, 64 in-neighbours at a time
                const int32_t j = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                const int32_t we0 = __shfl(e0, j), we1 = __shfl(e1, j);
                const uint64_t wmiss = __shfl(miss, j);
                uint64_t wacc = 0;
                for (int32_t base = we0; base < we1; base += 64) {
                    const int32_t e = base + lane;
                    wacc |= ms_wave_or(e < we1 ? (cur[a.pnbr[e]] & wmiss) : 0ull);
                    if (wacc == wmiss) break;  // uniform: every lane holds the same wacc
                }
                if (lane == j) acc = wacc;
            }
            if (in) nxt[v] = acc;
        }
    } else {
        for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nwords; w += nwaves) {
            const int32_t u = (int32_t)(w * 64) + lane;
            const uint64_t fu = u < a.V ? cur[u] : 0ull;
            int32_t e0 = 0, e1 = 0;
            if (fu) {
                e0 = a.off[u];
                e1 = a.off[u + 1];
                if (e1 - e0 >= kMsChunk) e1 = e0;  // a chunk item below
            }
            const bool wide = e1 - e0 >= kMsWide;
            if (!wide)
                for (int32_t e = e0; e < e1; ++e) ms_push_edge(a, nxt, fu, a.nbr[e]);
            uint64_t wm = __ballot(wide);
            while (wm) {
                const int32_t j = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                const int32_t we0 = __shfl(e0, j), we1 = __shfl(e1, j);
                const uint64_t wfu = __shfl(fu, j);
                for (int32_t e = we0 + lane; e < we1; e += 64) ms_push_edge(a, nxt, wfu, a.nbr[e]);
            }
        }
        for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < a.nwide; i += nwaves) {
            const int4 c = a.wide[i];
            const uint64_t fu = cur[c.x];
            if (!fu) continue;  // uniform
            for (int32_t e = c.y + lane; e < c.z; e += 64) ms_push_edge(a, nxt, fu, a.nbr[e]);
        }
    }
}

// Chunk items of the out-lists of kMsChunk edges or more.
__global__ void ms_wide_kernel(const int32_t *off, int32_t V, int4 *out, int32_t cap, int32_t *n) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int32_t e0 = off[v], e1 = off[v + 1];
        if (e1 - e0 < kMsChunk) continue;
        const int32_t k = (e1 - e0 + kMsChunk - 1) / kMsChunk;
        const int32_t p = atomicAdd(n, k);
        for (int32_t j = 0; j < k && p + j < cap; ++j) {
            const int32_t s0 = e0 + j * kMsChunk;
            out[p + j] = make_int4(v, s0, s0 + kMsChunk < e1 ? s0 + kMsChunk : e1, 0);
        }
    }
}

__global__ void __launch_bounds__(kMsBlock) ms_finalize_kernel(MsArgs a) {
    MsState *st = a.st;
    if (st->done) return;
    const int32_t l = st->level;
    uint64_t *nxt = ms_pick3(a.fr[0], a.fr[1], a.fr[2], (l + 1) % 3);
    uint64_t *clr = ms_pick3(a.fr[0], a.fr[1], a.fr[2], (l + 2) % 3);
    int32_t *depth = st->depth;
    const int32_t lane = lane_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t nwords = ((int64_t)a.V + 63) >> 6;
    long long front = 0, fdeg = 0;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nwords; w += nwaves) {
        const int32_t v = (int32_t)(w * 64) + lane;
        const bool in = v < a.V;
        uint64_t n = 0;
        if (in) {
            const uint64_t sv = a.seen[v];
            n = nxt[v] & ~sv;
            if (n) a.seen[v] = sv | n;
            nxt[v] = n;
            clr[v] = 0;
        }
        if (n) {
            front += 1;
            fdeg += a.off[v + 1] - a.off[v];
        }
        // depths: one masked, coalesced store per source bit present in the wave
        uint64_t any = ms_wave_or(n);
        while (any) {
            const int32_t b = __ffsll((unsigned long long)any) - 1;
            any &= any - 1;
            if ((n >> b) & 1ull) depth[(int64_t)b * a.V + v] = l + 1;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        front += __shfl_xor(front, o);
        fdeg += __shfl_xor(fdeg, o);
    }
    __shared__ long long rf[kMsBlock / 64], rd[kMsBlock / 64];
    if (lane == 0) {
        rf[threadIdx.x >> 6] = front;
        rd[threadIdx.x >> 6] = fdeg;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long f = 0, d = 0;
        for (int i = 0; i < kMsBlock / 64; ++i) {
            f += rf[i];
            d += rd[i];
        }
        a.slots[blockIdx.x] = f;  // one slot per workgroup: no atomics on one word
        a.slots[kMsMaxGrid + blockIdx.x] = d;
    }
}

__global__ void __launch_bounds__(kMsDecBlock) ms_decide_kernel(MsArgs a, int32_t nslots) {
    MsState *st = a.st;
    if (st->done) return;  // uniform
    const int32_t l = st->level;
    long long f = 0, d = 0;
    for (int32_t i = threadIdx.x; i < nslots; i += kMsDecBlock) {
        f += a.slots[i];
        d += a.slots[kMsMaxGrid + i];
        a.slots[i] = 0;
        a.slots[kMsMaxGrid + i] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        f += __shfl_xor(f, o);
        d += __shfl_xor(d, o);
    }
    __shared__ long long rf[kMsDecBlock / 64], rd[kMsDecBlock / 64];
    if (lane_id() == 0) {
        rf[threadIdx.x >> 6] = f;
        rd[threadIdx.x >> 6] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        f = 0;
        d = 0;
        for (int i = 0; i < kMsDecBlock / 64; ++i) {
            f += rf[i];
            d += rd[i];
        }
        st->front = f;
        st->fdeg = d;
        if (f == 0) {
            st->done = 1;
            st->levels = l + 1;
        } else {
            st->level = l + 1;
            st->pull = d >= a.E / kMsPullDiv && d > 0;
        }
    }
}

struct MsSeed {
    int32_t *depth;   // the group's depth rows
    int32_t nb;
};

// Clears the masks and the group's depth rows, seeds the sources (one
// thread per source; repeated vertices get several bits) and the state.
__global__ void ms_reset_kernel(MsArgs a, MsSeed sd) {
    const int64_t n = (int64_t)sd.nb * a.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n || i < a.V; i += stride) {
        if (i < n) sd.depth[i] = -1;
        if (i < a.V) {
            a.seen[i] = 0;
            a.fr[0][i] = 0;
            a.fr[1][i] = 0;
            a.fr[2][i] = 0;
        }
    }
}

__global__ void ms_seed_kernel(MsArgs a, MsSeed sd, const int32_t *src) {
    const int32_t b = threadIdx.x;
    if (b < sd.nb) {
        const int32_t s = src[b];
        atomicOr((unsigned long long *)&a.seen[s], 1ull << b);
        atomicOr((unsigned long long *)&a.fr[0][s], 1ull << b);
        sd.depth[(int64_t)b * a.V + s] = 0;
    }
    if (b == 0) {
        MsState *st = a.st;
        st->done = 0;
        st->level = 0;
        st->pull = 0;
        st->levels = 0;
        st->nb = sd.nb;
        st->depth = sd.depth;
        st->front = sd.nb;
        st->fdeg = 0;
    }
}

__global__ void ms_zero_slots_kernel(long long *slots, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) slots[i] = 0;
}

static int ms_grid() {
    static const int per_cu = resident_per_cu((const void *)ms_step_kernel, kMsBlock);
    int g = device_info().num_cus * per_cu;
    return g < kMsMaxGrid ? g : kMsMaxGrid;
}

static void ms_launch_level(const MsArgs &a, Launcher &L) {
    const int grid = ms_grid();
    L(ms_step_kernel, grid, kMsBlock, a);
    L(ms_finalize_kernel, grid, kMsBlock, a);
    L(ms_decide_kernel, 1, kMsDecBlock, a, grid);
}

// Arrays, state and recorded level blocks of one call.
struct MsRun {
    uint64_t *seen = nullptr;
    uint64_t *fr[3] = {nullptr, nullptr, nullptr};
    long long *slots = nullptr;
    MsState *st = nullptr;
    MsState *host = nullptr;  // pinned
    int32_t *d_src = nullptr;
    int4 *wide = nullptr;
    int32_t *nwide = nullptr;
    hipGraphExec_t exec = nullptr;
    ~MsRun() {
        if (exec) (void)hipGraphExecDestroy(exec);
        dfree(wide);
        dfree(nwide);
        dfree(seen);
        for (auto p : fr) dfree(p);
        dfree(slots);
        dfree(st);
        dfree(d_src);
        if (host) (void)hipHostFree(host);
    }
};

static int bfs_batched(gi_graph *g, const int32_t *h_src, int32_t n, int32_t *d_depth, gi_run_stats *stats,
                       hipStream_t s) {
    const gi_graph *pg = g;  // in-edges for pull steps
    if (g->directed) {
        gi_graph *t = nullptr;
        GI_TRY(gi_graph_transpose(g, &t));
        pg = t;
    }
    MsRun r;
    const size_t V = (size_t)g->V;
    GI_TRY(dalloc(&r.seen, V));
    for (auto &p : r.fr) GI_TRY(dalloc(&p, V));
    GI_TRY(dalloc(&r.slots, 2 * (size_t)kMsMaxGrid));
    GI_TRY(dalloc(&r.st, 1));
    GI_TRY(dalloc(&r.d_src, 64));
    if (hipHostMalloc((void **)&r.host, sizeof(MsState)) != hipSuccess) return fail(GI_ERR_OOM, "pinned allocation failed");
    MsArgs a{};
    a.st = r.st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.poff = pg->d_off;
    a.pnbr = pg->d_nbr;
    a.V = g->V;
    a.E = g->E;
    a.seen = r.seen;
    for (int i = 0; i < 3; ++i) a.fr[i] = r.fr[i];
    a.slots = r.slots;
    ms_zero_slots_kernel<<<16, 256, 0, s>>>(r.slots, 2 * kMsMaxGrid);
    GI_LAUNCH_CHECK();
    // a list of d >= kMsChunk edges makes ceil(d / kMsChunk) <= 2 d / kMsChunk items
    const int32_t wcap = (int32_t)(2 * g->E / kMsChunk + 1);
    GI_TRY(dalloc(&r.wide, (size_t)wcap));
    GI_TRY(dalloc(&r.nwide, 1));
    GI_HIP(hipMemsetAsync(r.nwide, 0, sizeof(int32_t), s));
    if (V) ms_wide_kernel<<<(int)(ceil_div((int64_t)V, 256) < 4096 ? ceil_div((int64_t)V, 256) : 4096), 256, 0, s>>>(
        g->d_off, g->V, r.wide, wcap, r.nwide);
    GI_LAUNCH_CHECK();
    int32_t nwide = 0;
    GI_HIP(hipMemcpyAsync(&nwide, r.nwide, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    a.wide = r.wide;
    a.nwide = nwide < wcap ? nwide : wcap;
    GI_TRY(build_graph(
        [&](Launcher &L) {
            for (int32_t i = 0; i < kMsBlockLevels; ++i) ms_launch_level(a, L);
        },
        &r.exec));
    EventTimer timer(s);
    int64_t levels = 0;
    for (int32_t g0 = 0; g0 < n; g0 += 64) {
        const int32_t nb = n - g0 < 64 ? n - g0 : 64;
        MsSeed sd{d_depth + (int64_t)g0 * g->V, nb};
        GI_HIP(hipMemcpyAsync(r.d_src, h_src + g0, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
        const int64_t work = (int64_t)nb * g->V > (int64_t)g->V ? (int64_t)nb * g->V : (int64_t)g->V;
        int grid = (int)ceil_div(work, 256);
        const int cap = persistent_grid(256);
        ms_reset_kernel<<<grid < cap ? grid : cap, 256, 0, s>>>(a, sd);
        ms_seed_kernel<<<1, 64, 0, s>>>(a, sd, r.d_src);
        GI_LAUNCH_CHECK();
        // a BFS has at most V levels: the bound only guards a state that never ends
        for (int64_t used = 0;; used += kMsBlockLevels) {
            if (used > (int64_t)g->V + kMsBlockLevels) return fail(GI_ERR_HIP, "batched BFS never reported done");
            GI_HIP(hipGraphLaunch(r.exec, s));
            GI_HIP(hipMemcpyAsync(r.host, r.st, sizeof(MsState), hipMemcpyDeviceToHost, s));
            GI_HIP(hipStreamSynchronize(s));
            if (r.host->done) break;
        }
        levels += r.host->levels;
    }
    const double ms = timer.stop();
    if (stats) {
        stats->iterations = (int32_t)(levels < INT32_MAX ? levels : INT32_MAX);
        stats->push_iterations = 0;
        stats->pull_iterations = 0;
        stats->reached = 0;
        stats->edges_traversed = 0;
        stats->ms = ms;
    }
    return GI_OK;
}

}  // namespace gi

using namespace gi;

extern "C" int gi_bfs_batched(gi_graph *g, const int32_t *sources, int32_t num_sources, int32_t *d_depth,
                              gi_run_stats *stats, void *stream) {
    if (!g || (num_sources > 0 && (!sources || !d_depth))) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (num_sources < 0) return fail(GI_ERR_RANGE, "negative source count");
    for (int32_t i = 0; i < num_sources; ++i)
        if (sources[i] < 0 || sources[i] >= g->V)
            return fail(GI_ERR_RANGE, "source %d out of range [0,%d)", sources[i], g->V);
    if (num_sources == 0) {
        if (stats) *stats = gi_run_stats{};
        return GI_OK;
    }
    return bfs_batched(g, sources, num_sources, d_depth, stats, (hipStream_t)stream);
}
