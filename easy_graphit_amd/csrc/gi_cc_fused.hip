// Fused connected components: the loop of apps/cc.cpp:53-78 (hook every
// edge, ID[ID[s]] min= ID[ID[d]] both ways, then pointer jumping until
// stable, until a sweep changes nothing) with every decision on the device
// and the rounds recorded as a hipGraph.
//
// Two schedules. EDGE_ONLY (the app's own, apps/cc.cpp:84-86) sweeps every
// edge each round. VERTEX_BASED first hooks the first two neighbours of every
// vertex (two vertex passes, each followed by shortcutting), samples the most
// frequent label, and then runs rounds in which only vertices outside that
// label hook their edges; on graphs with a giant component (RMAT) those
// rounds touch a small part of the graph. Both stop when a round lowers no
// label and give the same labels.
//
// Per EDGE_ONLY round (three launches):
//   * cc_hook: an edge-parallel sweep of the CSR. Edges are cut into chunks
//     of kCcChunk; the source of each chunk's first edge is precomputed once
//     per graph. A wave takes 256 edges at a time (4 per lane, loads in
//     flight together) and finds each edge's source from a window of the
//     next 64 offsets held one per lane (binary search by shuffles), so no
//     COO source array is read. Only the larger of the two parent labels is
//     lowered (atomicMin), and only when it is larger. The hook of (u, v)
//     is the hook of (v, u), and the graph is symmetric, so only the edges
//     with v > u hook: half the label gathers and atomics of a sweep. A
//     sweep that lowers nothing still proves every edge's two ends equal.
//     The half kept is the one a low id's own list holds, so in the
//     degree-ordered layout the hubs, whose lists come first, hook their
//     neighbours in the sweep's first chunks (RMAT-24: 2 rounds, 8.8 ->
//     6.5 ms; keeping v < u instead took 3 rounds and 15.8 ms).
//   * cc_compress: every vertex follows its label chain to the root and
//     stores it (full shortcutting; the fixed point is the same as the
//     reference's repeated single jumps).
//   * cc_control: ends the run when a sweep lowered nothing.
// The result is the component's minimum vertex id, as in the reference:
// labels are vertex ids of the same component and only decrease, and the
// minimum vertex never takes another label.
#include <functional>
#include <map>
#include <tuple>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kCcBlock = 256;
constexpr int32_t kCcChunk = 8192;   // edges per chunk (a wave's unit of work)
constexpr int kCcEdgesPerLane = 4;

struct CcState {
    int32_t changed;   // set by a hook sweep that lowered a label
    int32_t done;
    int32_t rounds;
    int32_t giant;     // label whose vertices the filtered rounds skip (-1: none)
};

struct CcArgs {
    const int32_t *off;
    const int32_t *nbr;
    const int32_t *chunk_v;  // source vertex of each chunk's first edge
    int32_t *id;
    CcState *st;
    int32_t V;
    int64_t E;
    int32_t chunks;
};

__global__ void cc_chunk_owner_kernel(const int32_t *off, int32_t V, int32_t chunks, int32_t *chunk_v) {
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < chunks; c += gridDim.x * blockDim.x) {
        const int64_t e = (int64_t)c * kCcChunk;
        // last v with off[v] <= e
        int32_t lo = 0, hi = V;  // answer in [lo, hi)
        while (hi - lo > 1) {
            int32_t mid = lo + (hi - lo) / 2;
            if ((int64_t)off[mid] <= e) lo = mid;
            else hi = mid;
        }
        chunk_v[c] = lo;
    }
}

__global__ void cc_init_kernel(int32_t *id, int32_t V, CcState *st) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) id[v] = v;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->changed = 0;
        st->done = 0;
        st->rounds = 0;
        st->giant = -1;
    }
}

// Source of edge e: the last v >= v0 with off[v] <= e. `wo` holds
// off[v0 + 1 + lane] (INT_MAX past V); returns -1 when the source lies past
// the window.
__device__ __forceinline__ int32_t window_owner(int32_t v0, int32_t wo, int32_t e) {
    const bool past = __shfl(wo, 63) <= e;  // all 64 window offsets <= e
    int32_t c = 0;  // number of window offsets <= e, when below 64
    for (int32_t step = 32; step > 0; step >>= 1) {
        int32_t t = c + step;
        if (__shfl(wo, t - 1) <= e) c = t;
    }
    return past ? -1 : v0 + c;
}

__device__ __forceinline__ int32_t global_owner(const int32_t *off, int32_t lo, int32_t V, int32_t e) {
    int32_t hi = V;
    while (hi - lo > 1) {
        int32_t mid = lo + (hi - lo) / 2;
        if (off[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kCcBlock) cc_hook_kernel(CcArgs a) {
    if (a.st->done) return;
    const int32_t lane = lane_id();
    const int32_t waves = gridDim.x * (kCcBlock / 64);
    bool changed = false;
    for (int32_t c = blockIdx.x * (kCcBlock / 64) + (threadIdx.x >> 6); c < a.chunks; c += waves) {
        const int32_t c0 = c * kCcChunk;
        const int32_t c1 = (int64_t)c0 + kCcChunk < a.E ? c0 + kCcChunk : (int32_t)a.E;
        int32_t vcur = a.chunk_v[c];
        for (int32_t b = c0; b < c1; b += 64 * kCcEdgesPerLane) {
            const int32_t wv = vcur + 1 + lane;
            const int32_t wo = wv <= a.V ? a.off[wv] : 0x7fffffff;
            int32_t u[kCcEdgesPerLane], d[kCcEdgesPerLane];
#pragma unroll
            for (int k = 0; k < kCcEdgesPerLane; ++k) {
                const int32_t e = b + 64 * k + lane;
                const bool valid = e < c1;
                d[k] = valid ? a.nbr[e] : -1;
                int32_t o = window_owner(vcur, wo, e);
                if (valid && o < 0) o = global_owner(a.off, vcur + 64, a.V, e);  // past the window
                u[k] = valid && d[k] > o ? o : -1;  // one direction of each symmetric pair
            }
            // next window starts at the source of the batch's last edge
            int32_t last = b + 64 * kCcEdgesPerLane - 1;
            last = last < c1 ? last : c1 - 1;
            const int32_t ol = window_owner(vcur, wo, last);
            vcur = ol >= 0 ? ol : global_owner(a.off, vcur + 64, a.V, last);
            int32_t sid[kCcEdgesPerLane], did[kCcEdgesPerLane];
#pragma unroll
            for (int k = 0; k < kCcEdgesPerLane; ++k) {
                sid[k] = u[k] >= 0 ? a.id[u[k]] : 0;
                did[k] = u[k] >= 0 ? a.id[d[k]] : 0;
            }
#pragma unroll
            for (int k = 0; k < kCcEdgesPerLane; ++k) {
                if (u[k] < 0) continue;
                const int32_t ps = a.id[sid[k]], pd = a.id[did[k]];  // apps/cc.cpp:23-29
                if (pd < ps) {
                    if (atomicMin(&a.id[sid[k]], pd) > pd) changed = true;
                } else if (ps < pd) {
                    if (atomicMin(&a.id[did[k]], ps) > ps) changed = true;
                }
            }
        }
    }
    if (__ballot(changed) && lane == 0 && !a.st->changed) a.st->changed = 1;
}

// One hook of edge (u, v) as the reference's UDF does it: ID[ID[u]] and
// ID[ID[v]] compared, the larger parent lowered by atomicMin
// (apps/cc.cpp:23-29). lu is ID[u], already read. True when it lowered.
__device__ __forceinline__ bool cc_hook(int32_t *id, int32_t lu, int32_t v) {
    const int32_t lv = id[v];
    const int32_t ps = id[lu], pd = id[lv];
    if (pd < ps) return atomicMin(&id[lu], pd) > pd;
    if (ps < pd) return atomicMin(&id[lv], ps) > ps;
    return false;
}

// Vertex-parallel hooks (VERTEX_BASED). Sampling pass r: every vertex hooks
// its r-th neighbour, so two passes join most of a giant component at the
// cost of two reads per vertex.
__global__ void __launch_bounds__(kCcBlock) cc_sample_kernel(CcArgs a, int32_t r) {
    if (a.st->done) return;
    for (int32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < a.V; u += gridDim.x * blockDim.x) {
        const int32_t e = a.off[u] + r;
        if (e < a.off[u + 1]) (void)cc_hook(a.id, a.id[u], a.nbr[e]);
    }
}

// The most frequent label among 256 spread-out vertices (one workgroup of
// kCcGiantBlock threads, one sample each, counted against all the others
// from LDS; the label only steers which vertices the filtered rounds skip,
// any label gives the same result). A giant component holding half the
// vertices gets about 128 of the samples; 1024 samples cost 22 us a run on
// RMAT-24, almost all of it the 1024 x 1024 comparisons.
constexpr int kCcGiantBlock = 256;
__global__ void __launch_bounds__(kCcGiantBlock) cc_giant_kernel(CcArgs a) {
    constexpr int kSamples = kCcGiantBlock;
    __shared__ int32_t lab[kSamples];
    __shared__ int32_t best_n[kCcGiantBlock / 64], best_l[kCcGiantBlock / 64];
    const int32_t i = threadIdx.x;
    const int32_t mine = a.id[(int32_t)(((uint64_t)i * 2654435761ull + 12345ull) % (uint64_t)a.V)];
    lab[i] = mine;
    __syncthreads();
    int32_t n = 0;
#pragma unroll 16
    for (int32_t j = 0; j < kSamples; ++j) n += lab[j] == mine;  // (the same j in every lane: a broadcast read)
    int32_t bn = n, bl = mine;
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t on = __shfl_xor(bn, o), ol = __shfl_xor(bl, o);
        if (on > bn || (on == bn && ol < bl)) {
            bn = on;
            bl = ol;
        }
    }
    if (lane_id() == 0) {
        best_n[threadIdx.x >> 6] = bn;
        best_l[threadIdx.x >> 6] = bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kCcGiantBlock / 64; ++w)
            if (best_n[w] > bn || (best_n[w] == bn && best_l[w] < bl)) {
                bn = best_n[w];
                bl = best_l[w];
            }
        a.st->giant = bl;
    }
}

// Filtered round: every vertex outside the giant label hooks all its edges;
// vertices inside it are skipped. In a symmetric graph an edge whose ends
// still differ has an end outside the giant label, whose own list holds the
// edge, so a round that lowers nothing still proves every edge's ends equal.
// Short lists are walked by their lane, long ones by the whole wave.
__global__ void __launch_bounds__(kCcBlock) cc_rest_kernel(CcArgs a) {
    if (a.st->done) return;
    const int32_t giant = a.st->giant, lane = lane_id();
    bool changed = false;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < a.V; base += stride) {
        const int32_t u = base + lane;
        int32_t lu = 0, e0 = 0, e1 = 0;
        if (u < a.V) {
            lu = a.id[u];
            if (lu != giant) {
                e0 = a.off[u];
                e1 = a.off[u + 1];
            }
        }
        const bool wide = e1 - e0 >= 64;
        if (!wide)
            for (int32_t e = e0; e < e1; ++e) changed |= cc_hook(a.id, lu, a.nbr[e]);
        uint64_t wm = __ballot(wide);
        while (wm) {
            const int32_t l = __ffsll((unsigned long long)wm) - 1;
            wm &= wm - 1;
            const int32_t w0 = __shfl(e0, l), w1 = __shfl(e1, l), wu = __shfl(lu, l);
            for (int32_t e = w0 + lane; e < w1; e += 64) changed |= cc_hook(a.id, wu, a.nbr[e]);
        }
    }
    if (__ballot(changed) && lane == 0 && !a.st->changed) a.st->changed = 1;
}

// Four consecutive vertices per thread, their chains followed side by side
// (a step's four loads in flight together).
__global__ void cc_compress_kernel(CcArgs a) {
    if (a.st->done) return;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t v0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x); v0 < a.V; v0 += 4 * stride) {
        int32_t y0[4], y[4], x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) y0[k] = v0 + k < a.V ? a.id[v0 + k] : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            y[k] = y0[k];
            x[k] = a.id[y[k]];
        }
        // labels strictly decrease along a chain, so each ends at a root
        bool more = x[0] != y[0] || x[1] != y[1] || x[2] != y[2] || x[3] != y[3];
        while (more) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (x[k] != y[k]) {
                    y[k] = x[k];
                    x[k] = a.id[y[k]];
                }
            more = x[0] != y[0] || x[1] != y[1] || x[2] != y[2] || x[3] != y[3];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (v0 + k < a.V && y[k] != y0[k]) a.id[v0 + k] = y[k];
    }
}

__global__ void cc_control_kernel(CcState *st) {
    if (st->done) return;
    st->rounds += 1;
    if (!st->changed) st->done = 1;
    st->changed = 0;
}

// ---------------------------------------------------------------- host side
struct CcPlan {
    int32_t chunks = 0;
    int32_t *chunk_v = nullptr;
    CcState *st = nullptr;
    int32_t *host = nullptr;  // pinned copy of the state
    // (rounds, labels, sampled, with the run's first launches)
    std::map<std::tuple<int32_t, int32_t *, bool, bool>, hipGraphExec_t> execs;
    int32_t last_rounds = 0;
};

static PlanRegistry<CcPlan> g_cc;

static void free_cc(CcPlan *p) {
    for (auto &kv : p->execs) (void)hipGraphExecDestroy(kv.second);
    dfree(p->chunk_v);
    dfree(p->st);
    if (p->host) (void)hipHostFree(p->host);
    delete p;
}

void release_cc(const gi_graph *g) {
    if (CcPlan *p = g_cc.take(g)) free_cc(p);
}

const int32_t *cc_giant_label(gi_graph *g) {
    CcPlan *p = g_cc.find(g);
    return p ? &p->st->giant : nullptr;
}

static int ensure_cc(gi_graph *g, CcPlan **out) {
    if ((*out = g_cc.find(g)) != nullptr) return GI_OK;
    if (g->E >= 0x7fffffffLL - kCcChunk) return fail(GI_ERR_RANGE, "fused CC needs E < 2^31");
    CcPlan *p = new CcPlan();
    p->chunks = (int32_t)ceil_div(g->E, kCcChunk);
    int rc = dalloc(&p->chunk_v, (size_t)p->chunks + 1);
    if (rc == GI_OK) rc = dalloc(&p->st, 1);
    if (rc == GI_OK && hipHostMalloc((void **)&p->host, sizeof(CcState)) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
    if (rc == GI_OK && p->chunks > 0) {
        int grid = (int)ceil_div(p->chunks, 256);
        cc_chunk_owner_kernel<<<grid, 256>>>(g->d_off, g->V, p->chunks, p->chunk_v);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(GI_ERR_HIP, "chunk owners");
    }
    if (rc != GI_OK) {
        free_cc(p);
        return rc;
    }
    g_cc.put(g, p);
    *out = p;
    return GI_OK;
}

static int hook_grid() {
    static const int per_cu = resident_per_cu((const void *)cc_hook_kernel, kCcBlock);
    return device_info().num_cus * per_cu;
}

static int vertex_grid(int32_t V) {
    const int g = (int)ceil_div(V, kCcBlock);
    return g < 8192 ? (g < 1 ? 1 : g) : 8192;
}

static void launch_round(const CcArgs &a, bool sampled, Launcher &L) {
    if (sampled) L(cc_rest_kernel, vertex_grid(a.V), kCcBlock, a);
    else L(cc_hook_kernel, hook_grid(), kCcBlock, a);
    int cg = (int)ceil_div(a.V, 256);
    L(cc_compress_kernel, cg < 8192 ? cg : 8192, 256, a);
    L(cc_control_kernel, 1, 1, a.st);
}

// The run is one recorded graph: the label init, for the sampled schedule
// its two sampling passes and the giant label, then a block of rounds sized
// by the last run (rounds after the end return at once); the state copy and
// the epilogue (the caller's, e.g. the labels' mapping back to original ids)
// are queued behind it, and the host waits once. A run longer than the block
// replays further blocks of rounds, each followed by the epilogue again.
int cc_fused_run(gi_graph *g, int32_t *d_labels, int32_t *rounds_out, bool sampled, hipStream_t s,
                 const std::function<int(hipStream_t)> *epilogue) {
    if (g->V == 0) {
        if (rounds_out) *rounds_out = 0;
        return epilogue ? (*epilogue)(s) : GI_OK;
    }
    if (g->directed) return fail(GI_ERR_UNSUPPORTED, "fused CC needs a symmetric graph");
    CcPlan *p = nullptr;
    GI_TRY(ensure_cc(g, &p));
    CcArgs a;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.chunk_v = p->chunk_v;
    a.id = d_labels;
    a.st = p->st;
    a.V = g->V;
    a.E = g->E;
    a.chunks = p->chunks;
    const int ig = (int)ceil_div(g->V, 256);
    int32_t block = p->last_rounds > 0 ? p->last_rounds : 4;
    bool first = true;
    for (int32_t done_rounds = 0; done_rounds < (1 << 20);) {
        if (block > 32) block = 32;
        auto key = std::make_tuple(block, d_labels, sampled, first);
        auto it = p->execs.find(key);
        if (it == p->execs.end()) {
            if (p->execs.size() >= 8) {
                for (auto &kv : p->execs) (void)hipGraphExecDestroy(kv.second);
                p->execs.clear();
            }
            hipGraphExec_t exec = nullptr;
            GI_TRY(build_graph(
                [&](Launcher &L) {
                    if (first) {
                        L(cc_init_kernel, ig < 8192 ? ig : 8192, 256, d_labels, g->V, p->st);
                        if (sampled) {  // two sampling passes, each followed by full shortcutting
                            for (int32_t r = 0; r < 2; ++r) {
                                L(cc_sample_kernel, vertex_grid(g->V), kCcBlock, a, r);
                                L(cc_compress_kernel, ig < 8192 ? ig : 8192, 256, a);
                            }
                            L(cc_giant_kernel, 1, kCcGiantBlock, a);
                        }
                    }
                    for (int32_t r = 0; r < block; ++r) launch_round(a, sampled, L);
                },
                &exec));
            it = p->execs.emplace(key, exec).first;
        }
        GI_HIP(hipGraphLaunch(it->second, s));
        GI_HIP(hipMemcpyAsync(p->host, p->st, sizeof(CcState), hipMemcpyDeviceToHost, s));
        if (epilogue) GI_TRY((*epilogue)(s));
        GI_HIP(hipStreamSynchronize(s));
        first = false;
        done_rounds += block;
        if (p->host[1]) {  // done
            p->last_rounds = p->host[2];
            if (rounds_out) *rounds_out = p->host[2];
            return GI_OK;
        }
        block *= 2;
    }
    return fail(GI_ERR_RANGE, "CC did not converge");
}

}  // namespace gi
