// Fused PageRank: the iteration loop of apps/pr.cpp:28-46 (contrib =
// rank / out-degree; new = sum of in-neighbour contribs; rank = (1-d)/N +
// d * new; no dangling redistribution) as a pull over the in-edge CSR with
// no atomics, recorded once as a hipGraph of all iterations.
//
// Work per iteration (two launches):
//   * pr_gather: work items are groups of 64 consecutive vertices and
//     512-edge chunks of the in-edge lists of wide vertices (in-degree >= 64).
//     A group's wave sums the in-edge lists of its narrow vertices
//     cooperatively (the lists are contiguous: 64 edges per pass, owner by
//     binary search over the lane prefix, a segmented scan across the wave,
//     accumulation per vertex in LDS) and writes their new rank and next
//     contribution. A chunk's wave writes one partial sum.
//   * pr_finalize: thread per wide vertex adds its chunk partials in chunk
//     order and writes its rank and next contribution.
// Every sum has a fixed order, so results are deterministic.
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "gi_internal.h"

namespace gi {

constexpr int kPrBlock = 256;
constexpr int32_t kPrNarrow = 64;      // in-degree below this: summed inside its group
constexpr int32_t kPrChunk = 512;      // edges per chunk of a wide vertex
constexpr int kPrPasses = 4;           // 64-edge passes of a group in flight together
// The gather keeps the first kPrHubs contributions in LDS (128 KB, one
// 1024-thread workgroup per CU): on a degree-ordered view these are the
// hubs, the sources of about half the in-edges of an RMAT graph, whose
// gathers then leave the L2 alone (random 4-byte gathers run at the L2's
// request rate).
constexpr int kPrGatherBlock = 1024;
constexpr int32_t kPrHubs = 32768;

struct PrArgs {
    const int32_t *poff;      // in-edge CSR
    const int32_t *pnbr;
    const int32_t *off;       // out-edge CSR offsets (out-degree)
    int32_t V;
    int32_t groups;           // ceil(V / 64)
    int32_t chunks;
    const int32_t *chunk_b;   // wide-vertex index of each chunk
    const int32_t *chunk_e0;  // first edge of each chunk
    const int32_t *wide;      // wide vertices
    const int32_t *wide_c0;   // first chunk of each wide vertex (nwide + 1)
    int32_t nwide;
    float *partial;           // one per chunk
    const float *contrib;     // this iteration's contributions
    float *contrib_next;
    float *rank;              // written when write_rank
    float damping;
    float beta;               // (1 - d) / N
    int32_t write_rank;
};

__device__ __forceinline__ void pr_store(const PrArgs &a, int32_t v, float sum) {
    float r = a.beta + a.damping * sum;
    if (a.write_rank) a.rank[v] = r;
    a.contrib_next[v] = r / (float)(a.off[v + 1] - a.off[v]);  // apps/pr.cpp:28-30
}

// contrib[u], from the workgroup's copy of the first kPrHubs entries when u
// is one of them
__device__ __forceinline__ float pr_contrib(const PrArgs &a, const float *hub, int32_t u) {
    return u < kPrHubs ? hub[u] : a.contrib[u];
}

__global__ void __launch_bounds__(kPrGatherBlock) pr_gather_kernel(PrArgs a) {
    __shared__ float acc_s[kPrGatherBlock / 64][64];
    __shared__ float4 hub4[kPrHubs / 4];
    float *hub = reinterpret_cast<float *>(hub4);
    {  // the first kPrHubs contributions (the hubs, on a degree-ordered view)
        const int32_t n4 = (a.V < kPrHubs ? a.V : kPrHubs) / 4;
        const float4 *c4 = reinterpret_cast<const float4 *>(a.contrib);
        for (int32_t i = threadIdx.x; i < n4; i += kPrGatherBlock) hub4[i] = c4[i];
        for (int32_t i = 4 * n4 + (int32_t)threadIdx.x; i < a.V && i < kPrHubs; i += kPrGatherBlock)
            hub[i] = a.contrib[i];
        __syncthreads();
    }
    const int32_t lane = lane_id(), wid = threadIdx.x >> 6;
    float *acc = acc_s[wid];
    const int32_t waves = gridDim.x * (kPrGatherBlock / 64);
    const int32_t items = a.groups + a.chunks;
    for (int32_t item = blockIdx.x * (kPrGatherBlock / 64) + wid; item < items; item += waves) {
        if (item < a.groups) {
            const int32_t v = item * 64 + lane;
            int32_t e0 = 0, len = 0;
            bool narrow = false;
            if (v < a.V) {
                e0 = a.poff[v];
                len = a.poff[v + 1] - e0;
                narrow = len < kPrNarrow;
                if (!narrow) len = 0;
            }
            acc[lane] = 0.f;
            int32_t incl = len;
            for (int o = 1; o < 64; o <<= 1) {
                int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            const int32_t total = __shfl(incl, 63);
            const int32_t pre = incl - len;
            __builtin_amdgcn_wave_barrier();
            // kPrPasses passes of 64 edges at a time: their index loads, then
            // their gathers, are in flight together; the sums are added in
            // pass order, as one pass at a time would
            for (int32_t base = 0; base < total; base += 64 * kPrPasses) {
                int32_t jj[kPrPasses], ee[kPrPasses], uu[kPrPasses];
                float xx[kPrPasses];
#pragma unroll
                for (int p = 0; p < kPrPasses; ++p) {
                    const int32_t g = base + 64 * p + lane;
                    int32_t j = 0;  // owner: last lane whose exclusive prefix is <= g
                    for (int32_t step = 32; step > 0; step >>= 1) {
                        int32_t t = j + step;
                        int32_t pt = __shfl(pre, t & 63);
                        if (t < 64 && pt <= g) j = t;
                    }
                    const int32_t ej = __shfl(e0, j), pj = __shfl(pre, j);  // every lane takes part
                    const bool valid = g < total;
                    jj[p] = valid ? j : 64;  // past the last owner
                    ee[p] = valid ? ej + (g - pj) : -1;
                }
#pragma unroll
                for (int p = 0; p < kPrPasses; ++p) uu[p] = ee[p] >= 0 ? a.pnbr[ee[p]] : -1;
#pragma unroll
                for (int p = 0; p < kPrPasses; ++p) xx[p] = uu[p] >= 0 ? pr_contrib(a, hub, uu[p]) : 0.f;
#pragma unroll
                for (int p = 0; p < kPrPasses; ++p) {
                    float x = xx[p];
                    const int32_t j = jj[p];
                    // segmented inclusive scan (segments = runs of equal owner)
                    for (int o = 1; o < 64; o <<= 1) {
                        float y = __shfl_up(x, o);
                        int32_t jo = __shfl_up(j, o);
                        if (lane >= o && jo == j) x += y;
                    }
                    const int32_t jn = __shfl_down(j, 1);
                    if (j < 64 && (lane == 63 || jn != j)) acc[j] += x;  // segment end
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if (narrow) pr_store(a, v, acc[lane]);
            __builtin_amdgcn_wave_barrier();
        } else {
            const int32_t k = item - a.groups;
            const int32_t b = a.chunk_b[k];
            const int32_t vw = a.wide[b];
            const int32_t e0 = a.chunk_e0[k];
            const int32_t end = a.poff[vw + 1];
            const int32_t e1 = e0 + kPrChunk < end ? e0 + kPrChunk : end;
            // the chunk's kPrChunk / 64 edges per lane: index loads, then
            // gathers, in flight together, added in edge order
            int32_t uu[kPrChunk / 64];
            float xx[kPrChunk / 64];
#pragma unroll
            for (int i = 0; i < kPrChunk / 64; ++i) {
                const int32_t e = e0 + lane + 64 * i;
                uu[i] = e < e1 ? a.pnbr[e] : -1;
            }
#pragma unroll
            for (int i = 0; i < kPrChunk / 64; ++i) xx[i] = uu[i] >= 0 ? pr_contrib(a, hub, uu[i]) : 0.f;
            float x = 0.f;
#pragma unroll
            for (int i = 0; i < kPrChunk / 64; ++i) x += xx[i];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
            if (lane == 0) a.partial[k] = x;
        }
    }
}

__global__ void pr_finalize_kernel(PrArgs a) {
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < a.nwide; b += gridDim.x * blockDim.x) {
        float sum = 0.f;
        for (int32_t k = a.wide_c0[b]; k < a.wide_c0[b + 1]; ++k) sum += a.partial[k];
        pr_store(a, a.wide[b], sum);
    }
}

__global__ void pr_init_kernel(const int32_t *off, int32_t V, float *contrib, float *rank) {
    const float r0 = (float)(1.0 / (double)V);  // apps/pr.cpp:43-46
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        rank[v] = r0;
        contrib[v] = r0 / (float)(off[v + 1] - off[v]);
    }
}

// ---------------------------------------------------------------- host side
struct PrPlan {
    const int32_t *poff = nullptr, *pnbr = nullptr;
    int32_t nwide = 0, chunks = 0;
    int32_t *wide = nullptr, *wide_c0 = nullptr, *chunk_b = nullptr, *chunk_e0 = nullptr;
    float *partial = nullptr;
    float *contrib[2] = {nullptr, nullptr};
    // recorded runs keyed by (iterations, output array, damping bits)
    std::map<std::tuple<int32_t, const float *, uint32_t>, hipGraphExec_t> execs;
};

static PlanRegistry<PrPlan> g_pr;

static void free_plan(PrPlan *p) {
    for (auto &kv : p->execs) (void)hipGraphExecDestroy(kv.second);
    dfree(p->wide);
    dfree(p->wide_c0);
    dfree(p->chunk_b);
    dfree(p->chunk_e0);
    dfree(p->partial);
    dfree(p->contrib[0]);
    dfree(p->contrib[1]);
    delete p;
}

void release_pr(const gi_graph *g) {
    if (PrPlan *p = g_pr.take(g)) free_plan(p);
}

static int ensure_plan(gi_graph *g, PrPlan **out) {
    if ((*out = g_pr.find(g)) != nullptr) return GI_OK;
    gi_graph *pg = g;
    if (g->directed) GI_TRY(gi_graph_transpose(g, &pg));
    const int32_t V = g->V;
    std::vector<int32_t> poff((size_t)V + 1);
    GI_HIP(hipMemcpy(poff.data(), pg->d_off, sizeof(int32_t) * (V + 1), hipMemcpyDeviceToHost));
    std::vector<int32_t> wide, wide_c0(1, 0), chunk_b, chunk_e0;
    for (int32_t v = 0; v < V; ++v) {
        int32_t d = poff[v + 1] - poff[v];
        if (d < kPrNarrow) continue;
        int32_t b = (int32_t)wide.size();
        wide.push_back(v);
        for (int32_t e = poff[v]; e < poff[v + 1]; e += kPrChunk) {
            chunk_b.push_back(b);
            chunk_e0.push_back(e);
        }
        wide_c0.push_back((int32_t)chunk_b.size());
    }
    PrPlan *p = new PrPlan();
    p->poff = pg->d_off;
    p->pnbr = pg->d_nbr;
    p->nwide = (int32_t)wide.size();
    p->chunks = (int32_t)chunk_b.size();
    int rc = dalloc(&p->wide, wide.size() + 1);
    if (rc == GI_OK) rc = dalloc(&p->wide_c0, wide_c0.size());
    if (rc == GI_OK) rc = dalloc(&p->chunk_b, chunk_b.size() + 1);
    if (rc == GI_OK) rc = dalloc(&p->chunk_e0, chunk_e0.size() + 1);
    if (rc == GI_OK) rc = dalloc(&p->partial, chunk_b.size() + 1);
    if (rc == GI_OK) rc = dalloc(&p->contrib[0], (size_t)V + 1);
    if (rc == GI_OK) rc = dalloc(&p->contrib[1], (size_t)V + 1);
    auto up = [&](int32_t *d, const std::vector<int32_t> &h) {
        if (rc == GI_OK && !h.empty() &&
            hipMemcpy(d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(GI_ERR_HIP, "plan upload");
    };
    up(p->wide, wide);
    up(p->wide_c0, wide_c0);
    up(p->chunk_b, chunk_b);
    up(p->chunk_e0, chunk_e0);
    if (rc != GI_OK) {
        free_plan(p);
        return rc;
    }
    g_pr.put(g, p);
    *out = p;
    return GI_OK;
}

static int gather_grid() {
    static const int per_cu = resident_per_cu((const void *)pr_gather_kernel, kPrGatherBlock);
    return device_info().num_cus * per_cu;
}

static void launch_iters(gi_graph *g, PrPlan *p, int32_t iters, float damping, float *rank,
                         Launcher &L) {
    PrArgs a;
    a.poff = p->poff;
    a.pnbr = p->pnbr;
    a.off = g->d_off;
    a.V = g->V;
    a.groups = (int32_t)ceil_div(g->V, 64);
    a.chunks = p->chunks;
    a.chunk_b = p->chunk_b;
    a.chunk_e0 = p->chunk_e0;
    a.wide = p->wide;
    a.wide_c0 = p->wide_c0;
    a.nwide = p->nwide;
    a.partial = p->partial;
    a.rank = rank;
    a.damping = damping;
    a.beta = (float)((1.0 - (double)damping) / (double)g->V);  // apps/pr.cpp:36-41
    int ig = (int)ceil_div(g->V, 256);
    L(pr_init_kernel, ig < 4096 ? ig : 4096, 256, g->d_off, g->V, p->contrib[0], rank);
    const int grid = gather_grid();
    const int fg = (int)ceil_div(p->nwide > 0 ? p->nwide : 1, 256);
    for (int32_t it = 0; it < iters; ++it) {
        a.contrib = p->contrib[it & 1];
        a.contrib_next = p->contrib[(it + 1) & 1];
        a.write_rank = it + 1 == iters;
        L(pr_gather_kernel, grid, kPrGatherBlock, a);
        if (p->nwide > 0) L(pr_finalize_kernel, fg < 4096 ? fg : 4096, 256, a);
    }
}

int pr_fused_run(gi_graph *g, int32_t iters, float damping, float *d_rank, hipStream_t s) {
    if (g->V == 0) return GI_OK;
    PrPlan *p = nullptr;
    GI_TRY(ensure_plan(g, &p));
    if (iters == 0) {
        int ig = (int)ceil_div(g->V, 256);
        pr_init_kernel<<<ig < 4096 ? ig : 4096, 256, 0, s>>>(g->d_off, g->V, p->contrib[0], d_rank);
        GI_LAUNCH_CHECK();
        return GI_OK;
    }
    // the whole run is one recorded graph per (iteration count, output, damping)
    uint32_t dbits = 0;
    memcpy(&dbits, &damping, sizeof(dbits));
    auto key = std::make_tuple(iters, (const float *)d_rank, dbits);
    auto it = p->execs.find(key);
    if (it == p->execs.end()) {
        if (p->execs.size() >= 8) {  // bounded cache
            for (auto &kv : p->execs) (void)hipGraphExecDestroy(kv.second);
            p->execs.clear();
        }
        hipGraphExec_t exec = nullptr;
        GI_TRY(build_graph([&](Launcher &L) { launch_iters(g, p, iters, damping, d_rank, L); }, &exec));
        it = p->execs.emplace(key, exec).first;
    }
    GI_HIP(hipGraphLaunch(it->second, s));
    return GI_OK;
}

}  // namespace gi
