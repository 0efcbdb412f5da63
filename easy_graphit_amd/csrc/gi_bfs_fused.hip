// Fused BFS engine: the level loop of apps/bfs.cpp with every schedule
// decision on the device.
//
// The reference's fused mode (fuse_kernel, operators.cpp:651-672) runs the
// whole while-loop inside one cooperative kernel with grid barriers; its
// hybrid schedule decides push vs pull on the host from a D2H copy of the
// frontier size (operators.cpp:526). Here each level is three launches on one
// stream: a persistent step kernel that switches on the chosen mode, a hub
// kernel that spreads the edges of very wide push vertices over the whole
// grid (a no-op on most levels), and a one-wave decide kernel (GraphIt's
// hybrid rule or the MLP selector, from device counters). Every step writes
// the next frontier in both representations (sparse queue and 64-bit-word
// bitmap), so a direction change needs no conversion pass, and keeps the
// counts (|F|, sum of frontier degrees, unvisited vertices and their
// degrees) that feed the decision.
#include <vector>

#include "gi_bfs_fused.h"
#include "gi_lb.h"

namespace gi {

__device__ __forceinline__ int32_t deg_of(const int32_t *off, int32_t v) {
    return off[v + 1] - off[v];
}

struct StepArgs {
    BfsState *st;
    const int32_t *off;
    const int32_t *nbr;
    int32_t V;
    int32_t words;
    int32_t *depth;
    int32_t *parent;
    uint64_t *visited;
    uint64_t *bm[3];
    int32_t *q[2];
};

struct StepCtx {
    int32_t level;
    int32_t *qout;
    int32_t *count_out;
    uint64_t *bm_out;
    unsigned long long *fdeg_out;
};

// Push discovery: claim the visited bit, record depth (and parent), append
// to the sparse queue, set the bitmap bit, accumulate |F'| and the degree
// sum. Every active lane of the wave calls it once per step.
__device__ __forceinline__ void push_edge(const StepArgs &a, const StepCtx &c, int32_t s,
                                          int32_t e, bool active) {
    int32_t d = 0;
    bool fresh = false;
    if (active) {
        d = a.nbr[e];
        if (!bit_test(a.visited, d) && bit_claim(a.visited, d)) {
            fresh = true;
            a.depth[d] = c.level + 1;
            if (a.parent) a.parent[d] = s;
            atomicOr((unsigned long long *)&c.bm_out[d >> 6], 1ull << (d & 63));
        }
    }
    uint64_t m = __ballot(fresh);
    if (m == 0) return;
    int32_t pos = wave_atomic_add(c.count_out, fresh);
    if (fresh) c.qout[pos] = d;
    unsigned long long dg = fresh ? (unsigned long long)deg_of(a.off, d) : 0ull;
    for (int o = 32; o > 0; o >>= 1) dg += __shfl_xor(dg, o);
    if (lane_id() == 0) atomicAdd(c.fdeg_out, dg);
}

// Lane-per-vertex edge walk; the wave loops while any lane has edges left.
__device__ __forceinline__ void lane_walk(const StepArgs &a, const StepCtx &c, int32_t s,
                                          int32_t e, int32_t e1) {
    while (__ballot(e < e1)) {
        bool live = e < e1;
        push_edge(a, c, s, e, live);
        e += live ? 1 : 0;
    }
}

__device__ __forceinline__ void wave_walk(const StepArgs &a, const StepCtx &c, int32_t s,
                                          int32_t e0, int32_t e1) {
    int32_t lane = lane_id();
    for (int32_t b = e0; b < e1; b += 64) push_edge(a, c, s, b + lane, b + lane < e1);
}

// ---------------------------------------------------------------- step kernel
constexpr int kStepBlock = 256;

__global__ void __launch_bounds__(kStepBlock) bfs_step_kernel(StepArgs a) {
    BfsState *st = a.st;
    if (st->done) return;
    StepCtx c;
    c.level = st->level;
    const int32_t mode = st->mode;
    const int32_t p = c.level & 1;
    const int32_t *qin = a.q[p];
    c.qout = a.q[p ^ 1];
    const int32_t n = st->count[p];
    c.count_out = &st->count[p ^ 1];
    const uint64_t *bm_in = a.bm[c.level % 3];
    c.bm_out = a.bm[(c.level + 1) % 3];
    uint64_t *bm_clear = a.bm[(c.level + 2) % 3];
    c.fdeg_out = &st->fdeg[p ^ 1];
    const int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t threads = gridDim.x * blockDim.x;
    const int32_t lane = lane_id();

    // the bitmap two levels ahead is free: clear it for the next step
    for (int32_t w = tid; w < a.words; w += threads) bm_clear[w] = 0;

    if (mode == BFS_PUSH_THREAD) {
        for (int32_t i = tid; i - lane < n; i += threads) {
            bool act = i < n;
            int32_t s = act ? qin[i] : 0;
            lane_walk(a, c, s, act ? a.off[s] : 0, act ? a.off[s + 1] : 0);
        }
    } else if (mode == BFS_PUSH_WAVE) {
        for (int32_t i = tid >> 6; i < n; i += threads >> 6) {
            int32_t s = qin[i];
            wave_walk(a, c, s, a.off[s], a.off[s + 1]);
        }
    } else if (mode == BFS_PUSH_TWC) {
        // lane bin (< 64), wave bin (< kHubDegree), grid bin (deferred)
        for (int32_t i = tid; i - lane < n; i += threads) {
            bool act = i < n;
            int32_t s = act ? qin[i] : 0;
            int32_t e0 = act ? a.off[s] : 0, e1 = act ? a.off[s + 1] : 0;
            int32_t dg = e1 - e0;
            bool hub = act && dg >= kHubDegree;
            int32_t hpos = wave_atomic_add(&st->hub_count, hub);
            if (hub && hpos < kMaxHubs) st->hubs[hpos] = s;
            bool wide = act && dg >= 64 && (!hub || hpos >= kMaxHubs);
            uint64_t wm = __ballot(wide);
            while (wm) {  // the wave walks each wide vertex together
                int32_t l = __ffsll((unsigned long long)wm) - 1;
                wm &= wm - 1;
                int32_t ws = __shfl(s, l), w0 = __shfl(e0, l), w1 = __shfl(e1, l);
                wave_walk(a, c, ws, w0, w1);
            }
            bool small = act && dg < 64;
            lane_walk(a, c, s, small ? e0 : 0, small ? e1 : 0);
        }
    } else if (mode == BFS_PUSH_TWCE || mode == BFS_PUSH_CM) {
        __shared__ int32_t sv[kStepBlock], soff[kStepBlock], sscan[kStepBlock + 1];
        for (int32_t base = blockIdx.x * kStepBlock; base < n; base += gridDim.x * kStepBlock) {
            int32_t i = base + (int32_t)threadIdx.x;
            int32_t dg = 0;
            if (i < n) {
                int32_t s = qin[i];
                int32_t o0 = a.off[s];
                dg = a.off[s + 1] - o0;
                sv[threadIdx.x] = s;
                soff[threadIdx.x] = o0;
            }
            int32_t total = 0;
            int32_t ex = block_exclusive_scan(dg, &total);
            sscan[threadIdx.x] = ex;
            int32_t width = n - base < kStepBlock ? n - base : kStepBlock;
            __syncthreads();
            if (mode == BFS_PUSH_CM) {
                // merge-path over the workgroup's concatenated edge lists
                for (int32_t k = threadIdx.x; k - (int32_t)threadIdx.x < total; k += kStepBlock) {
                    bool act = k < total;
                    int32_t s = 0, e = 0;
                    if (act) {
                        int32_t j = upper_bound_i32(sscan, 0, width, k) - 1;
                        s = sv[j];
                        e = soff[j] + (k - sscan[j]);
                    }
                    push_edge(a, c, s, e, act);
                }
            } else {
                // TWCE: workgroup granule for degree >= block, wave granule
                // for >= 64, lanes for the rest
                int32_t wid = threadIdx.x >> 6;
                for (int32_t j = 0; j < width; ++j) {  // workgroup granule (uniform)
                    int32_t dj = (j + 1 < width ? sscan[j + 1] : total) - sscan[j];
                    if (dj < kStepBlock) continue;
                    int32_t s = sv[j], o = soff[j];
                    for (int32_t b = 0; b < dj; b += kStepBlock) {
                        int32_t k = b + (int32_t)threadIdx.x;
                        push_edge(a, c, s, o + k, k < dj);
                    }
                }
                for (int32_t j = wid; j < width; j += kStepBlock / 64) {  // wave granule
                    int32_t dj = (j + 1 < width ? sscan[j + 1] : total) - sscan[j];
                    if (dj < 64 || dj >= kStepBlock) continue;
                    wave_walk(a, c, sv[j], soff[j], soff[j] + dj);
                }
                bool mine = (int32_t)threadIdx.x < width && dg > 0 && dg < 64;
                int32_t o = mine ? soff[threadIdx.x] : 0;
                lane_walk(a, c, mine ? sv[threadIdx.x] : 0, o, mine ? o + dg : 0);
            }
            __syncthreads();
        }
    } else {
        // pull over all vertices; waves own 64-vertex words of the bitmaps
        unsigned long long fdeg_acc = 0;
        for (int32_t w = tid >> 6; w < a.words; w += threads >> 6) {
            uint64_t vis = a.visited[w];
            if (vis == ~0ull) continue;  // wave-uniform
            uint64_t found = 0;
            if (mode == BFS_PULL_THREAD) {
                int32_t v = w * 64 + lane;
                bool want = v < a.V && !((vis >> lane) & 1ull);
                int32_t e = want ? a.off[v] : 0, e1 = want ? a.off[v + 1] : 0;
                bool hit = false;
                int32_t par = -1;
                for (; e < e1; ++e) {
                    int32_t u = a.nbr[e];
                    if ((bm_in[u >> 6] >> (u & 63)) & 1ull) {
                        hit = true;
                        par = u;
                        break;
                    }
                }
                found = __ballot(hit);
                if (hit) {
                    a.depth[v] = c.level + 1;
                    if (a.parent) a.parent[v] = par;
                    fdeg_acc += (unsigned long long)(e1 - a.off[v]);
                }
            } else {
                const int32_t sub = lane >> 3, g8 = lane & 7;
                for (int32_t r = 0; r < 8; ++r) {
                    int32_t slot = r * 8 + sub;
                    int32_t v = w * 64 + slot;
                    bool want = v < a.V && !((vis >> slot) & 1ull);
                    int32_t e0 = want ? a.off[v] : 0, e1 = want ? a.off[v + 1] : 0;
                    bool hit = false;
                    int32_t par = -1;
                    for (int32_t b = e0; b < e1; b += 8) {  // group-uniform trip count
                        int32_t e = b + g8;
                        bool h = false;
                        int32_t u = 0;
                        if (e < e1) {
                            u = a.nbr[e];
                            h = (bm_in[u >> 6] >> (u & 63)) & 1ull;
                        }
                        uint32_t gm = (uint32_t)(__ballot(h) >> (sub * 8)) & 0xffu;
                        if (gm) {
                            if (g8 == __ffs(gm) - 1) {
                                hit = true;
                                par = u;
                            }
                            break;
                        }
                    }
                    uint64_t hm = __ballot(hit);
                    for (int32_t sg = 0; sg < 8; ++sg)
                        if ((hm >> (sg * 8)) & 0xffull) found |= 1ull << (r * 8 + sg);
                    if (hit) {
                        a.depth[v] = c.level + 1;
                        if (a.parent) a.parent[v] = par;
                        fdeg_acc += (unsigned long long)(e1 - e0);
                    }
                }
            }
            if (found) {
                if (lane == 0) {
                    a.visited[w] = vis | found;  // this wave owns word w in pull steps
                    c.bm_out[w] = found;
                }
                bool mine = (found >> lane) & 1ull;
                int32_t pos = wave_atomic_add(c.count_out, mine);
                if (mine) c.qout[pos] = w * 64 + lane;
            }
        }
        for (int o = 32; o > 0; o >>= 1) fdeg_acc += __shfl_xor(fdeg_acc, o);
        if (lane == 0 && fdeg_acc) atomicAdd(c.fdeg_out, fdeg_acc);
    }
}

// Grid bin of PUSH_TWC: the edges of the deferred hubs, spread evenly over
// the grid. Every workgroup scans the hub degrees in LDS and walks its share
// of the concatenated edge range.
__global__ void __launch_bounds__(kStepBlock) bfs_hub_kernel(StepArgs a) {
    BfsState *st = a.st;
    if (st->done) return;
    int32_t nh = st->hub_count;
    if (nh <= 0) return;
    if (nh > kMaxHubs) nh = kMaxHubs;
    StepCtx c;
    c.level = st->level;
    const int32_t p = c.level & 1;
    c.qout = a.q[p ^ 1];
    c.count_out = &st->count[p ^ 1];
    c.bm_out = a.bm[(c.level + 1) % 3];
    c.fdeg_out = &st->fdeg[p ^ 1];
    __shared__ int32_t hscan[kMaxHubs + 1];
    __shared__ int32_t hv[kMaxHubs];
    // serial-chunk scan: each thread sums a contiguous slice, then one scan
    int32_t per = (nh + kStepBlock - 1) / kStepBlock;
    int32_t lo = threadIdx.x * per, hi = lo + per < nh ? lo + per : nh;
    int32_t part = 0;
    for (int32_t i = lo; i < hi; ++i) {
        int32_t s = st->hubs[i];
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
    int32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    int32_t k0 = blockIdx.x * chunk, k1 = k0 + chunk < total ? k0 + chunk : total;
    for (int32_t k = k0 + (int32_t)threadIdx.x; k - (int32_t)threadIdx.x < k1; k += kStepBlock) {
        bool act = k < k1;
        int32_t s = 0, e = 0;
        if (act) {
            int32_t j = upper_bound_i32(hscan, 0, nh, k) - 1;
            s = hv[j];
            e = a.off[s] + (k - hscan[j]);
        }
        push_edge(a, c, s, e, act);
    }
}

// ---------------------------------------------------------------- decide kernel
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}

// One wave, after step `level` (or after seeding, level = -1): closes the
// step's bookkeeping, ends the run on an empty frontier, else picks the next
// step's mode and resets its output counters.
__global__ void bfs_decide_kernel(BfsState *st, SelectorView sel, int32_t V, int64_t E) {
    if (st->done) return;
    const int32_t lane = lane_id();
    const int32_t next_level = st->level + 1;
    const int32_t p = next_level & 1;
    const int32_t n = st->count[p];
    const unsigned long long fdeg = st->fdeg[p];
    const int32_t prev_n = next_level > 0 ? st->count[p ^ 1] : 0;
    const long long unv = st->unvisited - n;
    const long long unv_e = st->unvisited_edges - (long long)fdeg;
    __shared__ float feat[kSelInputs];
    __shared__ float h1[kSelHidden];
    __shared__ float h2[kSelHidden];
    __shared__ float y[kSelMaxOut];
    if (n == 0) {
        if (lane == 0) {
            st->done = 1;
            st->levels = next_level;
        }
        return;
    }
    int32_t mode;
    bool learned = sel.mode == GI_SELECT_LEARNED && sel.w1 != nullptr;
    if (!learned) {
        if (sel.mode == GI_SELECT_S1) mode = sel.mode_s1;
        else if (sel.mode == GI_SELECT_S2) mode = sel.mode_s2;
        else mode = ((double)n <= (double)sel.threshold * (double)V) ? sel.mode_s1 : sel.mode_s2;
    } else {
        if (lane == 0) {
            float raw[kSelInputs];
            bfs_features(raw, n, fdeg, unv, unv_e, next_level, st->mode, prev_n, V, E);
            for (int k = 0; k < kSelInputs; ++k) feat[k] = (raw[k] - sel.feat_mean[k]) * sel.feat_scale[k];
        }
        __syncthreads();
        float acc = bf16_to_f32(sel.b1[lane]);
        for (int k = 0; k < kSelInputs; ++k) acc += feat[k] * bf16_to_f32(sel.w1[k * kSelHidden + lane]);
        h1[lane] = acc > 0.f ? acc : 0.f;
        __syncthreads();
        acc = bf16_to_f32(sel.b2[lane]);
        for (int k = 0; k < kSelHidden; ++k) acc += h1[k] * bf16_to_f32(sel.w2[k * kSelHidden + lane]);
        h2[lane] = acc > 0.f ? acc : 0.f;
        __syncthreads();
        if (lane < sel.n_out) {
            acc = bf16_to_f32(sel.b3[lane]);
            for (int k = 0; k < kSelHidden; ++k) acc += h2[k] * bf16_to_f32(sel.w3[k * sel.n_out + lane]);
            y[lane] = acc;
        }
        __syncthreads();
        mode = sel.mode_s1;
        float best = 3.4e38f;
        for (int j = 0; j < sel.n_out; ++j)
            if (((sel.allowed >> j) & 1) && y[j] < best) {
                best = y[j];
                mode = j;
            }
    }
    if (lane == 0) {
        st->unvisited = unv;
        st->unvisited_edges = unv_e;
        st->mode = mode;
        st->level = next_level;
        st->hub_count = 0;
        if (next_level < kBfsMaxLevels) {
            st->hist_mode[next_level] = mode;
            st->hist_next[next_level] = n;
        }
        st->count[p ^ 1] = 0;
        st->fdeg[p ^ 1] = 0;
    }
}

__global__ void bfs_seed_fused_kernel(BfsState *st, int32_t *depth, int32_t *parent,
                                      uint64_t *visited, uint64_t *bm0, int32_t *q0,
                                      const int32_t *off, int32_t src, int32_t V, int64_t E) {
    depth[src] = 0;
    if (parent) parent[src] = src;
    visited[src >> 6] |= 1ull << (src & 63);
    bm0[src >> 6] |= 1ull << (src & 63);
    q0[0] = src;
    st->count[0] = 1;
    st->count[1] = 0;
    st->fdeg[0] = (unsigned long long)(off[src + 1] - off[src]);
    st->fdeg[1] = 0;
    st->done = 0;
    st->level = -1;
    st->levels = 0;
    st->mode = BFS_PUSH_THREAD;
    st->unvisited = V;
    st->unvisited_edges = E;
    st->hub_count = 0;
}

// ---------------------------------------------------------------- host side
struct FusedWork {
    BfsState *st = nullptr;
    uint64_t *bm[3] = {nullptr, nullptr, nullptr};
    int32_t *q[2] = {nullptr, nullptr};
    int32_t *host_flag = nullptr;  // pinned copy of `done`
};

static std::map<const gi_graph *, FusedWork *> g_fused;

static int ensure_fused(gi_graph *g, FusedWork **out) {
    auto it = g_fused.find(g);
    if (it != g_fused.end()) { *out = it->second; return GI_OK; }
    FusedWork *w = new FusedWork();
    int64_t words = bitmap_words(g->V);
    int rc = dalloc(&w->st, 1);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = dalloc(&w->bm[i], (size_t)words + 1);
    for (int i = 0; i < 2 && rc == GI_OK; ++i) rc = dalloc(&w->q[i], (size_t)g->V + 64);
    if (rc == GI_OK && !g->d_visited) rc = dalloc(&g->d_visited, (size_t)words + 1);
    if (rc == GI_OK && hipHostMalloc((void **)&w->host_flag, sizeof(int32_t) * 4) != hipSuccess)
        rc = fail(GI_ERR_OOM, "pinned allocation failed");
    if (rc != GI_OK) {
        dfree(w->st);
        for (auto p : w->bm) dfree(p);
        for (auto p : w->q) dfree(p);
        delete w;
        return rc;
    }
    g_fused[g] = w;
    *out = w;
    return GI_OK;
}

void release_fused(const gi_graph *g) {
    auto it = g_fused.find(g);
    if (it == g_fused.end()) return;
    FusedWork *w = it->second;
    dfree(w->st);
    for (auto p : w->bm) dfree(p);
    for (auto p : w->q) dfree(p);
    if (w->host_flag) (void)hipHostFree(w->host_flag);
    delete w;
    g_fused.erase(it);
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

int bfs_fused_run(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, SelectorView sel,
                  int32_t *d_depth, int32_t *d_parent, hipStream_t s, FusedTrace *trace,
                  bool timed, int32_t *levels_out) {
    FusedWork *w = nullptr;
    GI_TRY(ensure_fused(g, &w));
    int64_t words = bitmap_words(g->V);
    sel.mode_s1 = fused_mode_of(h->s1);
    sel.mode_s2 = fused_mode_of(h->s2);
    sel.threshold = h->threshold;
    if (sel.mode != GI_SELECT_LEARNED) sel.mode = h->selector;
    GI_TRY(fill_i32(d_depth, g->V, -1, s));
    if (d_parent) GI_TRY(fill_i32(d_parent, g->V, -1, s));
    GI_TRY(fill_u64(g->d_visited, words, 0, s));
    for (int i = 0; i < 3; ++i) GI_TRY(fill_u64(w->bm[i], words, 0, s));
    bfs_seed_fused_kernel<<<1, 1, 0, s>>>(w->st, d_depth, d_parent, g->d_visited, w->bm[0], w->q[0],
                                          g->d_off, source, g->V, g->E);
    bfs_decide_kernel<<<1, 64, 0, s>>>(w->st, sel, g->V, g->E);
    GI_LAUNCH_CHECK();
    StepArgs a;
    a.st = w->st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.V = g->V;
    a.words = (int32_t)words;
    a.depth = d_depth;
    a.parent = d_parent;
    a.visited = g->d_visited;
    for (int i = 0; i < 3; ++i) a.bm[i] = w->bm[i];
    for (int i = 0; i < 2; ++i) a.q[i] = w->q[i];
    int grid = persistent_grid(kStepBlock, 8);
    int32_t *done_flag = &w->st->done;
    for (int32_t level = 0; level < kBfsMaxLevels; ++level) {
        hipEvent_t *ev = nullptr;
        if (trace && timed) {
            hipEvent_t e0, e1;
            GI_HIP(hipEventCreate(&e0));
            GI_HIP(hipEventCreate(&e1));
            trace->ev.push_back(e0);
            trace->ev.push_back(e1);
            ev = &trace->ev[trace->ev.size() - 2];
            GI_HIP(hipEventRecord(ev[0], s));
        }
        bfs_step_kernel<<<grid, kStepBlock, 0, s>>>(a);
        bfs_hub_kernel<<<grid, kStepBlock, 0, s>>>(a);
        if (ev) GI_HIP(hipEventRecord(ev[1], s));
        bfs_decide_kernel<<<1, 64, 0, s>>>(w->st, sel, g->V, g->E);
        GI_LAUNCH_CHECK();
        // one 4-byte read per level tells the host whether to go on
        GI_HIP(hipMemcpyAsync(w->host_flag, done_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GI_HIP(hipStreamSynchronize(s));
        if (*w->host_flag) break;
    }
    int32_t levels = 0;
    GI_HIP(hipMemcpyAsync(&levels, &w->st->levels, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GI_HIP(hipStreamSynchronize(s));
    if (levels_out) *levels_out = levels;
    if (trace) {
        int32_t n = levels < kBfsMaxLevels ? levels : kBfsMaxLevels;
        trace->modes.resize(n);
        trace->sizes.resize(n);
        GI_HIP(hipMemcpy(trace->modes.data(), w->st->hist_mode, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        GI_HIP(hipMemcpy(trace->sizes.data(), w->st->hist_next, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    }
    return GI_OK;
}

// ---------------------------------------------------------------- sweep
__global__ void bfs_force_mode_kernel(BfsState *st, int32_t mode) {
    int32_t p = (st->level & 1) ^ 1;
    st->mode = mode;
    st->count[p] = 0;
    st->fdeg[p] = 0;
    st->hub_count = 0;
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
    StepArgs a;
    a.st = w->st;
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.V = g->V;
    a.words = (int32_t)words;
    a.depth = depth;
    a.parent = nullptr;
    a.visited = g->d_visited;
    for (int i = 0; i < 3; ++i) a.bm[i] = w->bm[i];
    for (int i = 0; i < 2; ++i) a.q[i] = w->q[i];
    int grid = persistent_grid(kStepBlock, 8);
    if (rc == GI_OK) rc = fill_i32(depth, g->V, -1, s);
    if (rc == GI_OK) rc = fill_u64(g->d_visited, words, 0, s);
    for (int i = 0; i < 3 && rc == GI_OK; ++i) rc = fill_u64(w->bm[i], words, 0, s);
    if (rc == GI_OK) {
        bfs_seed_fused_kernel<<<1, 1, 0, s>>>(w->st, depth, nullptr, g->d_visited, w->bm[0], w->q[0],
                                              g->d_off, source, g->V, g->E);
        bfs_decide_kernel<<<1, 64, 0, s>>>(w->st, sel, g->V, g->E);
    }
    int32_t nl = 0, prev_mode = BFS_PUSH_THREAD, prev_n = 0;
    while (rc == GI_OK && nl < max_levels) {
        SweepHost h;
        if (hipMemcpyAsync(&h, w->st, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = fail(GI_ERR_HIP, "state read"); break; }
        if (h.done) break;
        int32_t p = h.level & 1;
        bfs_raw_features(features + (int64_t)nl * kSelInputs, h.count[p], h.fdeg[p], h.unvisited,
                         h.unvisited_edges, h.level, prev_mode, prev_n, g->V, g->E);
        // snapshot, then time every mode of the mask from the same state
        GI_HIP(hipMemcpyAsync(depth_bak, depth, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice, s));
        GI_HIP(hipMemcpyAsync(vis_bak, g->d_visited, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, s));
        float best = 3.4e38f;
        int32_t best_mode = -1;
        for (int32_t m = 0; m < BFS_NUM_MODES; ++m) {
            float *t = times_ms + (int64_t)nl * BFS_NUM_MODES + m;
            *t = -1.f;
            if (!((mode_mask >> m) & 1)) continue;
            float tmin = 3.4e38f;
            for (int32_t r = 0; r < reps; ++r) {
                GI_HIP(hipMemcpyAsync(depth, depth_bak, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice, s));
                GI_HIP(hipMemcpyAsync(g->d_visited, vis_bak, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, s));
                GI_TRY(fill_u64(w->bm[(h.level + 1) % 3], words, 0, s));
                bfs_force_mode_kernel<<<1, 1, 0, s>>>(w->st, m);
                GI_HIP(hipEventRecord(e0, s));
                bfs_step_kernel<<<grid, kStepBlock, 0, s>>>(a);
                bfs_hub_kernel<<<grid, kStepBlock, 0, s>>>(a);
                GI_HIP(hipEventRecord(e1, s));
                GI_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                GI_HIP(hipEventElapsedTime(&ms, e0, e1));
                tmin = ms < tmin ? ms : tmin;
            }
            *t = tmin;
            if (tmin < best) { best = tmin; best_mode = m; }
        }
        int32_t run = follow >= 0 ? follow : best_mode;
        GI_HIP(hipMemcpyAsync(depth, depth_bak, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice, s));
        GI_HIP(hipMemcpyAsync(g->d_visited, vis_bak, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, s));
        GI_TRY(fill_u64(w->bm[(h.level + 1) % 3], words, 0, s));
        bfs_force_mode_kernel<<<1, 1, 0, s>>>(w->st, run);
        bfs_step_kernel<<<grid, kStepBlock, 0, s>>>(a);
        bfs_hub_kernel<<<grid, kStepBlock, 0, s>>>(a);
        bfs_decide_kernel<<<1, 64, 0, s>>>(w->st, sel, g->V, g->E);
        GI_LAUNCH_CHECK();
        prev_mode = run;
        prev_n = h.count[p];
        ++nl;
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
