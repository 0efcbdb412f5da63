// Fused app drivers: the five reference apps (apps/{bfs,sssp,cc,pr,bc}.cpp)
// with their loops on the GPU side of the C ABI.
#include <climits>
#include <vector>

#include "gi_hash.h"

#include "gi_bfs_fused.h"
#include "gi_lb.h"

namespace gi {

static int ensure_work(gi_graph *g) {
    if (!g->work) GI_TRY(gi_frontier_create(g->V, 0, &g->work));
    if (!g->d_visited) GI_TRY(dalloc(&g->d_visited, (size_t)bitmap_words(g->V) + 1));
    return GI_OK;
}

__global__ void bfs_seed_kernel(int32_t *depth, int32_t *parent, uint64_t *visited, int32_t src) {
    depth[src] = 0;
    if (parent) parent[src] = src;
    visited[src >> 6] |= 1ull << (src & 63);
}

__global__ void count_reached_kernel(const int32_t *depth, int32_t V, int32_t *out) {
    int32_t c = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        c += depth[v] >= 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0) atomicAdd(out, c);
}

static void default_bfs_schedule(gi_hybrid_schedule *h) {
    gi_schedule_init(&h->s1);
    h->s1.direction = GI_PUSH;
    h->s1.load_balancing = GI_LB_TWCE;  // apps/bfs.cpp:86-88
    gi_schedule_init(&h->s2);
    h->s2.direction = GI_PULL;          // apps/bfs.cpp:75-78 ("power" mode s2)
    h->s2.pull_frontier_rep = GI_PULL_BITMAP;
    h->s2.load_balancing = GI_LB_VERTEX_BASED;
    h->s2.frontier_creation = GI_FC_BITMAP;
    h->threshold = 0.f;
    h->selector = GI_SELECT_S1;
}

// --------------------------------------------------------------- SSSP helpers
__global__ void sssp_window_min_kernel(const int32_t *dist, int32_t V, int32_t lower, int32_t *out) {
    int32_t best = INT_MAX;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        int32_t d = dist[v];
        if (d >= lower && d != INT_MAX && d < best) best = d;
    }
    for (int o = 32; o > 0; o >>= 1) {
        int32_t x = __shfl_xor(best, o);
        best = x < best ? x : best;
    }
    if (lane_id() == 0 && best != INT_MAX) atomicMin(out, best);
}

// update_nodes_special (gpu_priority_queue.h:272-288): enqueue [cur, cur+delta)
__global__ void sssp_window_enqueue_kernel(const int32_t *dist, int32_t V, int32_t lo, int32_t hi,
                                           int32_t *q, int32_t *count) {
    int64_t limit = ceil_div(V, 64) * 64;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < limit;
         v += (int64_t)gridDim.x * blockDim.x) {
        bool in = false;
        if (v < V) {
            int32_t d = dist[v];
            in = d >= lo && d < hi;
        }
        int32_t pos = wave_atomic_add(count, in);
        if (in) q[pos] = (int32_t)v;
    }
}

__global__ void set_i32_kernel(int32_t *p, int32_t v) { *p = v; }

// --------------------------------------------------------------- BC helpers
__global__ void bc_seed_kernel(double *paths, uint8_t *visited, int32_t src) {
    paths[src] = 1.0;
    visited[src] = 1;
}

__global__ void bc_accumulate_kernel(const float *dep, float *scores, int32_t V) {
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        scores[v] += dep[v];
}

static int grid_for(int64_t n, int block = 256) {
    int64_t want = ceil_div(n, block);
    int cap = persistent_grid(block);
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

}  // namespace gi

using namespace gi;

extern "C" {

}  // extern "C"

namespace gi {

struct LevelTrace {
    gi_level_record *recs = nullptr;
    int32_t max = 0;
    int32_t n = 0;
    std::vector<hipEvent_t> ev;
};

// The fused BFS loop shared by gi_bfs and gi_bfs_profile.
static int bfs_run(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, int32_t *d_depth,
                   int32_t *d_parent, gi_run_stats *stats, hipStream_t s, LevelTrace *trace) {
    if (!g || !d_depth) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (source < 0 || source >= g->V) return fail(GI_ERR_RANGE, "source %d out of range [0,%d)", source, g->V);
    gi_hybrid_schedule def;
    if (!h) { default_bfs_schedule(&def); h = &def; }
    GI_TRY(validate_schedule(&h->s1));
    GI_TRY(validate_schedule(&h->s2));
    if (h->s1.kernel_fusion == GI_FUSION_ENABLED) {
        // fused engine: schedule decisions on the device
        SelectorView sel{};
        if (h->selector == GI_SELECT_LEARNED) sel = selector_view();  // empty: heuristic fallback
        if (!sel.w1) sel.mode = h->selector == GI_SELECT_LEARNED ? GI_SELECT_HEURISTIC : h->selector;
        FusedTrace ft;
        bool want_trace = trace != nullptr || stats != nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (stats) {
            GI_HIP(hipEventCreate(&t0));
            GI_HIP(hipEventCreate(&t1));
            GI_HIP(hipEventRecord(t0, s));
        }
        // a degree-ordered layout: run on the view, map the results back
        gi_graph *rg = g;
        int32_t rsrc = source, *rdepth = d_depth, *rparent = d_parent;
        if (g->layout) {
            rg = layout_view(g);
            rsrc = layout_map_source(g, source);
            rdepth = layout_tmp(g, 0);
            rparent = d_parent ? layout_tmp(g, 1) : nullptr;
        }
        int32_t levels = 0;
        FusedEpilogue back = [&](hipStream_t st) {  // queued behind the level loop
            GI_TRY(layout_depth_back(g, rdepth, source, d_depth, st));
            return d_parent ? layout_map_back(g, rparent, d_parent, true, st) : GI_OK;
        };
        // on a symmetric graph's view the map-back reads the depths of the
        // vertices with edges only: the edgeless tail needs no clearing
        const int32_t depth_len = g->layout && !g->directed ? layout_iso_begin(g) : -1;
        int rc = bfs_fused_run(rg, rsrc, h, sel, rdepth, rparent, s, want_trace ? &ft : nullptr,
                               trace != nullptr, &levels, g->layout ? &back : nullptr, depth_len);
        float ms = 0.f;
        if (stats && rc == GI_OK) {
            GI_HIP(hipEventRecord(t1, s));
            GI_HIP(hipEventSynchronize(t1));
            GI_HIP(hipEventElapsedTime(&ms, t0, t1));
        }
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (rc == GI_OK && trace) {
            const int32_t recorded = (int32_t)ft.modes.size();  // the first kBfsMaxLevels levels
            for (int32_t i = 0; i < levels && i < recorded && trace->n < trace->max; ++i) {
                gi_level_record &r = trace->recs[trace->n++];
                r.level = i;
                r.direction = ft.modes[i] >= BFS_PULL_THREAD ? GI_PULL : GI_PUSH;
                r.load_balancing = 100 + ft.modes[i];  // fused mode id
                r.frontier_size = ft.sizes[i];
                r.next_size = i + 1 < levels && i + 1 < recorded ? ft.sizes[i + 1] : 0;
            }
            trace->ev.swap(ft.ev);
        }
        for (hipEvent_t e : ft.ev) (void)hipEventDestroy(e);
        GI_TRY(rc);
        if (stats) {
            stats->iterations = levels;
            stats->pull_iterations = ft.pull_levels;
            stats->push_iterations = levels - ft.pull_levels;
            stats->ms = ms;
            stats->edges_traversed = 0;
            GI_TRY(fused_reached(g, d_depth, &stats->reached, s));
        }
        return GI_OK;
    }
    GI_TRY(ensure_work(g));
    gi_frontier *f = g->work;
    EventTimer timer(s);
    GI_TRY(fill_i32(d_depth, g->V, -1, s));
    if (d_parent) GI_TRY(fill_i32(d_parent, g->V, -1, s));
    GI_TRY(fill_u64(g->d_visited, bitmap_words(g->V), 0, s));
    bfs_seed_kernel<<<1, 1, 0, s>>>(d_depth, d_parent, g->d_visited, source);
    GI_LAUNCH_CHECK();
    GI_TRY(gi_frontier_clear(f, s));
    GI_TRY(fill_i32(f->count_out(), 1, 0, s));
    GI_TRY(gi_frontier_add_vertex(f, source, s));
    int32_t n = 1, level = 0, npush = 0, npull = 0;
    gi_op_args dummy{};
    while (n > 0) {
        const gi_schedule *pick = &h->s1;
        if (h->selector == GI_SELECT_S2) pick = &h->s2;
        else if (h->selector == GI_SELECT_HEURISTIC || h->selector == GI_SELECT_LEARNED)
            pick = ((double)n <= (double)h->threshold * (double)g->V) ? &h->s1 : &h->s2;
        OpBfsVisit op{g->d_visited, d_depth, d_parent, level + 1};
        hipEvent_t *evs = nullptr;
        if (trace && trace->n < trace->max) {
            hipEvent_t e0, e1;
            GI_HIP(hipEventCreate(&e0));
            GI_HIP(hipEventCreate(&e1));
            trace->ev.push_back(e0);
            trace->ev.push_back(e1);
            evs = &trace->ev[trace->ev.size() - 2];
            gi_level_record &r = trace->recs[trace->n];
            r.level = level;
            r.direction = pick->direction;
            r.load_balancing = pick->load_balancing;
            r.frontier_size = n;
        }
        GI_TRY(edgeset_apply(g, pick, f, f, GI_OP_BFS_DEPTH, dummy, s, &op, evs));
        if (pick->direction == GI_PULL) ++npull; else ++npush;
        GI_TRY(read_count(f->count_in(), &n, s));
        if (evs) trace->recs[trace->n++].next_size = n;
        ++level;
    }
    int32_t ovf = 0;
    GI_TRY(read_count(f->overflow(), &ovf, s));
    if (ovf) return fail(GI_ERR_RANGE, "frontier overflow");
    double ms = timer.stop();
    if (stats) {
        stats->iterations = level;
        stats->push_iterations = npush;
        stats->pull_iterations = npull;
        stats->ms = ms;
        stats->edges_traversed = 0;
        int32_t *cnt = nullptr;
        GI_TRY(dalloc(&cnt, 1));
        int rc = fill_i32(cnt, 1, 0, s);
        if (rc == GI_OK) {
            count_reached_kernel<<<grid_for(g->V), 256, 0, s>>>(d_depth, g->V, cnt);
            rc = read_count(cnt, &stats->reached, s);
        }
        dfree(cnt);
        GI_TRY(rc);
    }
    return GI_OK;
}

// Algorithmic bytes of one BFS step, from the final depths.
// out[0] = sum over F of (12 + 8 deg) (push) or the pull terms
// 8 + 8 k_v over U; out[1] = |F'|.
__global__ void level_bytes_kernel(const int32_t *off, const int32_t *nbr, const int32_t *depth,
                                   int32_t V, int32_t level, int32_t pull,
                                   unsigned long long *out) {
    unsigned long long acc = 0, next = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        int32_t d = depth[v];
        if (d == level + 1) next += 1;
        if (!pull) {
            if (d == level) acc += 12ull + 8ull * (unsigned long long)(off[v + 1] - off[v]);
        } else if (d == -1 || d > level) {
            unsigned long long k = 0;
            for (int32_t e = off[v]; e < off[v + 1]; ++e) {
                ++k;
                if (depth[nbr[e]] == level) break;
            }
            acc += 8ull + 8ull * k;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o);
        next += __shfl_xor(next, o);
    }
    if (lane_id() == 0) {
        atomicAdd(&out[0], acc);
        atomicAdd(&out[1], next);
    }
}

__global__ void reached_edges_kernel(const int32_t *off, const int32_t *depth, int32_t V,
                                     unsigned long long *out) {
    unsigned long long acc = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        if (depth[v] >= 0) acc += (unsigned long long)(off[v + 1] - off[v]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane_id() == 0) atomicAdd(out, acc);
}

}  // namespace gi

extern "C" {

int gi_bfs(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, int32_t *d_depth,
           int32_t *d_parent, gi_run_stats *stats, void *stream) {
    return bfs_run(g, source, h, d_depth, d_parent, stats, (hipStream_t)stream, nullptr);
}

int gi_bfs_profile(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, int32_t *d_depth,
                   gi_level_record *records, int32_t max_records, int32_t *num_records,
                   void *stream) {
    if (!records || !num_records || max_records < 0) return fail(GI_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    LevelTrace tr;
    tr.recs = records;
    tr.max = max_records;
    gi_run_stats st{};
    int rc = bfs_run(g, source, h, d_depth, nullptr, &st, s, &tr);
    if (rc == GI_OK) {
        unsigned long long *acc = nullptr;
        rc = dalloc(&acc, 2);
        for (int32_t i = 0; i < tr.n && rc == GI_OK; ++i) {
            gi_level_record &r = records[i];
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, tr.ev[2 * i], tr.ev[2 * i + 1]) != hipSuccess) {
                rc = fail(GI_ERR_HIP, "event timing");
                break;
            }
            r.kernel_ms = ms;
            r.reserved = 0;
            rc = fill_u64((uint64_t *)acc, 2, 0, s);
            if (rc != GI_OK) break;
            level_bytes_kernel<<<grid_for(g->V), 256, 0, s>>>(g->d_off, g->d_nbr, d_depth, g->V, r.level,
                                                              r.direction == GI_PULL, acc);
            unsigned long long h2[2] = {0, 0};
            if (hipMemcpyAsync(h2, acc, sizeof(h2), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(GI_ERR_HIP, "level bytes");
                break;
            }
            double bytes = (double)h2[0] + 8.0 * (double)h2[1];
            if (r.direction == GI_PULL) bytes += 4.0 * (double)g->V;
            r.alg_bytes = bytes;
        }
        dfree(acc);
    }
    for (hipEvent_t e : tr.ev) (void)hipEventDestroy(e);
    *num_records = st.iterations;
    return rc;
}

int gi_graph_pick_sources(const gi_graph *g, uint64_t seed, int32_t count, int32_t *out) {
    if (!g || (count > 0 && !out)) return fail(GI_ERR_INVALID, "null argument");
    if (g->E == 0) return fail(GI_ERR_RANGE, "graph has no edges");
    std::vector<int32_t> off((size_t)g->V + 1);
    GI_HIP(hipMemcpy(off.data(), g->d_off, sizeof(int32_t) * (g->V + 1), hipMemcpyDeviceToHost));
    uint64_t draw = 0;
    for (int32_t got = 0; got < count;) {
        int32_t v = (int32_t)uniform_below(hash64(seed, 6, draw++), (uint32_t)g->V);
        if (off[v + 1] > off[v]) out[got++] = v;
    }
    return GI_OK;
}

int gi_reached_edges(const gi_graph *g, const int32_t *d_depth, int64_t *out, void *stream) {
    if (!g || !d_depth || !out) return fail(GI_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long *acc = nullptr;
    GI_TRY(dalloc(&acc, 1));
    int rc = fill_u64((uint64_t *)acc, 1, 0, s);
    if (rc == GI_OK) reached_edges_kernel<<<grid_for(g->V), 256, 0, s>>>(g->d_off, d_depth, g->V, acc);
    unsigned long long h1 = 0;
    if (rc == GI_OK && (hipMemcpyAsync(&h1, acc, sizeof(h1), hipMemcpyDeviceToHost, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
        rc = fail(GI_ERR_HIP, "reached edges");
    dfree(acc);
    *out = (int64_t)h1;
    return rc;
}

// ------------------------------------------------------------ priority queue
}  // extern "C"

struct gi_pq {
    gi_graph *g = nullptr;
    int32_t *prio = nullptr;
    int32_t delta = 1;
    int32_t cur = 0;
    int32_t window_upper = 0;
    bool ready_dequeued = false;
    gi_frontier *f = nullptr;
    int32_t *d_min = nullptr;
};

namespace gi {
// dequeueReadySet's refill (gpu_priority_queue.h:125-148): the smallest
// priority >= cur+delta becomes cur, and every vertex in [cur, cur+delta)
// enters the frontier.
static int pq_refill(gi_pq *q, hipStream_t s) {
    q->window_upper = q->cur + q->delta;
    GI_TRY(fill_i32(q->d_min, 1, INT_MAX, s));
    sssp_window_min_kernel<<<grid_for(q->g->V), 256, 0, s>>>(q->prio, q->g->V, q->window_upper, q->d_min);
    GI_LAUNCH_CHECK();
    GI_TRY(read_count(q->d_min, &q->cur, s));
    gi_frontier *f = q->f;
    GI_TRY(fill_i32(f->count_in(), 1, 0, s));
    f->fmt = GI_FMT_SPARSE;
    if (q->cur == INT_MAX) return GI_OK;
    int64_t hi = (int64_t)q->cur + q->delta;
    sssp_window_enqueue_kernel<<<grid_for(q->g->V), 256, 0, s>>>(
        q->prio, q->g->V, q->cur, hi > INT_MAX ? INT_MAX : (int32_t)hi, f->q[f->side], f->count_in());
    GI_LAUNCH_CHECK();
    return GI_OK;
}
}  // namespace gi

extern "C" {

int gi_pq_create(gi_graph *g, int32_t *d_priorities, int32_t initial_priority, int32_t delta,
                 int32_t source, gi_pq **out) {
    if (!g || !d_priorities || !out) return fail(GI_ERR_INVALID, "null argument");
    if (delta <= 0) return fail(GI_ERR_RANGE, "delta must be positive");
    if (source < -1 || source >= g->V) return fail(GI_ERR_RANGE, "source %d out of range", source);
    gi_pq *q = new gi_pq();
    q->g = g;
    q->prio = d_priorities;
    q->delta = delta;
    q->cur = initial_priority;
    int rc = gi_frontier_create(g->V, 0, &q->f);
    if (rc == GI_OK) rc = dalloc(&q->d_min, 1);
    if (rc == GI_OK && source != -1) rc = gi_frontier_add_vertex(q->f, source, nullptr);
    if (rc != GI_OK) { gi_pq_destroy(q); return rc; }
    *out = q;
    return GI_OK;
}

int gi_pq_finished(gi_pq *q, int32_t *finished, void *stream) {
    if (!q || !finished) return fail(GI_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (q->cur == INT_MAX) { *finished = 1; return GI_OK; }
    int32_t n = 0;
    if (!q->ready_dequeued) {
        GI_TRY(gi_frontier_size(q->f, &n, stream));
        if (n == 0) {
            GI_TRY(pq_refill(q, s));
            q->ready_dequeued = true;
            *finished = q->cur == INT_MAX;
            return GI_OK;
        }
    }
    *finished = 0;
    return GI_OK;
}

int gi_pq_dequeue(gi_pq *q, gi_frontier **frontier, void *stream) {
    if (!q || !frontier) return fail(GI_ERR_INVALID, "null argument");
    *frontier = q->f;
    if (q->ready_dequeued) {
        q->ready_dequeued = false;
        return GI_OK;
    }
    int32_t n = 0;
    GI_TRY(gi_frontier_size(q->f, &n, stream));
    if (n == 0) GI_TRY(pq_refill(q, (hipStream_t)stream));
    q->ready_dequeued = false;
    return GI_OK;
}

int gi_pq_cutoff(gi_pq *q, int32_t *cutoff) {
    if (!q || !cutoff) return fail(GI_ERR_INVALID, "null argument");
    int64_t c = (int64_t)q->cur + q->delta;
    *cutoff = c > INT_MAX ? INT_MAX : (int32_t)c;
    return GI_OK;
}

int gi_pq_destroy(gi_pq *q) {
    if (!q) return GI_OK;
    if (q->f) gi_frontier_destroy(q->f);
    dfree(q->d_min);
    delete q;
    return GI_OK;
}

// apps/sssp.cpp:34-57 with the priority queue above.
int gi_sssp(gi_graph *g, int32_t source, int32_t delta, const gi_schedule *sched,
            int32_t *d_dist, gi_run_stats *stats, void *stream) {
    if (!g || !d_dist) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (!g->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    if (source < 0 || source >= g->V) return fail(GI_ERR_RANGE, "source %d out of range", source);
    if (delta <= 0) return fail(GI_ERR_RANGE, "delta must be positive");
    gi_schedule def;
    if (!sched) {  // apps/sssp.cpp:77-79
        gi_schedule_init(&def);
        def.direction = GI_PUSH;
        def.frontier_creation = GI_FC_SPARSE;
        def.load_balancing = GI_LB_TWCE;
        sched = &def;
    }
    GI_TRY(validate_schedule(sched));
    if (sched->direction != GI_PUSH) return fail(GI_ERR_UNSUPPORTED, "SSSP runs push schedules");
    hipStream_t s = (hipStream_t)stream;
    if (sched->kernel_fusion == GI_FUSION_ENABLED) {
        // fused engine: windows and phases decided on the device
        EventTimer timer(s);
        int32_t counts[4] = {0, 0, 0, 0};
        // native vertex order even when a degree-ordered copy exists: there the
        // hubs share a few cache lines and the relaxations' atomics serialize
        GI_TRY(sssp_fused_run(g, source, delta, d_dist, counts, s, sched->max_cta));
        double ms = timer.stop();
        if (stats) {
            stats->iterations = counts[0];
            stats->push_iterations = counts[0] - counts[1];
            stats->pull_iterations = counts[1];
            stats->ms = ms;
            stats->reached = 0;
            stats->edges_traversed = 0;
        }
        return GI_OK;
    }
    gi_pq *pq = nullptr;
    GI_TRY(gi_pq_create(g, d_dist, 0, delta, source, &pq));
    EventTimer timer(s);
    int rc = fill_i32(d_dist, g->V, INT_MAX, s);
    if (rc == GI_OK) set_i32_kernel<<<1, 1, 0, s>>>(d_dist + source, 0);
    gi_schedule run = *sched;
    run.frontier_creation = GI_FC_SPARSE;
    int32_t iters = 0, fin = 0;
    while (rc == GI_OK) {
        rc = gi_pq_finished(pq, &fin, s);
        if (rc != GI_OK || fin) break;
        gi_frontier *f = nullptr;
        rc = gi_pq_dequeue(pq, &f, s);
        gi_op_args a{};
        a.data0 = d_dist;
        if (rc == GI_OK) rc = gi_pq_cutoff(pq, &a.i0);
        if (rc == GI_OK) rc = edgeset_apply(g, &run, f, f, GI_OP_SSSP_RELAX, a, s, nullptr);
        ++iters;
    }
    double ms = timer.stop();
    gi_pq_destroy(pq);
    GI_TRY(rc);
    if (stats) {
        stats->iterations = iters;
        stats->push_iterations = iters;
        stats->pull_iterations = 0;
        stats->ms = ms;
        stats->reached = 0;
        stats->edges_traversed = 0;
    }
    return GI_OK;
}

int gi_cc(gi_graph *g, const gi_schedule *sched, int32_t *d_labels, gi_run_stats *stats,
          void *stream) {
    if (!g || !d_labels) return fail(GI_ERR_INVALID, "null argument");
    gi_schedule def;
    if (!sched) {  // apps/cc.cpp:84-86
        gi_schedule_init(&def);
        def.direction = GI_PUSH;
        def.load_balancing = GI_LB_EDGE_ONLY;
        sched = &def;
    }
    hipStream_t s = (hipStream_t)stream;
    if (sched->kernel_fusion == GI_FUSION_ENABLED) {
        // fused engine: hook sweeps and shortcutting with device-side termination
        EventTimer timer(s);
        int32_t rounds = 0;
        const bool sampled = sched->load_balancing == GI_LB_VERTEX_BASED;
        if (g->layout) {  // degree-ordered view: labels come back as original minimum ids
            gi_graph *view = layout_view(g);
            const std::function<int(hipStream_t)> back = [&](hipStream_t st) {
                // the sampled run's giant label is the most frequent one
                return layout_labels_back(g, layout_tmp(g, 0), d_labels, st, sampled ? cc_giant_label(view) : nullptr);
            };
            GI_TRY(cc_fused_run(view, layout_tmp(g, 0), &rounds, sampled, s, &back));
        } else {
            GI_TRY(cc_fused_run(g, d_labels, &rounds, sampled, s));
        }
        double ms = timer.stop();
        if (stats) {
            stats->iterations = rounds;
            stats->push_iterations = rounds;
            stats->pull_iterations = 0;
            stats->ms = ms;
            stats->reached = g->V;
            stats->edges_traversed = sampled ? 0 : (int64_t)rounds * g->E;  // sampled rounds skip most edges
        }
        return GI_OK;
    }
    EventTimer timer(s);
    int32_t *upd = nullptr;
    GI_TRY(dalloc(&upd, 2));
    gi_op_args a{};
    a.data0 = d_labels;
    a.data1 = upd;
    int rc = gi_vertexset_apply(g, nullptr, GI_VOP_IOTA_I32, &a, s);
    int32_t flags[2] = {1, 1};
    int32_t sweeps = 0;
    while (rc == GI_OK && flags[0]) {
        rc = fill_i32(upd, 2, 0, s);
        if (rc == GI_OK) rc = edgeset_apply(g, sched, nullptr, nullptr, GI_OP_CC_HOOK, a, s, nullptr);
        ++sweeps;
        flags[1] = 1;
        while (rc == GI_OK && flags[1]) {
            rc = fill_i32(upd + 1, 1, 0, s);
            if (rc == GI_OK) rc = gi_vertexset_apply(g, nullptr, GI_VOP_CC_PJUMP, &a, s);
            if (rc == GI_OK && hipMemcpyAsync(flags, upd, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) rc = fail(GI_ERR_HIP, "read flags");
            if (rc == GI_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(GI_ERR_HIP, "sync");
        }
    }
    dfree(upd);
    GI_TRY(rc);
    double ms = timer.stop();
    if (stats) {
        stats->iterations = sweeps;
        stats->push_iterations = sweeps;
        stats->pull_iterations = 0;
        stats->ms = ms;
        stats->reached = g->V;
        stats->edges_traversed = (int64_t)sweeps * g->E;
    }
    return GI_OK;
}

int gi_pagerank(gi_graph *g, int32_t iters, float damping, const gi_schedule *sched,
                float *d_rank, gi_run_stats *stats, void *stream) {
    if (!g || !d_rank) return fail(GI_ERR_INVALID, "null argument");
    if (iters < 0) return fail(GI_ERR_RANGE, "negative iteration count");
    gi_schedule def;
    if (!sched) {  // apps/pr.cpp:78-81
        gi_schedule_init(&def);
        def.direction = GI_PULL;
        def.load_balancing = GI_LB_EDGE_ONLY;
        def.edge_blocking = GI_BLOCKED;
        def.block_size = 0x42000;
        sched = &def;
    }
    hipStream_t s = (hipStream_t)stream;
    if (sched->kernel_fusion == GI_FUSION_ENABLED) {
        // fused engine: deterministic pull over the in-edges, one recorded graph
        if (sched->direction != GI_PULL) return fail(GI_ERR_UNSUPPORTED, "fused PageRank pulls");
        EventTimer timer(s);
        if (g->layout) {  // degree-ordered view; ranks gathered back (bit copies)
            float *rv = reinterpret_cast<float *>(layout_tmp(g, 0));
            GI_TRY(pr_fused_run(layout_view(g), iters, damping, rv, s));
            GI_TRY(layout_map_back(g, reinterpret_cast<const int32_t *>(rv), reinterpret_cast<int32_t *>(d_rank),
                                   false, s));
        } else {
            GI_TRY(pr_fused_run(g, iters, damping, d_rank, s));
        }
        double ms = timer.stop();
        if (stats) {
            stats->iterations = iters;
            stats->push_iterations = 0;
            stats->pull_iterations = iters;
            stats->ms = ms;
            stats->reached = g->V;
            stats->edges_traversed = (int64_t)iters * g->E;
        }
        return GI_OK;
    }
    float *nw = nullptr, *contrib = nullptr;
    GI_TRY(dalloc(&nw, (size_t)g->V));
    int rc = dalloc(&contrib, (size_t)g->V);
    EventTimer timer(s);
    gi_op_args a{};
    a.data0 = d_rank;
    a.data1 = nw;
    if (rc == GI_OK) rc = gi_vertexset_apply(g, nullptr, GI_VOP_PR_RESET, &a, s);
    for (int32_t it = 0; it < iters && rc == GI_OK; ++it) {
        gi_op_args c{};
        c.data0 = d_rank;
        c.data1 = contrib;
        rc = gi_vertexset_apply(g, nullptr, GI_VOP_PR_CONTRIB, &c, s);
        gi_op_args e{};
        e.data0 = nw;
        e.data1 = contrib;
        if (rc == GI_OK) rc = edgeset_apply(g, sched, nullptr, nullptr, GI_OP_PR_PULL, e, s, nullptr);
        gi_op_args u{};
        u.data0 = d_rank;
        u.data1 = nw;
        u.f0 = damping;
        if (rc == GI_OK) rc = gi_vertexset_apply(g, nullptr, GI_VOP_PR_UPDATE, &u, s);
    }
    double ms = timer.stop();
    dfree(nw);
    dfree(contrib);
    GI_TRY(rc);
    if (stats) {
        stats->iterations = iters;
        stats->push_iterations = 0;
        stats->pull_iterations = iters;
        stats->ms = ms;
        stats->reached = g->V;
        stats->edges_traversed = (int64_t)iters * g->E;
    }
    return GI_OK;
}

static int bc_check_args(gi_graph *g, const int32_t *h_sources, int32_t num_sources, float *d_scores) {
    if (!g || !d_scores || (num_sources > 0 && !h_sources)) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    for (int32_t i = 0; i < num_sources; ++i)
        if (h_sources[i] < 0 || h_sources[i] >= g->V) return fail(GI_ERR_RANGE, "source %d out of range", h_sources[i]);
    return GI_OK;
}

// Sources in batches on the device-driven engine (batch 1: fused single
// source); d_scores zeroed first.
static int bc_on_device(gi_graph *g, const int32_t *h_sources, int32_t num_sources, int32_t batch,
                        float *d_scores, gi_run_stats *stats, hipStream_t s) {
    EventTimer timer(s);
    GI_TRY(fill_f32(d_scores, g->V, 0.f, s));
    int64_t levels = 0;
    GI_TRY(bc_batched(g, h_sources, num_sources, batch, d_scores, s, &levels));
    double ms = timer.stop();
    if (stats) {
        stats->iterations = num_sources;
        stats->push_iterations = levels;
        stats->pull_iterations = 0;
        stats->ms = ms;
        stats->reached = 0;
        stats->edges_traversed = 0;
    }
    return GI_OK;
}

// One source at a time with a host round trip per level and per list
// operation, as the app runs unfused (apps/bc.cpp:64-120): push steps with
// the schedule's load balancing.
static int bc_unfused(gi_graph *g, const int32_t *h_sources, int32_t num_sources, const gi_schedule &sched,
                      float *d_scores, gi_run_stats *stats, hipStream_t s) {
    double *paths = nullptr;
    float *dep = nullptr;
    uint8_t *visited = nullptr;
    gi_frontier *f = nullptr;
    gi_frontier_list *list = nullptr;
    int rc = dalloc(&paths, (size_t)g->V);
    if (rc == GI_OK) rc = dalloc(&dep, (size_t)g->V);
    if (rc == GI_OK) rc = dalloc(&visited, (size_t)g->V + 64);
    if (rc == GI_OK) rc = gi_frontier_create(g->V, 0, &f);
    if (rc == GI_OK) rc = gi_frontier_list_create(g->V + 1, &list);
    EventTimer timer(s);
    if (rc == GI_OK) rc = fill_f32(d_scores, g->V, 0.f, s);
    int32_t total_rounds = 0;
    for (int32_t si = 0; si < num_sources && rc == GI_OK; ++si) {
        int32_t src = h_sources[si];
        gi_op_args va{};
        va.data0 = dep;
        va.data1 = paths;
        va.data2 = visited;
        rc = gi_vertexset_apply(g, nullptr, GI_VOP_BC_RESET, &va, s);
        bc_seed_kernel<<<1, 1, 0, s>>>(paths, visited, src);
        // every source leaves the level list empty (L+2 inserts, L+2 retrieves)
        int32_t left = 0;
        if (rc == GI_OK) rc = gi_frontier_list_levels(list, &left);
        if (rc == GI_OK && left != 0) rc = fail(GI_ERR_INVALID, "level list not drained");
        if (rc == GI_OK) rc = gi_frontier_clear(f, s);
        if (rc == GI_OK) rc = gi_frontier_add_vertex(f, src, s);
        if (rc == GI_OK) rc = gi_frontier_list_insert(list, f, s);
        int32_t round = 0, n = 1;
        while (rc == GI_OK && n != 0) {  // forward, apps/bc.cpp:93-99
            ++round;
            gi_op_args fa{};
            fa.data0 = paths;
            fa.data1 = visited;
            rc = edgeset_apply(g, &sched, f, f, GI_OP_BC_FORWARD, fa, s, nullptr);
            gi_op_args ma{};
            ma.data2 = visited;
            ma.i0 = 1;
            if (rc == GI_OK) rc = gi_vertexset_apply(g, f, GI_VOP_BC_MARK, &ma, s);
            if (rc == GI_OK) rc = gi_frontier_list_insert(list, f, s);
            if (rc == GI_OK) rc = read_count(f->count_in(), &n, s);
        }
        gi_op_args ua{};
        ua.data2 = visited;
        ua.i0 = 0;
        if (rc == GI_OK) rc = gi_vertexset_apply(g, nullptr, GI_VOP_BC_MARK, &ua, s);
        if (rc == GI_OK) rc = gi_frontier_list_retrieve(list, f, s);
        if (rc == GI_OK) rc = gi_frontier_list_retrieve(list, f, s);
        if (rc == GI_OK) rc = gi_vertexset_apply(g, f, GI_VOP_BC_BACKWARD_V, &va, s);
        --round;
        while (rc == GI_OK && round > 0) {  // backward, apps/bc.cpp:108-114
            gi_op_args ba{};
            ba.data0 = dep;
            ba.data1 = visited;
            rc = edgeset_apply(g, &sched, f, nullptr, GI_OP_BC_BACKWARD, ba, s, nullptr);
            if (rc == GI_OK) rc = gi_frontier_list_retrieve(list, f, s);
            if (rc == GI_OK) rc = gi_vertexset_apply(g, f, GI_VOP_BC_BACKWARD_V, &va, s);
            --round;
        }
        if (rc == GI_OK) rc = gi_vertexset_apply(g, nullptr, GI_VOP_BC_FINAL, &va, s);
        if (rc == GI_OK) bc_accumulate_kernel<<<grid_for(g->V), 256, 0, s>>>(dep, d_scores, g->V);
        total_rounds += round;
    }
    double ms = timer.stop();
    dfree(paths);
    dfree(dep);
    dfree(visited);
    gi_frontier_destroy(f);
    gi_frontier_list_destroy(list);
    GI_TRY(rc);
    if (stats) {
        stats->iterations = num_sources;
        stats->push_iterations = total_rounds;
        stats->pull_iterations = 0;
        stats->ms = ms;
        stats->reached = 0;
        stats->edges_traversed = 0;
    }
    return GI_OK;
}

int gi_bc(gi_graph *g, const int32_t *h_sources, int32_t num_sources, int32_t batch,
          float *d_scores, gi_run_stats *stats, void *stream) {
    GI_TRY(bc_check_args(g, h_sources, num_sources, d_scores));
    hipStream_t s = (hipStream_t)stream;
    if (batch > 1) return bc_on_device(g, h_sources, num_sources, batch, d_scores, stats, s);
    gi_schedule sched;  // the app's schedule, apps/bc.cpp:149-151
    gi_schedule_init(&sched);
    sched.direction = GI_PUSH;
    sched.load_balancing = GI_LB_TWCE;
    return bc_unfused(g, h_sources, num_sources, sched, d_scores, stats, s);
}

int gi_bc_schedule(gi_graph *g, const int32_t *h_sources, int32_t num_sources, const gi_schedule *sched,
                   float *d_scores, gi_run_stats *stats, void *stream) {
    GI_TRY(bc_check_args(g, h_sources, num_sources, d_scores));
    hipStream_t s = (hipStream_t)stream;
    gi_schedule def;
    gi_schedule_init(&def);
    def.direction = GI_PUSH;
    def.load_balancing = GI_LB_TWCE;
    const gi_schedule &sc = sched ? *sched : def;
    if (sc.kernel_fusion == GI_FUSION_ENABLED) return bc_on_device(g, h_sources, num_sources, 1, d_scores, stats, s);
    if (sc.direction != GI_PUSH)
        return fail(GI_ERR_UNSUPPORTED, "unfused BC runs push steps; pulled levels need kernel_fusion");
    GI_TRY(validate_schedule(&sc));
    return bc_unfused(g, h_sources, num_sources, sc, d_scores, stats, s);
}

}  // extern "C"

// ----------------------------------------------------- BFS without allocation
// gi_bfs_static: the unfused push BFS of gi_bfs (the same edgeset_apply and
// load-balanced kernels) with its frontier queues, counters and visited
// bitmap in a fixed device array instead of memory the library allocates,
// for callers that cannot allocate. Layout of `scratch` (int32 units):
//   [0, 64)                      frontier counts and overflow flag
//   [64, 64 + Vmax)              queue 0
//   [64 + Vmax, 64 + 2 Vmax)     queue 1
//   [64 + 2 Vmax, ...)           visited bitmap, Vmax / 64 words of 64 bits
// Each GPU holds one copy until the process ends; every call on a GPU uses
// that GPU's copy, so calls that overlap on one GPU race.
namespace gi {
namespace {

constexpr int64_t kStaticCounts = 64;
constexpr int64_t kStaticQueue = GI_BFS_STATIC_MAX_VERTICES;
constexpr int64_t kStaticWords = kStaticCounts + 2 * kStaticQueue + 2 * (kStaticQueue / 64);
static_assert(kStaticQueue % 64 == 0, "whole bitmap words");
static_assert((kStaticCounts + 2 * kStaticQueue) % 2 == 0, "the bitmap starts 8-byte aligned");

}  // namespace
}  // namespace gi

static __device__ __align__(256) int32_t scratch[gi::kStaticWords];

namespace gi {
namespace {

__global__ void set_one_kernel_static(int32_t *q, int32_t *count, int32_t v) {
    q[0] = v;
    *count = 1;
}

// Device address of this GPU's `scratch` (the array's name is a host-side
// stand-in): looked up once per device, with that device current.
int static_scratch(int32_t **out) {
    static std::mutex mu;
    static std::map<int, int32_t *> by_device;
    int dev = 0;
    GI_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = by_device.find(dev);
    if (it == by_device.end()) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GI_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        it = by_device.emplace(dev, static_cast<int32_t *>(p)).first;
    }
    *out = it->second;
    return GI_OK;
}

}  // namespace
}  // namespace gi

extern "C" {

int gi_bfs_static(gi_graph *g, int32_t source, const gi_schedule *sched, int32_t *d_depth,
                  int32_t *d_parent, void *stream) {
    if (!g || !d_depth) return fail(GI_ERR_INVALID, "null argument");
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (g->V > GI_BFS_STATIC_MAX_VERTICES)
        return fail(GI_ERR_RANGE, "gi_bfs_static handles up to %d vertices (graph has %d)",
                    GI_BFS_STATIC_MAX_VERTICES, g->V);
    if (source < 0 || source >= g->V) return fail(GI_ERR_RANGE, "source %d out of range [0,%d)", source, g->V);
    gi_schedule s1;
    if (sched) {
        s1 = *sched;
    } else {  // the app's schedule, apps/bfs.cpp:86-88
        gi_schedule_init(&s1);
        s1.direction = GI_PUSH;
        s1.load_balancing = GI_LB_TWCE;
    }
    GI_TRY(validate_schedule(&s1));
    // push steps whose load balancing needs no per-graph memory: TWC and
    // STRICT bin into graph-owned scratch, EDGE_ONLY builds a COO array
    if (s1.direction != GI_PUSH || s1.kernel_fusion == GI_FUSION_ENABLED ||
        s1.frontier_creation != GI_FC_SPARSE ||
        (s1.load_balancing != GI_LB_VERTEX_BASED && s1.load_balancing != GI_LB_TWCE &&
         s1.load_balancing != GI_LB_CM && s1.load_balancing != GI_LB_WM))
        return fail(GI_ERR_UNSUPPORTED,
                    "gi_bfs_static runs unfused push steps with a sparse frontier and VERTEX_BASED, WM, "
                    "TWCE or CM load balancing");
    s1.deduplication = GI_DEDUP_DISABLED;  // the visited bitmap already admits each vertex once
    hipStream_t s = (hipStream_t)stream;
    int32_t *base = nullptr;
    GI_TRY(static_scratch(&base));
    // a frontier over the fixed array, built on the host stack
    gi_frontier f;
    f.V = g->V;
    f.cap = g->V;
    f.fmt = GI_FMT_SPARSE;
    f.side = 0;
    f.counts = base;
    f.q[0] = base + kStaticCounts;
    f.q[1] = base + kStaticCounts + kStaticQueue;
    f.words = bitmap_words(g->V);
    uint64_t *visited = reinterpret_cast<uint64_t *>(base + kStaticCounts + 2 * kStaticQueue);
    // the array keeps the previous call's contents: clear what starts at zero
    GI_TRY(fill_i32(base, kStaticCounts, 0, s));
    GI_TRY(fill_u64(visited, f.words, 0, s));
    GI_TRY(fill_i32(d_depth, g->V, -1, s));
    if (d_parent) GI_TRY(fill_i32(d_parent, g->V, -1, s));
    bfs_seed_kernel<<<1, 1, 0, s>>>(d_depth, d_parent, visited, source);
    set_one_kernel_static<<<1, 1, 0, s>>>(f.q[0], f.counts, source);
    GI_LAUNCH_CHECK();
    int32_t n = 1, level = 0;
    gi_op_args dummy{};
    while (n > 0) {
        OpBfsVisit op{visited, d_depth, d_parent, level + 1};
        GI_TRY(edgeset_apply(g, &s1, &f, &f, GI_OP_BFS_DEPTH, dummy, s, &op, nullptr));
        GI_TRY(read_count(f.count_in(), &n, s));
        ++level;
    }
    int32_t ovf = 0;
    GI_TRY(read_count(f.overflow(), &ovf, s));
    if (ovf) return fail(GI_ERR_RANGE, "frontier overflow");
    return GI_OK;
}

}  // extern "C"
