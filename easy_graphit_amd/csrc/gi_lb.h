// Load-balanced edge traversal templates: the seven load-balancing schedules
// of SimpleGPUSchedule (include/graphit/schedule.h:53-61) as wave64 HIP
// kernels, parameterised by an operator functor (the compiled form of an
// app's UDF) and by direction.
//
//   VERTEX_BASED  lane per enumerated vertex        (operators.cpp:72-129)
//   TWC           global degree bins: lane / wave / workgroup per vertex
//                 (the reference declares TWC with twc_*_bin scratch,
//                 graph.h:48-54, but asserts at run time, operators.cpp:645)
//   TWCE          per-workgroup edge split into workgroup-, wave- and
//                 8-lane-granule chunks through LDS queues (operators.cpp:201-357;
//                 the warp granule of 32 becomes the wave granule of 64)
//   WM            one wave per vertex (declared only in the reference)
//   CM            workgroup degree scan + merge-path search (operators.cpp:359-492;
//                 here the edge weight is passed, where the reference passes 0)
//   STRICT        grid-wide degree prefix sum, equal edge ranges per workgroup
//                 (declared only; strict_* scratch, graph.h:56-59)
//   EDGE_ONLY     grid-stride over the COO edge list, optionally blocked by
//                 destination (operators.cpp:131-200)
//
// Direction (operators.cpp:550-620): PUSH enumerates sources, filters
// destinations with the op's filter and applies op(src, dst). PULL runs on
// the transposed graph, enumerates the vertices passing the op's filter,
// re-checks that filter before every edge (the early break of
// operators.cpp:85-87), keeps neighbours that are members of the input
// frontier, and applies op(neighbour, vertex).
#pragma once

#include "gi_internal.h"

namespace gi {

// ============================================================ operators
// Each op: filter(v) = the `to` filter; apply(src, dst, w) performs the
// tracked write and returns true when dst must enter the output frontier.

struct OpBfsParent {  // apps/bfs.cpp:23-29 + CAS patch-up (graphit.cpp:55-122)
    int32_t *parent;
    __device__ bool filter(int32_t v) const { return parent[v] == -1; }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        return atomicCAS(&parent[d], -1, s) == -1;
    }
};

struct OpBfsDepth {  // depth-recording BFS: first claim wins, value is level+1
    int32_t *depth;
    int32_t next;
    __device__ bool filter(int32_t v) const { return depth[v] == -1; }
    __device__ bool apply(int32_t, int32_t d, int32_t) const {
        return atomicCAS(&depth[d], -1, next) == -1;
    }
};

// Internal BFS op of the fused driver: a 64-bit-word visited bitmap carries
// the filter (2 MiB at 16.8M vertices stays in L2), depth and parent are
// plain stores after a successful claim.
struct OpBfsVisit {
    uint64_t *visited;
    int32_t *depth;
    int32_t *parent;  // optional
    int32_t next;
    __device__ bool filter(int32_t v) const { return !bit_test(visited, v); }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        if (!bit_claim(visited, d)) return false;
        depth[d] = next;
        if (parent) parent[d] = s;
        return true;
    }
};

struct OpSsspRelax {  // apps/sssp.cpp:24-27 + updatePriorityMin (graphit_types.h:426-430)
    int32_t *dist;
    int32_t cutoff;
    __device__ bool filter(int32_t) const { return true; }
    __device__ bool apply(int32_t s, int32_t d, int32_t w) const {
        int32_t ds = dist[s];
        if (ds == INT32_MAX) return false;
        int32_t nd = ds + w;
        if (nd >= dist[d]) return false;  // writeMin's test before the atomic
        int32_t old = atomicMin(&dist[d], nd);
        return nd < old && nd < cutoff;
    }
};

struct OpCcHook {  // apps/cc.cpp:23-38
    int32_t *ID;
    int32_t *updates;
    __device__ bool filter(int32_t) const { return true; }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        int32_t src_id = ID[s], dst_id = ID[d];
        int32_t p_src = ID[src_id], p_dst = ID[dst_id];
        if (p_dst < ID[src_id]) atomicMin(&ID[src_id], p_dst);
        if (p_src < ID[dst_id]) atomicMin(&ID[dst_id], p_src);
        if (p_dst != ID[dst_id] || p_src != ID[src_id])
            if (updates[0] == 0) updates[0] = 1;
        return false;
    }
};

struct OpPrAdd {  // apps/pr.cpp:32-34 (+= on a SHARED vertex is an atomic add)
    float *new_rank;
    const float *contrib;
    __device__ bool filter(int32_t) const { return true; }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        atomicAdd(&new_rank[d], contrib[s]);
        return false;
    }
};

struct OpBcForward {  // apps/bc.cpp:34-40
    double *paths;
    const uint8_t *visited;
    __device__ bool filter(int32_t v) const { return visited[v] == 0; }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        atomicAdd(&paths[d], paths[s]);
        return true;
    }
};

struct OpBcBackward {  // apps/bc.cpp:54-56
    float *dep;
    const uint8_t *visited;
    __device__ bool filter(int32_t v) const { return visited[v] == 0; }
    __device__ bool apply(int32_t s, int32_t d, int32_t) const {
        atomicAdd(&dep[d], dep[s]);
        return false;
    }
};

// ============================================================ context
template <class Op, bool PULL>
struct Ctx {
    DGraph g;
    Op op;
    Member in;  // PULL: membership of the input frontier
    Sink out;

    __device__ __forceinline__ bool src_ok(int32_t s) const {
        if constexpr (PULL) return op.filter(s);
        else return true;
    }
    // Every active lane of the wave calls this once per step; `active` false
    // lanes only take part in the output aggregation.
    __device__ __forceinline__ void edge(int32_t s, int32_t e, bool active) const {
        bool emit = false;
        int32_t id = 0;
        if (active) {
            int32_t d = g.nbr[e];
            int32_t w = g.w ? g.w[e] : 1;
            if constexpr (PULL) {
                if (member(in, d)) {
                    emit = op.apply(d, s, w);
                    id = s;
                }
            } else {
                if (op.filter(d)) {
                    emit = op.apply(s, d, w);
                    id = d;
                }
            }
        }
        sink_emit(out, emit, id);
    }
};

// ============================================================ VERTEX_BASED
template <class Op, bool PULL>
__global__ void lb_vertex_kernel(Ctx<Op, PULL> c, VEnum en) {
    int32_t n = en.size();
    int32_t stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int32_t s = en.at(i);
        int32_t e1 = c.g.off[s + 1];
        for (int32_t e = c.g.off[s]; e < e1; ++e) {
            if (!c.src_ok(s)) break;
            c.edge(s, e, true);
        }
    }
}

// ============================================================ WM
template <class Op, bool PULL>
__global__ void lb_wave_kernel(Ctx<Op, PULL> c, VEnum en) {
    int32_t n = en.size();
    int32_t lane = lane_id();
    int32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int32_t waves = (gridDim.x * blockDim.x) >> 6;
    for (int32_t i = wave; i < n; i += waves) {
        int32_t s = en.at(i);
        int32_t e0 = c.g.off[s], e1 = c.g.off[s + 1];
        for (int32_t base = e0; base < e1; base += 64) {
            if (!c.src_ok(s)) break;  // wave-uniform after the filter settles
            int32_t e = base + lane;
            c.edge(s, e, e < e1);
        }
    }
}

// ============================================================ TWC
// Phase 1: classify enumerated vertices into three global bins by degree.
// Phase 2 (one kernel): small bin lane-per-vertex, mid bin wave-per-vertex,
// large bin workgroup-per-vertex.
static __global__ void twc_bin_kernel(VEnum en, const int32_t *off, int32_t *bins, int32_t cap,
                               int32_t *counts, int32_t large_min) {
    int32_t n = en.size();
    int64_t limit = ceil_div(n, 64) * 64;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < limit;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t s = 0, deg = 0;
        bool ok = i < n;
        if (ok) {
            s = en.at((int32_t)i);
            deg = off[s + 1] - off[s];
        }
        bool sm = ok && deg < 64, md = ok && deg >= 64 && deg < large_min, lg = ok && deg >= large_min;
        int32_t p0 = wave_atomic_add(&counts[0], sm);
        int32_t p1 = wave_atomic_add(&counts[1], md);
        int32_t p2 = wave_atomic_add(&counts[2], lg);
        if (sm) bins[p0] = s;
        if (md) bins[cap + p1] = s;
        if (lg) bins[2 * cap + p2] = s;
    }
}

template <class Op, bool PULL>
__global__ void lb_twc_kernel(Ctx<Op, PULL> c, const int32_t *bins, int32_t cap,
                              const int32_t *counts) {
    int32_t n0 = counts[0], n1 = counts[1], n2 = counts[2];
    int32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t threads = gridDim.x * blockDim.x;
    // small: lane per vertex
    for (int32_t i = tid; i < n0; i += threads) {
        int32_t s = bins[i];
        int32_t e1 = c.g.off[s + 1];
        for (int32_t e = c.g.off[s]; e < e1; ++e) {
            if (!c.src_ok(s)) break;
            c.edge(s, e, true);
        }
    }
    // mid: wave per vertex
    int32_t lane = lane_id();
    for (int32_t i = tid >> 6; i < n1; i += threads >> 6) {
        int32_t s = bins[cap + i];
        int32_t e0 = c.g.off[s], e1 = c.g.off[s + 1];
        for (int32_t base = e0; base < e1; base += 64) {
            if (!c.src_ok(s)) break;
            c.edge(s, base + lane, base + lane < e1);
        }
    }
    // large: workgroup per vertex
    for (int32_t i = blockIdx.x; i < n2; i += gridDim.x) {
        int32_t s = bins[2 * cap + i];
        int32_t e0 = c.g.off[s], e1 = c.g.off[s + 1];
        for (int32_t base = e0; base < e1; base += blockDim.x) {
            if (!c.src_ok(s)) break;
            int32_t e = base + (int32_t)threadIdx.x;
            c.edge(s, e, e < e1);
        }
    }
}

// ============================================================ TWCE
constexpr int kTwceGroup = 8;   // lanes per vertex in stage 1 (STAGE_1_SIZE)
constexpr int kMaxBlock = 1024;

template <class Op, bool PULL>
__global__ void lb_twce_kernel(Ctx<Op, PULL> c, VEnum en) {
    __shared__ int32_t wq_v[kMaxBlock / kTwceGroup], wq_off[kMaxBlock / kTwceGroup],
        wq_len[kMaxBlock / kTwceGroup];
    __shared__ int32_t bq_v[kMaxBlock / kTwceGroup], bq_off[kMaxBlock / kTwceGroup],
        bq_len[kMaxBlock / kTwceGroup];
    __shared__ int32_t cnt[2];
    const int32_t B = blockDim.x;
    const int32_t per_block = B / kTwceGroup;
    int32_t n = en.size();
    int32_t lane = lane_id();
    int32_t g8 = threadIdx.x % kTwceGroup;
    for (int32_t base = blockIdx.x * per_block; base < n; base += gridDim.x * per_block) {
        if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
        __syncthreads();
        int32_t idx = base + (int32_t)threadIdx.x / kTwceGroup;
        bool has = idx < n;
        int32_t s = 0, small0 = 0, small1 = 0;
        if (has) {
            s = en.at(idx);
            int32_t o0 = c.g.off[s], o1 = c.g.off[s + 1];
            int32_t deg = o1 - o0;
            int32_t big = deg / B * B;
            int32_t rem = deg - big;
            int32_t mid = rem / 64 * 64;
            if (g8 == 0 && big > 0) {
                int32_t p = atomicAdd(&cnt[1], 1);
                bq_v[p] = s; bq_off[p] = o0; bq_len[p] = big;
            }
            if (g8 == 0 && mid > 0) {
                int32_t p = atomicAdd(&cnt[0], 1);
                wq_v[p] = s; wq_off[p] = o0 + big; wq_len[p] = mid;
            }
            small0 = o0 + big + mid;
            small1 = o1;
        }
        __syncthreads();
        // stage 1: 8 lanes per vertex over the remainder (< 64 edges)
        if (has) {
            for (int32_t e = small0 + g8; e < small1; e += kTwceGroup) {
                if (!c.src_ok(s)) break;
                c.edge(s, e, true);
            }
        }
        // stage 2: waves pop wave-granule chunks
        int32_t nbq = cnt[1];
        for (;;) {
            int32_t item = 0;
            if (lane == 0) item = atomicSub(&cnt[0], 1) - 1;
            item = __shfl(item, 0);
            if (item < 0) break;
            int32_t v = wq_v[item], o = wq_off[item], len = wq_len[item];
            for (int32_t k = 0; k < len; k += 64) {
                if (!c.src_ok(v)) break;
                c.edge(v, o + k + lane, true);
            }
        }
        // stage 3: the whole workgroup walks each workgroup-granule chunk
        for (int32_t j = 0; j < nbq; ++j) {
            int32_t v = bq_v[j], o = bq_off[j], len = bq_len[j];
            for (int32_t k = 0; k < len; k += B) {
                if (!c.src_ok(v)) break;
                c.edge(v, o + k + (int32_t)threadIdx.x, true);
            }
        }
        __syncthreads();
    }
}

// ============================================================ block scan
// Exclusive scan of one int per thread over the workgroup; *total gets the sum.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t x, int32_t *total) {
    __shared__ int32_t wave_sums[kMaxBlock / 64 + 1];
    int32_t lane = lane_id();
    int32_t wid = threadIdx.x >> 6;
    int32_t nw = blockDim.x >> 6;
    int32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        int32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wave_sums[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int32_t ws = lane < nw ? wave_sums[lane] : 0;
        int32_t wi = ws;
        for (int o = 1; o < 64; o <<= 1) {
            int32_t y = __shfl_up(wi, o);
            if (lane >= o) wi += y;
        }
        if (lane < nw) wave_sums[lane] = wi - ws;
        if (lane == 63) wave_sums[kMaxBlock / 64] = wi;
    }
    __syncthreads();
    int32_t r = incl - x + wave_sums[wid];
    *total = wave_sums[kMaxBlock / 64];
    __syncthreads();
    return r;
}

// ============================================================ CM
template <class Op, bool PULL>
__global__ void lb_cm_kernel(Ctx<Op, PULL> c, VEnum en) {
    __shared__ int32_t sv[kMaxBlock], soff[kMaxBlock], sscan[kMaxBlock + 1];
    const int32_t B = blockDim.x;
    int32_t n = en.size();
    for (int32_t base = blockIdx.x * B; base < n; base += gridDim.x * B) {
        int32_t i = base + (int32_t)threadIdx.x;
        int32_t deg = 0;
        if (i < n) {
            int32_t s = en.at(i);
            int32_t o0 = c.g.off[s];
            deg = c.g.off[s + 1] - o0;
            sv[threadIdx.x] = s;
            soff[threadIdx.x] = o0;
        }
        int32_t total = 0;
        int32_t ex = block_exclusive_scan(deg, &total);
        sscan[threadIdx.x] = ex;
        int32_t width = n - base < B ? n - base : B;
        __syncthreads();
        for (int32_t k = threadIdx.x; k < total; k += B) {
            int32_t j = upper_bound_i32(sscan, 0, width, k) - 1;
            int32_t s = sv[j];
            bool ok = c.src_ok(s);
            c.edge(s, soff[j] + (k - sscan[j]), ok);
        }
        __syncthreads();
    }
}

// ============================================================ STRICT
// S1: per-chunk degree sums; S2: scan of chunk sums (one workgroup);
// S3: per-element exclusive prefix; E: each workgroup takes an equal range
// of the grid's total edge count and locates owners by binary search.
static __global__ void strict_chunk_sum_kernel(VEnum en, const int32_t *off, int32_t *chunk_sum,
                                        int32_t chunks) {
    int32_t n = en.size();
    int32_t per = (int32_t)ceil_div(n, chunks);
    for (int32_t b = blockIdx.x; b < chunks; b += gridDim.x) {
        int32_t lo = b * per, hi = lo + per < n ? lo + per : n;
        int32_t acc = 0;
        for (int32_t i = lo + (int32_t)threadIdx.x; i < hi; i += blockDim.x) {
            int32_t s = en.at(i);
            acc += off[s + 1] - off[s];
        }
        int32_t total = 0;
        block_exclusive_scan(acc, &total);
        if (threadIdx.x == 0) chunk_sum[b] = total;
    }
}

static __global__ void strict_chunk_scan_kernel(int32_t *chunk_sum, int32_t chunks, int32_t *grand) {
    int32_t carry = 0;
    for (int32_t base = 0; base < chunks; base += blockDim.x) {
        int32_t i = base + (int32_t)threadIdx.x;
        int32_t x = i < chunks ? chunk_sum[i] : 0;
        int32_t total = 0;
        int32_t ex = block_exclusive_scan(x, &total);
        if (i < chunks) chunk_sum[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

static __global__ void strict_prefix_kernel(VEnum en, const int32_t *off, const int32_t *chunk_sum,
                                     int32_t chunks, int32_t *prefix) {
    int32_t n = en.size();
    int32_t per = (int32_t)ceil_div(n, chunks);
    for (int32_t b = blockIdx.x; b < chunks; b += gridDim.x) {
        int32_t lo = b * per, hi = lo + per < n ? lo + per : n;
        int32_t carry = chunk_sum[b];
        for (int32_t base = lo; base < hi; base += blockDim.x) {
            int32_t i = base + (int32_t)threadIdx.x;
            int32_t d = 0;
            if (i < hi) {
                int32_t s = en.at(i);
                d = off[s + 1] - off[s];
            }
            int32_t total = 0;
            int32_t ex = block_exclusive_scan(d, &total);
            if (i < hi) prefix[i] = carry + ex;
            carry += total;
        }
    }
}

template <class Op, bool PULL>
__global__ void lb_strict_kernel(Ctx<Op, PULL> c, VEnum en, const int32_t *prefix,
                                 const int32_t *grand) {
    int32_t n = en.size();
    int32_t total = *grand;
    if (n == 0 || total == 0) return;
    int32_t per = (int32_t)ceil_div(total, gridDim.x);
    int32_t lo = blockIdx.x * per;
    if (lo >= total) return;
    int32_t hi = lo + per < total ? lo + per : total;
    // owners of lo and hi-1 bound every search of this workgroup
    int32_t jlo = upper_bound_i32(prefix, 0, n, lo) - 1;
    int32_t jhi = upper_bound_i32(prefix, jlo, n, hi - 1);
    for (int32_t k = lo + (int32_t)threadIdx.x; k - (int32_t)threadIdx.x < hi; k += blockDim.x) {
        bool act = k < hi;
        int32_t s = 0, e = 0;
        if (act) {
            int32_t j = upper_bound_i32(prefix, jlo, jhi, k) - 1;
            s = en.at(j);
            e = c.g.off[s] + (k - prefix[j]);
            act = c.src_ok(s);
        }
        c.edge(s, e, act);
    }
}

// ============================================================ EDGE_ONLY
// Grid-stride over every edge entry (COO). Push from a frontier keeps edges
// whose source is a member (the reference's edge-only path ignores the
// frontier, operators.cpp:176-200; membership makes it usable as a
// frontier schedule). Pull keeps edges whose (transposed) source passes the
// op filter and whose neighbour is a member.
template <class Op, bool PULL>
__global__ void lb_edge_kernel(Ctx<Op, PULL> c, Member src_member) {
    int64_t E = c.g.E;
    int64_t limit = ceil_div(E, 64) * 64;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < limit;
         e += (int64_t)gridDim.x * blockDim.x) {
        bool act = e < E;
        int32_t s = 0;
        if (act) {
            s = c.g.esrc[e];
            if constexpr (PULL) act = c.op.filter(s);
            else act = member(src_member, s);
        }
        c.edge(s, (int32_t)e, act);
    }
}

// ============================================================ host launcher
struct LaunchParams {
    hipStream_t stream;
    int block;        // workgroup size
    int grid;         // persistent grid
    int32_t lb;       // gi_load_balance
    gi_graph *graph;  // traversed graph (for scratch)
    Member src_member;
    hipEvent_t ev0;   // optional: recorded around the traversal kernel only
    hipEvent_t ev1;
};

template <class Op, bool PULL>
int launch_lb(const Ctx<Op, PULL> &c, VEnum en, const LaunchParams &p) {
    hipStream_t s = p.stream;
    int32_t *bins = nullptr, *counts = nullptr, *prefix = nullptr, *grand = nullptr;
    int32_t cap = p.graph->V;
    // preparation kernels (binning, scans)
    if (p.lb == GI_LB_TWC) {
        GI_TRY(ensure_scratch(p.graph, scratch_lb_len(cap)));
        bins = p.graph->d_scratch;
        counts = bins + 3 * (int64_t)cap;
        GI_TRY(fill_i32(counts, 3, 0, s));
        twc_bin_kernel<<<p.grid, 256, 0, s>>>(en, c.g.off, bins, cap, counts, p.block);
        GI_LAUNCH_CHECK();
    } else if (p.lb == GI_LB_STRICT) {
        const int32_t chunks = 1024;
        GI_TRY(ensure_scratch(p.graph, scratch_lb_len(cap)));
        prefix = p.graph->d_scratch;
        int32_t *chunk_sum = prefix + cap + 64;
        grand = chunk_sum + chunks;
        strict_chunk_sum_kernel<<<chunks, 256, 0, s>>>(en, c.g.off, chunk_sum, chunks);
        strict_chunk_scan_kernel<<<1, 1024, 0, s>>>(chunk_sum, chunks, grand);
        strict_prefix_kernel<<<chunks, 256, 0, s>>>(en, c.g.off, chunk_sum, chunks, prefix);
        GI_LAUNCH_CHECK();
    }
    if (p.ev0) GI_HIP(hipEventRecord(p.ev0, s));
    switch (p.lb) {
        case GI_LB_VERTEX_BASED:
            lb_vertex_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, en);
            break;
        case GI_LB_WM:
            lb_wave_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, en);
            break;
        case GI_LB_TWC:
            lb_twc_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, bins, cap, counts);
            break;
        case GI_LB_TWCE:
            lb_twce_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, en);
            break;
        case GI_LB_CM:
            lb_cm_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, en);
            break;
        case GI_LB_STRICT:
            lb_strict_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, en, prefix, grand);
            break;
        case GI_LB_EDGE_ONLY:
            lb_edge_kernel<Op, PULL><<<p.grid, p.block, 0, s>>>(c, p.src_member);
            break;
        default:
            return fail(GI_ERR_UNSUPPORTED, "unknown load balancing %d", p.lb);
    }
    GI_LAUNCH_CHECK();
    if (p.ev1) GI_HIP(hipEventRecord(p.ev1, s));
    return GI_OK;
}

// Dispatch on direction for one op type.
template <class Op>
int launch_op(const Op &op, bool pull, const DGraph &g, const Member &in, const Sink &out,
              VEnum en, const LaunchParams &p) {
    if (pull) {
        Ctx<Op, true> c{g, op, in, out};
        return launch_lb(c, en, p);
    }
    Ctx<Op, false> c{g, op, in, out};
    return launch_lb(c, en, p);
}

}  // namespace gi

namespace gi {

// Enumerator for PULL (create_reverse_frontier, operators.cpp:493-512):
// compact the vertices passing the op filter into q.
template <class Op>
__global__ void reverse_frontier_kernel(Op op, int32_t V, int32_t *q, int32_t *count) {
    int64_t limit = ceil_div(V, 64) * 64;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < limit;
         v += (int64_t)gridDim.x * blockDim.x) {
        bool ok = v < V && op.filter((int32_t)v);
        int32_t pos = wave_atomic_add(count, ok);
        if (ok) q[pos] = (int32_t)v;
    }
}

// Arguments of one edgeset apply after the host dispatcher resolved the
// graph, enumerator and output.
struct ApplyPlan {
    bool pull;
    DGraph g;
    Member in;
    Sink out;
    VEnum en;
    LaunchParams lp;
    bool reverse;     // build the pull enumerator from the op filter first
    int32_t *rev_q;   // scratch for it
    int32_t *rev_count;
};

template <class Op>
int run_plan(const Op &op, ApplyPlan plan) {
    if (plan.reverse) {
        GI_TRY(fill_i32(plan.rev_count, 1, 0, plan.lp.stream));
        int grid = plan.lp.grid;
        reverse_frontier_kernel<Op><<<grid, 256, 0, plan.lp.stream>>>(op, plan.g.V, plan.rev_q, plan.rev_count);
        GI_LAUNCH_CHECK();
        plan.en.q = plan.rev_q;
        plan.en.count = plan.rev_count;
    }
    return launch_op(op, plan.pull, plan.g, plan.in, plan.out, plan.en, plan.lp);
}

// Implemented per operator family in gi_ops_*.hip.
int run_bfs_family(int32_t op, const gi_op_args &a, const ApplyPlan &plan);
int run_other_family(int32_t op, const gi_op_args &a, const ApplyPlan &plan);
int run_bfs_visit(const OpBfsVisit &op, const ApplyPlan &plan);

// The dispatcher's two halves (gi_ops.hip), for callers that bring their own
// operator functor: resolve graph view, enumerator, input representation and
// output sink (may_duplicate: the op can emit a vertex twice, so a sparse
// output is deduplicated), then, after run_plan(op, plan), swap the output.
int edgeset_plan(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to, bool may_duplicate,
                 hipStream_t st, hipEvent_t *events, ApplyPlan *out, int32_t *out_fmt);
int edgeset_finish(gi_frontier *to, int32_t fmt, hipStream_t st);

}  // namespace gi
