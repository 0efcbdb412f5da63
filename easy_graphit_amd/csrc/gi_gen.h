// Runtime of the HIP code the front end generates from staged GraphIt
// programs (frontend/, pipeline/graphit.h): the handle types of the
// program's variables and the calls its statements make. Generated
// translation units include this header, are compiled with hipcc for gfx950
// and link libgraphit_amd.so; every traversal goes through the library's
// dispatcher and load-balanced kernels (gi_lb.h) with the program's UDFs as
// operator functors. Failures throw gi_gen::failure; the generated entry
// point catches it and returns an error code.
#pragma once

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "gi_lb.h"

namespace gi {
int frontier_convert(gi_frontier *f, int32_t fmt, hipStream_t s);
}

namespace gi_gen {

struct failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what) {
    if (rc != GI_OK) throw failure(std::string(what) + ": " + gi_last_error());
}
inline void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw failure(std::string(what) + ": " + hipGetErrorString(e));
}

// GraphT (graph.h:17-88): the device CSR and its sizes, by value in kernels.
struct GraphT {
    gi_graph *h = nullptr;
    int num_vertices = 0;
    int num_edges = 0;
    const int *d_src_offsets = nullptr;
    const int *d_edge_dst = nullptr;
    const int *d_edge_weight = nullptr;
    int *out_degrees = nullptr;
};

__global__ inline void gen_degrees_kernel(const int *off, int V, int *deg) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        deg[v] = off[v + 1] - off[v];
}

inline GraphT wrap_graph(gi_graph *h) {
    GraphT g;
    g.h = h;
    int32_t V = 0;
    int64_t E = 0;
    check(gi_graph_num_vertices(h, &V), "graph size");
    check(gi_graph_num_edges(h, &E), "graph size");
    if (E > INT_MAX) throw failure("graph has more than 2^31 - 1 edges");
    g.num_vertices = V;
    g.num_edges = (int)E;
    const int32_t *off = nullptr, *nbr = nullptr, *w = nullptr;
    check(gi_graph_device_csr(h, &off, &nbr, &w), "graph arrays");
    g.d_src_offsets = off;
    g.d_edge_dst = nbr;
    g.d_edge_weight = w;
    void *deg = nullptr;
    check(gi_malloc(&deg, sizeof(int) * (size_t)(V > 0 ? V : 1)), "degrees");
    g.out_degrees = static_cast<int *>(deg);
    if (V > 0) {
        int grid = (V + 255) / 256;
        gen_degrees_kernel<<<grid < 4096 ? grid : 4096, 256>>>(off, V, g.out_degrees);
        check_hip(hipGetLastError(), "degrees kernel");
    }
    return g;
}

// load_graph (graph.h:236-338): a graph file; weights are read when the
// file has them (1 otherwise), as the reference's loader.
inline GraphT load_graph(const char *path) {
    gi_graph *h = nullptr;
    check(gi_graph_load(path, 0, &h), "load_graph");
    return wrap_graph(h);
}

inline GraphT transposed(const GraphT &g) {
    gi_graph *t = nullptr;
    check(gi_graph_transpose(g.h, &t), "transpose");
    return t == g.h ? g : wrap_graph(t);
}

inline GraphT blocked(const GraphT &g, int block_size) {
    gi_graph *b = nullptr;
    check(gi_graph_blocked(g.h, block_size, &b), "blocked graph");
    GraphT r = g;
    r.h = b;
    return r;
}

template <class T>
T *alloc(int n) {
    void *p = nullptr;
    check(gi_malloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)), "allocate");
    return static_cast<T *>(p);
}

// host access to one element of a device array
template <class T>
T load(const T *p, int i) {
    T v;
    check_hip(hipMemcpy(&v, p + i, sizeof(T), hipMemcpyDeviceToHost), "element read");
    return v;
}
template <class T, class U>
void store(T *p, int i, U v) {
    T x = static_cast<T>(v);
    check_hip(hipMemcpy(p + i, &x, sizeof(T), hipMemcpyHostToDevice), "element write");
}

// VertexSubset (vertex_frontier.h:26-166)
inline gi_frontier *new_vertex_subset(int num_vertices) {
    gi_frontier *f = nullptr;
    check(gi_frontier_create(num_vertices, 0, &f), "new_vertex_subset");
    return f;
}
inline void add_vertex(gi_frontier *f, int v) { check(gi_frontier_add_vertex(f, v, nullptr), "addVertex"); }
inline int frontier_size(gi_frontier *f) {
    int32_t n = 0;
    check(gi_frontier_size(f, &n, nullptr), "size");
    return n;
}

// VertexFrontierList (list.h:18-116)
inline gi_frontier_list *new_frontier_list(int num_vertices) {
    gi_frontier_list *l = nullptr;
    check(gi_frontier_list_create(num_vertices, &l), "new_frontier_list");
    return l;
}
inline void frontier_list_insert(gi_frontier_list *l, gi_frontier *f) {
    check(gi_frontier_list_insert(l, f, nullptr), "frontier list insert");
}
inline void frontier_list_retrieve(gi_frontier_list *l, gi_frontier *f) {
    check(gi_frontier_list_retrieve(l, f, nullptr), "frontier list retrieve");
}

// timers (gpu_intrinsics.h:66-80): seconds of wall time, device drained
inline std::chrono::steady_clock::time_point &timer_start() {
    static std::chrono::steady_clock::time_point t;
    return t;
}
inline void start_timer() {
    check_hip(hipDeviceSynchronize(), "timer");
    timer_start() = std::chrono::steady_clock::now();
}
inline float stop_timer() {
    check_hip(hipDeviceSynchronize(), "timer");
    return std::chrono::duration<float>(std::chrono::steady_clock::now() - timer_start()).count();
}
inline void print_time(float t) { printf("%f\n", t); }

// schedules: gi_schedule field for field (SimpleGPUSchedule)
inline gi_schedule schedule(int dir, int pull_rep, int creation, int dedup, int lb, int blocking, int block_size,
                            int fusion, int max_cta, int cta_size) {
    gi_schedule s;
    gi_schedule_init(&s);
    s.direction = dir;
    s.pull_frontier_rep = pull_rep;
    s.frontier_creation = creation;
    s.deduplication = dedup;
    s.load_balancing = lb;
    s.edge_blocking = blocking;
    s.block_size = block_size;
    s.kernel_fusion = fusion;
    s.max_cta = max_cta;
    s.cta_size = cta_size;
    return s;
}
struct Hybrid {
    gi_schedule s1, s2;
    float threshold;
};
inline Hybrid hybrid(const gi_schedule &a, const gi_schedule &b, float t) { return Hybrid{a, b, t}; }

// ---------------------------------------------------------------- UDF atomics
// writeMin (gpu_intrinsics.h:27-38): test, then atomicMin; true when lowered
__device__ inline bool write_min(int *p, int v) {
    if (*p <= v) return false;
    return atomicMin(p, v) > v;
}
__device__ inline bool write_min(float *p, float v) {
    float cur = *p;
    while (v < cur) {
        int old = atomicCAS(reinterpret_cast<int *>(p), __float_as_int(cur), __float_as_int(v));
        if (old == __float_as_int(cur)) return true;
        cur = __int_as_float(old);
    }
    return false;
}
template <class T, class U>
__device__ inline void write_sum(T *p, U v) {
    atomicAdd(p, static_cast<T>(v));
}
// compare-and-swap patch-up (graphit.cpp:95-110): true when *p was `old`
template <class U>
__device__ inline bool cas(int *p, int old, U v) {
    return atomicCAS(p, old, static_cast<int>(v)) == old;
}
template <class U>
__device__ inline bool cas(unsigned char *p, int old, U v) {
    // a byte inside its aligned 32-bit word
    unsigned int *w = reinterpret_cast<unsigned int *>(reinterpret_cast<uintptr_t>(p) & ~uintptr_t(3));
    const int sh = (int)((reinterpret_cast<uintptr_t>(p) & 3) * 8);
    unsigned int cur = *w;
    for (;;) {
        if (((cur >> sh) & 0xffu) != (unsigned int)(old & 0xff)) return false;
        const unsigned int nv = (cur & ~(0xffu << sh)) | ((unsigned int)(static_cast<int>(v) & 0xff) << sh);
        const unsigned int got = atomicCAS(w, cur, nv);
        if (got == cur) return true;
        cur = got;
    }
}

// ---------------------------------------------------------------- priority queue
// PrioQueue<int> (gpu_priority_queue.h:21-247) over the library's gi_pq; the
// current window's cutoff is kept in device memory for the UDFs.
struct PrioQueue {
    gi_pq *h = nullptr;
    int *prio = nullptr;
    int *d_cutoff = nullptr;
};
inline void pq_init(PrioQueue &q, const GraphT &g, int *prio, int initial, int delta, int source) {
    if (q.h) check(gi_pq_destroy(q.h), "priority queue");
    q.h = nullptr;
    check(gi_pq_create(g.h, prio, initial, delta, source, &q.h), "priority queue init");
    q.prio = prio;
    if (!q.d_cutoff) q.d_cutoff = alloc<int>(1);
}
inline int pq_finished(PrioQueue &q) {
    int32_t f = 0;
    check(gi_pq_finished(q.h, &f, nullptr), "finished");
    return f;
}
inline gi_frontier *pq_dequeue(PrioQueue &q) {
    gi_frontier *f = nullptr;
    check(gi_pq_dequeue(q.h, &f, nullptr), "dequeue");
    return f;
}
__device__ inline bool update_priority_min(int *prio, const int *cutoff, int v, int np) {
    return write_min(&prio[v], np) && np < *cutoff;
}

// ---------------------------------------------------------------- operators
template <class Op>
__global__ void gen_vertex_all_kernel(Op op, int V) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) op(v);
}
template <class Op>
__global__ void gen_vertex_list_kernel(Op op, const int *q, const int *count) {
    const int n = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) op(q[i]);
}

// vertexset_apply (operators.cpp:10-65) over every vertex
template <class Op>
void vertexset_apply(const GraphT &g, Op op) {
    if (g.num_vertices <= 0) return;
    const int grid = gi::persistent_grid(256);
    gen_vertex_all_kernel<Op><<<grid, 256>>>(op, g.num_vertices);
    check_hip(hipGetLastError(), "vertexset_apply");
}
// ... over the members of a frontier (its sparse form)
template <class Op>
void vertexset_apply(gi_frontier *f, Op op) {
    check(gi::frontier_convert(f, GI_FMT_SPARSE, nullptr), "vertexset_apply");
    const int grid = gi::persistent_grid(256);
    gen_vertex_list_kernel<Op><<<grid, 256>>>(op, f->q[f->side], f->count_in());
    check_hip(hipGetLastError(), "vertexset_apply");
}

// edgeset_apply (operators.cpp:513-650) with a generated operator functor;
// a sparse output is deduplicated (a UDF may add a vertex more than once)
template <class Op>
void edgeset_apply(const GraphT &g, const gi_schedule &s, gi_frontier *from, gi_frontier *to, Op op) {
    gi::ApplyPlan plan{};
    int32_t fmt = GI_FMT_SPARSE;
    check(gi::edgeset_plan(g.h, &s, from, to, true, nullptr, nullptr, &plan, &fmt), "edgeset_apply");
    check(gi::run_plan(op, plan), "edgeset_apply");
    check(gi::edgeset_finish(to, fmt, nullptr), "edgeset_apply");
}
// the hybrid rule (operators.cpp:516-532): |from| <= t * V runs s1, else s2
template <class Op>
void edgeset_apply(const GraphT &g, const Hybrid &h, gi_frontier *from, gi_frontier *to, Op op) {
    const gi_schedule *pick = &h.s1;
    if (from) {
        const int n = frontier_size(from);
        pick = (double)n <= (double)h.threshold * (double)g.num_vertices ? &h.s1 : &h.s2;
    }
    edgeset_apply(g, *pick, from, to, op);
}
// apply_priority (operators.h:73-103): the queue's frontier is the output
template <class Op>
void edgeset_apply_priority(const GraphT &g, const gi_schedule &s, gi_frontier *from, PrioQueue &q, Op op) {
    int32_t cutoff = 0;
    check(gi_pq_cutoff(q.h, &cutoff), "prio_cutoff");
    check_hip(hipMemcpy(q.d_cutoff, &cutoff, sizeof(int), hipMemcpyHostToDevice), "prio_cutoff");
    edgeset_apply(g, s, from, from, op);
}
template <class Op>
void edgeset_apply_priority(const GraphT &g, const Hybrid &h, gi_frontier *from, PrioQueue &q, Op op) {
    const int n = frontier_size(from);
    edgeset_apply_priority(g, (double)n <= (double)h.threshold * (double)g.num_vertices ? h.s1 : h.s2, from, q,
                           op);
}

}  // namespace gi_gen
