/*
 * graphit_amd.h — C ABI of the MI355X-native GraphIt GPU runtime.
 *
 * This is the drop-in boundary for the reference's generated-code → runtime
 * interface (namespace graphit_runtime, header-only CUDA in runtime/) and for
 * the edgeset/vertexset operators that the staged DSL lowers into kernels
 * (src/graphit/operators.cpp). Every entry point names the reference
 * interface it replaces (reference-relative file:line).
 *
 * Conventions
 *   - Return value: GI_OK (0) or a negative GI_ERR_* code; gi_last_error()
 *     gives a message. Nothing calls exit() (the reference calls exit(-1) in
 *     cudaCheckLastError, runtime/infra_gpu/support.h:4-10).
 *   - Handles own their device memory and are released by gi_*_destroy.
 *   - Host arrays passed in are copied. Vertex data (depth, parent, dist,
 *     labels, ranks, scores) are caller-owned DEVICE pointers of the stated
 *     type and length num_vertices.
 *   - `stream` is a hipStream_t (NULL = default stream). Work is queued
 *     asynchronously on it; calls documented as "synchronises" wait for it.
 *   - A handle is used by one host thread at a time; calls on different
 *     graphs may run on different threads at once. The selector model is
 *     process-wide: gi_selector_load/unload must not overlap other calls.
 *   - int32 vertex ids and int32 CSR offsets, as the reference
 *     (runtime/infra_gpu/graph.h:20-35); graphs with >= 2^31 edge entries are
 *     rejected with GI_ERR_RANGE.
 */
#ifndef GRAPHIT_AMD_H
#define GRAPHIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GI_OK 0
#define GI_ERR_INVALID (-1)
#define GI_ERR_HIP (-2)
#define GI_ERR_OOM (-3)
#define GI_ERR_UNSUPPORTED (-4)
#define GI_ERR_RANGE (-5)
#define GI_ERR_IO (-6)
#define GI_ERR_UNDERFLOW (-7)

/* Message for the last failing call on this thread. */
const char *gi_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int32_t gi_version(void);

/* ------------------------------------------------------ memory and streams
 * The reference's cudaMalloc<T> / cudaMemcpyHostToDevice<T> /
 * cudaMemcpyDeviceToHost<T> wrappers (runtime/gpu_intrinsics.h:13-25) and the
 * default stream they use; callers without a HIP runtime of their own (the
 * Python harness) allocate vertex data through these. */
int gi_malloc(void **ptr, uint64_t bytes);
int gi_free(void *ptr);
int gi_memcpy_h2d(void *dst, const void *src, uint64_t bytes, void *stream);
/* Synchronises. */
int gi_memcpy_d2h(void *dst, const void *src, uint64_t bytes, void *stream);
int gi_memcpy_d2d(void *dst, const void *src, uint64_t bytes, void *stream);
/* Fill count int32 / float / uint8 elements with a value (kernel, not a memset
 * node, so it is safe under stream capture). */
int gi_fill_i32(int32_t *ptr, int64_t count, int32_t value, void *stream);
int gi_fill_f32(float *ptr, int64_t count, float value, void *stream);
int gi_fill_u8(uint8_t *ptr, int64_t count, uint8_t value, void *stream);
int gi_stream_create(void **stream);
int gi_stream_destroy(void *stream);
int gi_stream_synchronize(void *stream);
/* Number of visible devices (0 without a GPU; never fails). */
int32_t gi_device_count(void);
/* Device timer (start_timer / stop_timer, gpu_intrinsics.h:66-80, with
 * hipEvents on a stream instead of gettimeofday). */
int gi_timer_start(void *stream);
int gi_timer_stop(void *stream, float *ms);

/* ------------------------------------------------------------------ schedule
 * Mirrors SimpleGPUSchedule (include/graphit/schedule.h:26-118); enum values
 * keep the reference's declaration order. */
enum gi_direction { GI_PUSH = 0, GI_PULL = 1 };
enum gi_pull_rep { GI_PULL_BITMAP = 0, GI_PULL_BOOLMAP = 1 };
enum gi_frontier_creation { GI_FC_SPARSE = 0, GI_FC_BITMAP = 1, GI_FC_BOOLMAP = 2 };
enum gi_dedup { GI_DEDUP_DISABLED = 0, GI_DEDUP_ENABLED = 1 };
enum gi_load_balance {
    GI_LB_VERTEX_BASED = 0, GI_LB_TWC = 1, GI_LB_TWCE = 2, GI_LB_WM = 3,
    GI_LB_CM = 4, GI_LB_STRICT = 5, GI_LB_EDGE_ONLY = 6
};
enum gi_edge_blocking { GI_BLOCKED = 0, GI_UNBLOCKED = 1 };
enum gi_kernel_fusion { GI_FUSION_DISABLED = 0, GI_FUSION_ENABLED = 1 };

typedef struct gi_schedule {
    int32_t direction;          /* gi_direction */
    int32_t pull_frontier_rep;  /* gi_pull_rep (default BOOLMAP, schedule.h:92) */
    int32_t frontier_creation;  /* gi_frontier_creation */
    int32_t deduplication;      /* gi_dedup */
    int32_t load_balancing;     /* gi_load_balance */
    int32_t edge_blocking;      /* gi_edge_blocking */
    int32_t block_size;         /* vertices per destination block (EDGE_ONLY+BLOCKED) */
    int32_t kernel_fusion;      /* gi_kernel_fusion */
    int32_t max_cta;            /* workgroups for grid-stride kernels (0 = auto) */
    int32_t cta_size;           /* workgroup size: 64..1024, multiple of 64 */
} gi_schedule;

/* SimpleGPUSchedule() defaults (schedule.h:87-97, schedule.cpp:5-6). */
void gi_schedule_init(gi_schedule *s);

/* HybridGPUSchedule (schedule.h:122-139) plus the selector that decides
 * between s1 and s2 each iteration. GI_SELECT_HEURISTIC is GraphIt's rule
 * (|frontier| <= threshold * V runs s1, operators.cpp:526 / operators.h:116-132).
 * GI_SELECT_LEARNED runs the on-device MLP loaded with gi_selector_load; the
 * heuristic is its fallback when no model is loaded. */
enum gi_selector_mode { GI_SELECT_HEURISTIC = 0, GI_SELECT_LEARNED = 1,
                        GI_SELECT_S1 = 2, GI_SELECT_S2 = 3 };
typedef struct gi_hybrid_schedule {
    gi_schedule s1;
    gi_schedule s2;
    float threshold;   /* bindThreshold / configThreshold value (schedule.cpp:9-12) */
    int32_t selector;  /* gi_selector_mode */
} gi_hybrid_schedule;

/* --------------------------------------------------------------------- graph
 * Replaces GraphT<int> and load_graph (runtime/infra_gpu/graph.h:17-88,
 * 236-338). CSR: offsets[V+1], neighbours[E] sorted per vertex, optional
 * int32 weights[E]. */
typedef struct gi_graph gi_graph;

/* Copy a host CSR (already squished) to the device. */
int gi_graph_from_csr(int32_t num_vertices, int64_t num_edges, const int32_t *offsets,
                      const int32_t *nbrs, const int32_t *weights, int32_t directed,
                      gi_graph **out);
/* Build a CSR on the device from a host edge list with GAPBS builder
 * semantics: optional symmetrize, sort, dedup, self-loop removal
 * (runtime/infra_gapbs/builder.h:134-246). */
int gi_graph_from_edges(int32_t num_vertices, int64_t num_pairs, const int32_t *src,
                        const int32_t *dst, int32_t symmetrize, gi_graph **out);
/* Seeded generators run on the device, then the same CSR build.
 * R-MAT (a,b,c)=(0.57,0.19,0.19) with vertex-id permutation
 * (runtime/infra_gapbs/generator.h:74-107); uniform random
 * (generator.h:57-72); width x width 4-neighbour grid plus num_random
 * uniform long-range pairs. Spec of the hash is in the oracle sources. */
int gi_graph_rmat(int32_t scale, int32_t edge_factor, uint64_t seed, int32_t symmetrize,
                  gi_graph **out);
int gi_graph_uniform(int32_t num_vertices, int64_t num_pairs, uint64_t seed,
                     int32_t symmetrize, gi_graph **out);
int gi_graph_grid(int32_t width, int64_t num_random, uint64_t seed, gi_graph **out);
/* Weights uniform in [wmin, wmax] as a function of the unordered endpoint
 * pair (the role of Generator::InsertWeights, generator.h:124-139). */
int gi_graph_set_weights(gi_graph *g, uint64_t seed, int32_t wmin, int32_t wmax);
int gi_graph_num_vertices(const gi_graph *g, int32_t *out);
int gi_graph_num_edges(const gi_graph *g, int64_t *out);
/* Copy the CSR back to host arrays (weights may be NULL). Synchronises. */
int gi_graph_read_csr(const gi_graph *g, int32_t *offsets, int32_t *nbrs, int32_t *weights);
/* Device pointers of the CSR (borrowed; valid until destroy). */
int gi_graph_device_csr(const gi_graph *g, const int32_t **offsets, const int32_t **nbrs,
                        const int32_t **weights);
/* get_transposed_graph (graph.h:72-75, builtin_transpose graph.h:165-234),
 * built on the device and cached; the returned handle is owned by g. */
int gi_graph_transpose(gi_graph *g, gi_graph **out);
/* get_blocked_graph (graph.h:77-87, block_graph_edges graph.h:105-163):
 * edges grouped into buckets by destination block, cached per block size;
 * owned by g. Unlike the reference, g's own arrays stay valid. */
int gi_graph_blocked(gi_graph *g, int32_t block_size, gi_graph **out);
int gi_graph_destroy(gi_graph *g);

/* ------------------------------------------------------------- graph files
 * Replaces load_graph (runtime/infra_gpu/graph.h:236-303) and the GAPBS
 * reader and builder it calls (runtime/infra_gapbs/reader.h:51-375,
 * builder.h:62-246). The format comes from the suffix: .el ("u v", weight
 * 1), .wel ("u v w"), .gr (DIMACS "a u v w" lines, ids as written), .graph
 * (METIS, ids from 1, fmt 1 = weights), .mtx (MatrixMarket coordinate, ids
 * from 1, symmetric matrices both ways), .sg / .wsg (GAPBS serialized CSR,
 * taken as stored), .graphit_bin (the reference's cache). Text formats are
 * built as GAPBS builds them: V = largest id + 1, lists sorted, a repeated
 * neighbour kept once with its least weight, self-loops dropped; directed
 * unless GI_LOAD_SYMMETRIZE (load_graph never symmetrizes). Host code only:
 * no device is touched by gi_read_graph_file. */
enum gi_load_flags { GI_LOAD_SYMMETRIZE = 1, GI_LOAD_CACHE = 2 };
typedef struct gi_host_csr {
    int32_t num_vertices;
    int64_t num_edges;
    int32_t directed;
    int32_t *offsets;  /* num_vertices + 1, malloc'd */
    int32_t *nbrs;     /* num_edges */
    int32_t *weights;  /* num_edges; NULL for .sg */
} gi_host_csr;
/* Read a graph file into host arrays (free with gi_host_csr_free). */
int gi_read_graph_file(const char *path, int32_t flags, gi_host_csr *out);
void gi_host_csr_free(gi_host_csr *c);
/* Write the reference's .graphit_bin cache (graph.h:286-293): int32 V,
 * int32 E, sources[E], neighbours[E], weights[E] (1 when absent),
 * offsets[V+1]. */
int gi_write_graphit_bin(const gi_host_csr *c, const char *path);
/* Read a graph file and build the device graph. With GI_LOAD_CACHE the
 * reference's cache next to the file ("<path>.i.graphit_bin", or
 * "<path>.sym.i.graphit_bin" when symmetrized) is read when present and
 * written when not (graph.h:239-293). */
int gi_graph_load(const char *path, int32_t flags, gi_graph **out);

/* Vertex layout used by the fused engines (no reference counterpart: the
 * reference keeps the file's vertex order, graph.h:236-338). DEGREE builds an
 * internal copy renumbered by decreasing degree; fused BFS (single- and
 * multi-source), fused CC and fused PageRank then run on it and map their
 * results back to the original ids, so outputs are unchanged (fused SSSP
 * keeps the native order). NATIVE drops the copy. Costs one more CSR in
 * device memory. */
enum gi_layout { GI_LAYOUT_NATIVE = 0, GI_LAYOUT_DEGREE = 1 };
int gi_graph_set_layout(gi_graph *g, int32_t layout);
int32_t gi_graph_layout(const gi_graph *g);

/* ------------------------------------------------------------------ frontier
 * Replaces VertexFrontier (runtime/infra_gpu/vertex_frontier.h:26-166) and
 * the conversions of vertex_representation.h:9-158. One frontier holds a
 * sparse queue, a bitmap (64-bit words) and a boolmap, double-buffered. */
typedef struct gi_frontier gi_frontier;
enum gi_format { GI_FMT_SPARSE = 0, GI_FMT_BITMAP = 1, GI_FMT_BOOLMAP = 2 };

/* create_new_vertex_set (vertex_frontier.h:126-166); init_all fills 0..V-1
 * (initialize_frontier_all, vertex_frontier.h:119-125). */
int gi_frontier_create(int32_t num_vertices, int32_t init_all, gi_frontier **out);
int gi_frontier_destroy(gi_frontier *f);
/* builtin_addVertex (vertex_frontier.h:168-175); queued on stream. */
int gi_frontier_add_vertex(gi_frontier *f, int32_t v, void *stream);
/* Empty the input side. */
int gi_frontier_clear(gi_frontier *f, void *stream);
/* size_host (vertex_frontier.h:66-68, 109-113). Synchronises. */
int gi_frontier_size(gi_frontier *f, int32_t *out, void *stream);
/* to_sparse_host / to_bitmap / to_boolmap (runtime.cpp:19-21,
 * vertex_representation.h:72-158): converts the input side in place. */
int gi_frontier_convert(gi_frontier *f, int32_t format, void *stream);
int gi_frontier_format(const gi_frontier *f, int32_t *format);
/* Copy the input side as a sparse list (converted first). Synchronises. */
int gi_frontier_read(gi_frontier *f, int32_t *out, int32_t capacity, int32_t *count,
                     void *stream);
/* swap_queues / swap_bitmaps / swap_bytemaps (vertex_frontier.h:217-316):
 * output side becomes input, new output side emptied. */
int gi_frontier_swap(gi_frontier *f, int32_t format, void *stream);
/* dedup_frontier_perfect (vertex_frontier.h:353-386). */
int gi_frontier_dedup(gi_frontier *f, void *stream);

/* ------------------------------------------------------------------ operators
 * edgeset_apply (src/graphit/operators.cpp:513-650, operators.h:38-175).
 * A C ABI cannot carry device lambdas, so each app UDF of the reference is a
 * compiled operator id; its vertex data come in gi_op_args. */
enum gi_op {
    /* apps/bfs.cpp:23-29: if parent[dst]==-1, CAS parent[dst]=src, enqueue.
     * data0 = int32 parent[V]. */
    GI_OP_BFS_PARENT = 0,
    /* BFS that records depth: data0 = int32 depth[V], i0 = current level;
     * dst filter depth[dst]==-1, write depth[dst]=i0+1 once, enqueue. */
    GI_OP_BFS_DEPTH = 1,
    /* apps/sssp.cpp:24-27 + updatePriorityMin (graphit_types.h:426-430):
     * nd = SP[src]+w; writeMin(SP[dst], nd) && nd < i0 (cutoff) -> enqueue.
     * data0 = int32 SP[V]. */
    GI_OP_SSSP_RELAX = 2,
    /* apps/cc.cpp:23-38 hook: data0 = int32 ID[V], data1 = int32 updates[2]. */
    GI_OP_CC_HOOK = 3,
    /* apps/pr.cpp:32-34: new_rank[dst] += contrib[src];
     * data0 = float new_rank[V], data1 = float contrib[V]. */
    GI_OP_PR_PULL = 4,
    /* apps/bc.cpp:34-36: num_paths[dst] += num_paths[src], filter
     * visited[dst]==0; data0 = double num_paths[V], data1 = uint8 visited[V]. */
    GI_OP_BC_FORWARD = 5,
    /* apps/bc.cpp:54-56: dependences[dst] += dependences[src], filter
     * visited[dst]==0; data0 = float dependences[V], data1 = uint8 visited[V]. */
    GI_OP_BC_BACKWARD = 6
};
typedef struct gi_op_args {
    void *data0;
    void *data1;
    void *data2;
    int32_t i0;
    int32_t i1;
    float f0;
    float f1;
} gi_op_args;

/* edgeset_apply(s).from(from).[to(filter)].apply(...) when `to` is NULL, or
 * .apply_modified(graph, to, tracked, udf, allow_dupes) when it is not
 * (operators.h:112-173): the tracked writes enqueue into `to` in the
 * schedule's frontier_creation format and `to` is swapped afterwards.
 * `from` NULL enumerates all vertices. Direction PULL uses the transposed
 * graph, converts `from` to the pull representation and enumerates the
 * vertices passing the op's filter (create_reverse_frontier,
 * operators.cpp:493-512). `from` and `to` may be the same frontier. */
int gi_edgeset_apply(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to,
                     int32_t op, const gi_op_args *args, void *stream);
/* The hybrid form (operators.cpp:516-532): picks s1 or s2 by the selector. */
int gi_edgeset_apply_hybrid(gi_graph *g, const gi_hybrid_schedule *h, gi_frontier *from,
                            gi_frontier *to, int32_t op, const gi_op_args *args,
                            void *stream);

/* vertexset_apply (operators.cpp:10-65) with the apps' vertex UDFs. */
enum gi_vop {
    GI_VOP_FILL_I32 = 0,      /* data0[v] = i0 (bfs reset, sssp reset, cc init uses iota) */
    GI_VOP_IOTA_I32 = 1,      /* data0[v] = v (apps/cc.cpp:39-41) */
    GI_VOP_CC_PJUMP = 2,      /* apps/cc.cpp:43-50; data0 = ID, data1 = updates[2] */
    GI_VOP_PR_CONTRIB = 3,    /* apps/pr.cpp:28-30; data0 = old, data1 = contrib */
    GI_VOP_PR_UPDATE = 4,     /* apps/pr.cpp:36-41; data0 = old, data1 = new, f0 = damp */
    GI_VOP_PR_RESET = 5,      /* apps/pr.cpp:43-46; data0 = old, data1 = new */
    GI_VOP_BC_RESET = 6,      /* apps/bc.cpp:28-32; data0 = dep, data1 = paths, data2 = visited */
    GI_VOP_BC_MARK = 7,       /* apps/bc.cpp:42-47; data2 = visited, i0 = value */
    GI_VOP_BC_BACKWARD_V = 8, /* apps/bc.cpp:49-52; data0 = dep, data1 = paths, data2 = visited */
    GI_VOP_BC_FINAL = 9       /* apps/bc.cpp:58-63; data0 = dep, data1 = paths */
};
/* Over all vertices of g when f is NULL, else over f's members. */
int gi_vertexset_apply(gi_graph *g, gi_frontier *f, int32_t vop, const gi_op_args *args,
                       void *stream);

/* ------------------------------------------------------------- frontier list
 * VertexFrontierList (runtime/infra_gpu/list.h:18-116): stack of levels. */
typedef struct gi_frontier_list gi_frontier_list;
int gi_frontier_list_create(int32_t max_elems, gi_frontier_list **out);
int gi_frontier_list_insert(gi_frontier_list *l, gi_frontier *f, void *stream);
/* Pops the top level into f's input side; GI_ERR_UNDERFLOW when empty
 * (the reference asserts, list.h:90-92). */
int gi_frontier_list_retrieve(gi_frontier_list *l, gi_frontier *f, void *stream);
int gi_frontier_list_levels(const gi_frontier_list *l, int32_t *out);
int gi_frontier_list_destroy(gi_frontier_list *l);

/* ------------------------------------------------------------ priority queue
 * GPUPriorityQueue<int> (runtime/infra_gpu/gpu_priority_queue.h:21-247). */
typedef struct gi_pq gi_pq;
/* init(graph, prio, initial_priority, delta, src) (gpu_priority_queue.h:33-47). */
int gi_pq_create(gi_graph *g, int32_t *d_priorities, int32_t initial_priority,
                 int32_t delta, int32_t source, gi_pq **out);
/* finished() (gpu_priority_queue.h:63-75). Synchronises. */
int gi_pq_finished(gi_pq *q, int32_t *finished, void *stream);
/* dequeue() (gpu_priority_queue.h:115-156): borrowed frontier of the
 * current window [cur, cur+delta). */
int gi_pq_dequeue(gi_pq *q, gi_frontier **frontier, void *stream);
/* prio_cutoff() (gpu_priority_queue.h:29-31). */
int gi_pq_cutoff(gi_pq *q, int32_t *cutoff);
int gi_pq_destroy(gi_pq *q);

/* --------------------------------------------------------- fused app drivers
 * Whole algorithms with the loop on the GPU side of the boundary (the fused
 * mode of fuse_kernel, operators.cpp:651-672). Per-run statistics are
 * returned in gi_run_stats when stats != NULL. */
typedef struct gi_run_stats {
    int32_t iterations;       /* levels / buckets / sweeps */
    int32_t push_iterations;
    int32_t pull_iterations;
    int32_t reached;          /* vertices reached */
    int64_t edges_traversed;  /* edge entries read by the traversal kernels */
    double ms;                /* device time of the run (hipEvents) */
} gi_run_stats;

/* BFS (apps/bfs.cpp). d_depth required; d_parent optional (CAS parent as
 * the reference UDF). h may be NULL: push TWCE (apps/bfs.cpp:86-88). */
int gi_bfs(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, int32_t *d_depth,
           int32_t *d_parent, gi_run_stats *stats, void *stream);
/* Per-source BFS depths for many sources at once: d_depth is
 * [num_sources][V] (source-major; -1 where a source does not reach v). The
 * BFS app (apps/bfs.cpp:44-60) runs one source per call; here groups of up
 * to 64 sources advance together, one 64-bit mask per vertex, with
 * direction-optimizing push/pull levels recorded as hipGraph blocks. A
 * directed graph pulls over its transpose. stats: iterations = levels summed
 * over the groups, ms = device time. Synchronises. */
int gi_bfs_batched(gi_graph *g, const int32_t *sources, int32_t num_sources, int32_t *d_depth,
                   gi_run_stats *stats, void *stream);
/* Delta-stepping SSSP from many sources at once: d_dist is
 * [num_sources][V] (source-major; INT32_MAX where unreached). The SSSP app
 * (apps/sssp.cpp:34-57, gpu_priority_queue.h:115-185) runs one source per
 * call; here passes of up to `batch` sources (<= 64; 0 = 64) run together,
 * one wave lane per source over vertex-major distance rows, sharing one
 * window of width delta: light steps, one heavy step per window, push or
 * pull per step. A pass shares its window among its sources, so it favours
 * wider windows than one source does; the distances do not depend on delta.
 * Distances equal gi_sssp's. stats: iterations = relax steps summed over the
 * passes (pull_iterations of them pulled), ms = device time. Synchronises. */
int gi_sssp_batched(gi_graph *g, const int32_t *sources, int32_t num_sources, int32_t delta, int32_t batch,
                    int32_t *d_dist, gi_run_stats *stats, void *stream);
/* BFS that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device or host memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It runs
 * gi_bfs's unfused push loop (s: push, sparse frontier creation, load
 * balancing VERTEX_BASED, WM, TWCE or CM; NULL = push TWCE) with the
 * frontier queues, counters and visited bitmap in a fixed device array of
 * the library (about 17 MB per GPU, held until the process ends), so the
 * graph may have at most GI_BFS_STATIC_MAX_VERTICES vertices (larger graphs
 * and other schedules return an error). The caller passes d_depth (and
 * optionally d_parent); results are gi_bfs's. Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU race. Synchronises once
 * per level (a 4-byte read of the frontier size). */
#define GI_BFS_STATIC_MAX_VERTICES (1 << 21)
int gi_bfs_static(gi_graph *g, int32_t source, const gi_schedule *s, int32_t *d_depth,
                  int32_t *d_parent, void *stream);
/* BFS with a record per level: direction and load balancing used, frontier
 * sizes, the traversal kernel's device time (hipEvents around that launch)
 * and the step's algorithmic HBM bytes:
 *   push: 12|F| + 8*sum_{u in F} deg(u) + 8|F'|
 *   pull: 4V + 8|U| + 8*sum_{v in U} k_v + 8|F'|
 * (U = vertices unvisited before the step, k_v = neighbours examined up to
 * and including the first frontier member, deg(v) when none). The counts
 * are computed after the run from the depths; the run itself is gi_bfs.
 * *num_records gets the number of levels (records beyond max_records are
 * dropped). Synchronises. */
typedef struct gi_level_record {
    int32_t level;
    int32_t direction;
    int32_t load_balancing;
    int32_t frontier_size;
    int32_t next_size;
    int32_t reserved;
    double kernel_ms;
    double alg_bytes;
} gi_level_record;
int gi_bfs_profile(gi_graph *g, int32_t source, const gi_hybrid_schedule *h, int32_t *d_depth,
                   gi_level_record *records, int32_t max_records, int32_t *num_records,
                   void *stream);
/* Sources for traversal benchmarks: uniform draws skipping degree-0
 * vertices (SourcePicker, runtime/infra_gapbs/benchmark.h:48-71) from the
 * counter-based generator (stream 6 of seed). */
int gi_graph_pick_sources(const gi_graph *g, uint64_t seed, int32_t count, int32_t *out);
/* Number of CSR entries whose source vertex has d_depth >= 0 (twice the
 * undirected edges of the reached component on a symmetric graph). */
int gi_reached_edges(const gi_graph *g, const int32_t *d_depth, int64_t *out, void *stream);

/* ------------------------------------------------------- learned selector
 * No counterpart in the reference (its only schedule choice is the fixed
 * hybrid rule, operators.cpp:526). With kernel_fusion ENABLED in s1, gi_bfs
 * runs the fused engine: at the start of each level's step launch the device picks one of
 * GI_BFS_MODES step modes, either by GraphIt's rule over (s1, s2) or, with
 * selector GI_SELECT_LEARNED and a loaded model, by a 16-64-64-K MLP that
 * predicts each mode's log time (BF16 weights, FP32 accumulation); outputs
 * are identical whichever mode runs. Modes: 0 push lane-per-vertex, 1 push
 * wave-per-vertex, 2 push TWCE, 3 push CM, 4 push TWC with a grid-wide hub
 * bin, 5 pull lane-per-vertex, 6 pull 8-lanes-per-vertex. */
#define GI_BFS_MODES 7
/* Parameters: w1[16][64], b1[64], w2[64][64], b2[64], w3[64][n_out], b3[n_out]
 * (row-major [in][out]), FP32; stored on the device as BF16. Features are
 * normalised as (x - feat_mean) * feat_scale before the first layer. */
int32_t gi_selector_param_count(int32_t n_out);
int gi_selector_load(const float *params, int32_t n_out, int32_t allowed_mask,
                     const float *feat_mean, const float *feat_scale);
int gi_selector_clear(void);
int32_t gi_selector_loaded(void);
/* One forward + backward pass of the FP32 training network on the device:
 * masked mean squared error between outputs and T over entries where M = 1.
 * X[batch][16] (already normalised), T and M [batch][n_out]. */
int gi_selector_loss_grad(const float *X, const float *T, const float *M, int32_t batch,
                          int32_t n_out, const float *params, float *grads, float *loss);
/* Full-batch Adam for `epochs` steps on the device; params updated in
 * place; loss_history[epochs] optional. */
int gi_selector_train(const float *X, const float *T, const float *M, int32_t batch,
                      int32_t n_out, int32_t epochs, float lr, float weight_decay,
                      float *params, float *loss_history);
/* Bin thresholds of the fused push steps (chosen by the project's autotuner,
 * easy_graphit_amd/autotune.py): vertices with degree < lane_max are walked
 * by one lane, wider ones by a wave, and in TWC mode those with degree >=
 * hub_min by the whole grid; blocks_per_cu sets the persistent grid
 * (1..8 workgroups of 256 per CU). Per graph; no reference counterpart. */
int gi_bfs_set_bins(gi_graph *g, int32_t lane_max, int32_t hub_min, int32_t blocks_per_cu);
/* Multi-source BFS with the fused engine (kernel_fusion must be ENABLED in
 * h->s1): d_depth[v] = distance to the nearest source. Over a graph that is
 * a disjoint union with one source per component (many small graphs packed
 * in one CSR) this runs every component's BFS in the same launches. The level
 * loop is a recorded hipGraph replayed until the frontier is empty. The
 * reference only seeds one source (apps/bfs.cpp:49-52). */
int gi_bfs_multi(gi_graph *g, const int32_t *sources, int32_t num_sources,
                 const gi_hybrid_schedule *h, int32_t *d_depth, int32_t *d_parent,
                 int32_t *levels, void *stream);
/* Many small uniform random graphs packed as one symmetric CSR:
 * num_graphs graphs, vertex counts log-uniform in [vmin, vmax], avg_degree
 * edge entries per vertex (avg_degree/2 undirected pairs, squished). Vertex
 * id ranges are contiguous; vertex_offsets[num_graphs+1] (host) receives
 * them. */
int gi_graph_uniform_batch(int32_t num_graphs, int32_t vmin, int32_t vmax, int32_t avg_degree,
                           uint64_t seed, int32_t *vertex_offsets, gi_graph **out);

/* Dataset recorder: a BFS from source where every level is run once per
 * mode of mode_mask from the same snapshot and timed (min of reps); the run
 * then continues with `follow` (or the fastest mode when follow < 0).
 * features[max_levels][16] (raw, unnormalised), times_ms[max_levels][7]
 * (-1 for modes not measured). Synchronises. */
int gi_bfs_sweep(gi_graph *g, int32_t source, int32_t mode_mask, int32_t follow, int32_t reps,
                 float *features, float *times_ms, int32_t max_levels, int32_t *num_levels,
                 void *stream);

/* Delta-stepping SSSP (apps/sssp.cpp:34-57); d_dist int32, INT_MAX unreached.
 * Default schedule (s == NULL): the app's PUSH TWCE with the bucketed
 * priority queue (apps/sssp.cpp:77-79). With kernel_fusion ENABLED the
 * windows and relaxation phases run on the device (near frontier as a
 * bitmap, pending vertices as per-unit bounds, one recorded launch per step,
 * no host synchronisation per step; max_cta > 0 sets the workgroups of the
 * step launches, 0 the resident grid). */
int gi_sssp(gi_graph *g, int32_t source, int32_t delta, const gi_schedule *s,
            int32_t *d_dist, gi_run_stats *stats, void *stream);
/* Measurement hook of the fused SSSP engine (no reference counterpart):
 * runs the engine's step launches one at a time, outside a graph, and times
 * each with HIP events. pull_min: |F| above which a relax step pulls (< 0:
 * the engine's default). kernel_ms[max_slots][6]: push relax step, pull
 * relax step (one launch, in the column of its direction, 0 in the other),
 * hub bin (0 for a graph whose out-degrees stay below the hub width: not
 * launched), 0, scan-and-relax step of a new window, 0. info[max_slots][5]:
 * the step's record: mode (0 relax / 1 scan-and-relax / 2 heavy-edge step
 * of a settled window), pull flag, |F| of a relax step, window end hi,
 * out-edges of F. Synchronises. */
int gi_sssp_trace(gi_graph *g, int32_t source, int32_t delta, int32_t pull_min, int32_t *d_dist,
                  int32_t max_slots, float *kernel_ms, int64_t *info, int32_t *num_slots, void *stream);
/* Connected components, hook + pointer jumping (apps/cc.cpp:53-78); labels
 * are the minimum vertex id of each component. Default schedule (s == NULL):
 * the app's PUSH EDGE_ONLY (apps/cc.cpp:84-86). With kernel_fusion ENABLED
 * (symmetric graphs) the rounds run as a recorded graph with device-side
 * termination and full shortcutting per round: EDGE_ONLY sweeps every edge
 * each round; VERTEX_BASED first hooks every vertex's first two neighbours,
 * then runs rounds over the vertices outside the most frequent label only
 * (an edge whose ends differ always has such an end). Same labels either
 * way. */
int gi_cc(gi_graph *g, const gi_schedule *s, int32_t *d_labels, gi_run_stats *stats,
          void *stream);
/* PageRank (apps/pr.cpp:49-69): iters iterations, FP32; d_rank = old_rank.
 * Default schedule (s == NULL): the app's PULL EDGE_ONLY BLOCKED 0x42000
 * (apps/pr.cpp:78-81). With kernel_fusion ENABLED (direction PULL) the whole
 * run is one recorded graph of deterministic pull sweeps over the in-edges
 * (no atomics). */
int gi_pagerank(gi_graph *g, int32_t iters, float damping, const gi_schedule *s,
                float *d_rank, gi_run_stats *stats, void *stream);
/* Brandes BC from num_sources sources (apps/bc.cpp:64-125), summed into
 * d_scores (float, zeroed first). batch > 1: passes of up to min(batch, 64)
 * sources advance level-synchronously, one lane per source over vertex-major
 * rows, the level loops on the device (recorded); batch <= 1: one source at
 * a time as the app runs unfused (push TWCE, a host round trip per level,
 * apps/bc.cpp:149-151). */
int gi_bc(gi_graph *g, const int32_t *h_sources, int32_t num_sources, int32_t batch,
          float *d_scores, gi_run_stats *stats, void *stream);
/* BC one source at a time under a schedule (apps/bc.cpp:64-120 with its s1
 * and to_fuse). s == NULL: the app's push TWCE, unfused. kernel_fusion
 * ENABLED: the fused mode (fuse_kernel, apps/bc.cpp:92-113): the forward and
 * backward level loops run on the device, push or pull per level (pull only
 * on symmetric graphs). Otherwise push steps with s's load balancing (a PULL
 * direction needs kernel_fusion: GI_ERR_UNSUPPORTED). */
int gi_bc_schedule(gi_graph *g, const int32_t *h_sources, int32_t num_sources, const gi_schedule *s,
                   float *d_scores, gi_run_stats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif
