// Internal definitions shared by the HIP translation units of libgraphit_amd.
// gfx950 (CDNA4) only: wave64, 64-bit ballots, 64-bit bitmap words.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/graphit_amd.h"

namespace gi {

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

#define GI_HIP(call)                                                              \
    do {                                                                          \
        hipError_t _e = (call);                                                   \
        if (_e != hipSuccess)                                                     \
            return ::gi::fail(GI_ERR_HIP, "%s failed: %s (%s:%d)", #call,         \
                              hipGetErrorString(_e), __FILE__, __LINE__);         \
    } while (0)

#define GI_TRY(call)                   \
    do {                               \
        int _rc = (call);              \
        if (_rc != GI_OK) return _rc;  \
    } while (0)

#define GI_LAUNCH_CHECK() GI_HIP(hipGetLastError())

// ------------------------------------------------------------ launch graphs
// Where a sequence of kernel launches goes: queued on a stream, or appended
// to a hipGraph built node by node, each node after the one before (the
// order the stream gives). The fused engines build the graphs they replay
// this way, not by stream capture: this HIP runtime invalidates a capture in
// progress when another thread makes a device-wide synchronising call (a
// null-stream copy, a device synchronise), so captures failed when two
// threads ran engines on two graphs.
struct Launcher {
    hipStream_t stream = nullptr;   // eager launches when graph is null
    hipGraph_t graph = nullptr;     // the graph being built
    hipGraphNode_t last = nullptr;  // its last node
    hipError_t err = hipSuccess;    // first failure while adding nodes
    std::vector<std::shared_ptr<void>> keep;  // argument values, alive until instantiation

    explicit Launcher(hipStream_t s = nullptr) : stream(s) {}

    template <class... P, class... A>
    void operator()(void (*kernel)(P...), dim3 grid, dim3 block, A &&...args) {
        using Vals = std::tuple<std::decay_t<P>...>;
        void *ptrs[sizeof...(P) + 1];
        if (!graph) {  // a failure is left for hipGetLastError, as a <<<>>> launch leaves it
            Vals vals(std::forward<A>(args)...);
            set_ptrs(vals, ptrs, std::index_sequence_for<P...>{});
            (void)hipLaunchKernel(reinterpret_cast<const void *>(kernel), grid, block, ptrs, 0, stream);
            return;
        }
        if (err != hipSuccess) return;
        auto vals = std::make_shared<Vals>(std::forward<A>(args)...);
        keep.push_back(vals);
        set_ptrs(*vals, ptrs, std::index_sequence_for<P...>{});
        hipKernelNodeParams np = {};
        np.func = reinterpret_cast<void *>(kernel);
        np.gridDim = grid;
        np.blockDim = block;
        np.sharedMemBytes = 0;
        np.kernelParams = ptrs;
        np.extra = nullptr;
        hipGraphNode_t node = nullptr;
        err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &np);
        if (err == hipSuccess) last = node;
    }

  private:
    template <class T, size_t... I>
    static void set_ptrs(T &t, void **p, std::index_sequence<I...>) {
        ((p[I] = (void *)&std::get<I>(t)), ...);
    }
};

// Device time between construction and stop() on one stream (HIP events).
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) (void)hipEventRecord(a, s);
    }
    double stop() {
        float ms = 0.f;
        if (a && b && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess)
            (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Instantiates the graph of the launches `record(Launcher &)` makes.
template <class F>
int build_graph(F &&record, hipGraphExec_t *out) {
    Launcher L;
    hipError_t e = hipGraphCreate(&L.graph, 0);
    if (e != hipSuccess) return fail(GI_ERR_HIP, "graph create failed: %s", hipGetErrorString(e));
    record(L);
    e = L.err;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, L.graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(L.graph);
    if (e != hipSuccess) return fail(GI_ERR_HIP, "graph build failed: %s", hipGetErrorString(e));
    *out = exec;
    return GI_OK;
}

// Workgroups of `kernel` (launched with `block` threads) resident per CU, 4
// when the runtime cannot say. Callers keep it in a function-local static,
// whose initialisation is thread-safe.
inline int resident_per_cu(const void *kernel, int block) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
        per_cu = 4;
    return per_cu;
}

// Per-graph plans of one engine, keyed by the graph handle. The map itself
// is locked, so threads that each use their own graph may call in parallel;
// a plan belongs to its graph and is used by that graph's one thread at a
// time.
template <class Plan>
class PlanRegistry {
  public:
    Plan *find(const gi_graph *g) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = map_.find(g);
        return it == map_.end() ? nullptr : it->second;
    }
    void put(const gi_graph *g, Plan *p) {
        std::lock_guard<std::mutex> lock(mu_);
        map_[g] = p;
    }
    // removes the graph's plan and hands it to the caller (nullptr: none)
    Plan *take(const gi_graph *g) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = map_.find(g);
        if (it == map_.end()) return nullptr;
        Plan *p = it->second;
        map_.erase(it);
        return p;
    }

  private:
    std::mutex mu_;
    std::map<const gi_graph *, Plan *> map_;
};

// ----------------------------------------------------------------- device
struct DeviceInfo {
    int num_cus = 256;
    int max_threads_per_cu = 2048;
};
const DeviceInfo &device_info();
// Grid size for persistent grid-stride kernels: enough workgroups to fill
// every CU several times over.
int persistent_grid(int block, int per_cu = 8);

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t bitmap_words(int64_t n) { return (n + 63) / 64; }

// Device allocation helpers (return GI codes).
int dmalloc(void **p, size_t bytes);
template <class T>
int dalloc(T **p, size_t count) {
    return dmalloc(reinterpret_cast<void **>(p), count * sizeof(T) + 1);
}
void dfree(void *p);

// Small fill kernels (graph-capture safe, unlike hipMemsetAsync nodes).
int fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t s);
int fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);
int fill_u8(uint8_t *p, int64_t n, uint8_t v, hipStream_t s);
int fill_f32(float *p, int64_t n, float v, hipStream_t s);
int fill_f64(double *p, int64_t n, double v, hipStream_t s);

// Device-wide exclusive scan of int32 counts into int64-safe int32 prefix
// (n + 1 outputs, out[n] = total); rocPRIM under the hood, workspace cached.
int exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s);

// ----------------------------------------------------------------- views
struct DGraph {
    int32_t V;
    int32_t E;
    const int32_t *off;   // [V+1]
    const int32_t *nbr;   // [E]
    const int32_t *w;     // [E] or null (weight 1)
    const int32_t *esrc;  // [E] COO source or null
};

enum SinkMode : int32_t { SINK_NONE = 0, SINK_SPARSE = 1, SINK_BITMAP = 2, SINK_BOOLMAP = 3 };

// Where tracked writes go (the output side of a frontier).
struct Sink {
    int32_t mode;
    int32_t cap;         // sparse capacity
    int32_t *q;          // sparse queue
    uint64_t *bm;        // bitmap words
    uint8_t *bmap;       // boolmap bytes
    int32_t *count;      // element count
    int32_t *dedup;      // per-vertex epoch (sparse dedup) or null
    int32_t epoch;
    int32_t *overflow;   // set when a sparse append exceeds cap
};

// Enumerator: all vertices (q == null), or the members of a sparse queue
// whose length lives on the device.
struct VEnum {
    const int32_t *q;
    const int32_t *count;
    int32_t V;  // vertex count, and the capacity of q
    __device__ int32_t size() const {
        if (!q) return V;
        int32_t n = *count;
        return n < V ? n : V;
    }
    __device__ int32_t at(int32_t i) const { return q ? q[i] : i; }
};

// Membership test of the input side (pull's dst filter, edge-only's src filter).
struct Member {
    int32_t mode;        // GI_FMT_BITMAP or GI_FMT_BOOLMAP (or -1 = everyone)
    const uint64_t *bm;
    const uint8_t *bmap;
};

// ------------------------------------------------------- wave64 primitives
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated atomic add: every active lane passes its amount; returns
// the lane's exclusive offset from the single atomic the wave issues.
__device__ __forceinline__ int32_t wave_atomic_add(int32_t *ctr, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    int leader = __ffsll((unsigned long long)mask) - 1;
    int32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(ctr, (int32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (int32_t)lanes_below(mask);
}

__device__ __forceinline__ bool bit_test(const uint64_t *bm, int32_t v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// Returns true when this call set the bit.
__device__ __forceinline__ bool bit_claim(uint64_t *bm, int32_t v) {
    unsigned long long bit = 1ull << (v & 63);
    unsigned long long old = atomicOr((unsigned long long *)&bm[v >> 6], bit);
    return (old & bit) == 0;
}

// Byte flag set through a 32-bit atomic on the containing word; true when
// this call set it (exact counts, where the reference's bytemap enqueue
// over-counts, vertex_frontier.h:195-201).
__device__ __forceinline__ bool byte_claim(uint8_t *bmap, int32_t v) {
    uint32_t *word = reinterpret_cast<uint32_t *>(bmap + (v & ~3));
    uint32_t bit = 1u << (8 * (v & 3));
    uint32_t old = atomicOr(word, bit);
    return (old & (0xffu << (8 * (v & 3)))) == 0;
}

__device__ __forceinline__ bool member(const Member &m, int32_t v) {
    if (m.mode == GI_FMT_BITMAP) return bit_test(m.bm, v);
    if (m.mode == GI_FMT_BOOLMAP) return m.bmap[v] != 0;
    return true;
}

// Called by every active lane of the wave (pred may be false).
__device__ __forceinline__ void sink_emit(const Sink &s, bool pred, int32_t v) {
    if (s.mode == SINK_NONE) return;
    if (s.mode == SINK_SPARSE) {
        if (pred && s.dedup) pred = atomicMax(&s.dedup[v], s.epoch) < s.epoch;
        int32_t pos = wave_atomic_add(s.count, pred);
        if (pred) {
            if (pos < s.cap) s.q[pos] = v;
            else atomicOr(s.overflow, 1);
        }
    } else if (s.mode == SINK_BITMAP) {
        bool fresh = pred && bit_claim(s.bm, v);
        wave_atomic_add(s.count, fresh);
    } else {
        bool fresh = pred && byte_claim(s.bmap, v);
        wave_atomic_add(s.count, fresh);
    }
}

// Upper bound in a sorted int32 array: first index i in [lo, hi) with a[i] > x.
__device__ __forceinline__ int32_t upper_bound_i32(const int32_t *a, int32_t lo, int32_t hi,
                                                   int32_t x) {
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

}  // namespace gi

namespace gi {
struct GraphLayout;
}

// ----------------------------------------------------------- host handles
struct gi_graph {
    int32_t V = 0;
    int64_t E = 0;
    int32_t directed = 0;
    int32_t *d_off = nullptr;
    int32_t *d_nbr = nullptr;
    int32_t *d_w = nullptr;
    int32_t *d_esrc = nullptr;       // COO source ids, built on demand
    int32_t *d_bucket_end = nullptr;  // blocked graphs: bucket end offsets
    int32_t num_buckets = 0;
    int32_t block_size = -1;
    int32_t max_degree = -1;
    int32_t *d_scratch = nullptr;    // load-balance scratch (TWC bins, STRICT prefix)
    int64_t scratch_len = 0;
    gi_frontier *work = nullptr;     // frontier reused by the fused drivers
    uint64_t *d_visited = nullptr;   // visited bitmap of the fused BFS
    gi_graph *transposed = nullptr;  // owned unless is_transpose_of is set
    gi_graph *transpose_owner = nullptr;
    std::map<int32_t, gi_graph *> blocked;  // owned
    gi::GraphLayout *layout = nullptr;     // degree-ordered view (gi_graph_set_layout), owned
    bool borrowed = false;  // owned by another graph
    bool degree_view = false;  // the degree-ordered view of another graph (lists in increasing new id)

    gi::DGraph view() const {
        gi::DGraph g;
        g.V = V;
        g.E = (int32_t)E;
        g.off = d_off;
        g.nbr = d_nbr;
        g.w = d_w;
        g.esrc = d_esrc;
        return g;
    }
};

struct gi_frontier {
    int32_t V = 0;
    int32_t cap = 0;
    int32_t fmt = GI_FMT_SPARSE;  // format of the input side
    int32_t side = 0;             // index of the input side
    int32_t *q[2] = {nullptr, nullptr};
    uint64_t *bm[2] = {nullptr, nullptr};
    uint8_t *bmap[2] = {nullptr, nullptr};
    int32_t *counts = nullptr;    // [2] element counts + [2] overflow flags
    int32_t *dedup = nullptr;     // per-vertex epochs
    int32_t epoch = 0;
    int64_t words = 0;

    int32_t *count_in() const { return counts + side; }
    int32_t *count_out() const { return counts + (side ^ 1); }
    int32_t *overflow() const { return counts + 2; }
};

namespace gi {
int ensure_coo(gi_graph *g, hipStream_t s);
int ensure_max_degree(gi_graph *g, hipStream_t s);
int ensure_scratch(gi_graph *g, int64_t len);
int bc_batched(gi_graph *g, const int32_t *h_src, int32_t num_sources, int32_t batch, float *d_scores,
               hipStream_t s, int64_t *levels);
// Batches of more than one source (gi_bc_vmajor.hip): passes of up to 64
// sources, one lane each over vertex-major rows.
int bc_vertex_major(gi_graph *g, const int32_t *h_src, int32_t num_sources, int32_t batch, float *d_scores,
                    hipStream_t s, int64_t *levels);
// Scratch layout per graph: load-balance region first (TWC bins 3V+3, or the
// STRICT prefix V + chunk sums), then the pull enumerator (count, queue).
inline int64_t scratch_lb_len(int64_t V) {
    int64_t twc = 3 * V + 64, strict = V + 64 + 1024 + 64;
    return twc > strict ? twc : strict;
}
inline int64_t scratch_total_len(int64_t V) { return scratch_lb_len(V) + 64 + V + 64; }
int validate_schedule(const gi_schedule *s);
int plan_grid(const gi_schedule *s);
struct OpBfsVisit;
int edgeset_apply(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to,
                  int32_t op, const gi_op_args &a, hipStream_t st, const OpBfsVisit *visit,
                  hipEvent_t *events = nullptr);
int frontier_convert(gi_frontier *f, int32_t fmt, hipStream_t s);
int frontier_clear_out(gi_frontier *f, int32_t fmt, hipStream_t s);
Sink make_sink(gi_frontier *f, int32_t fmt, bool dedup);
Member make_member(const gi_frontier *f);
int read_count(const int32_t *d_count, int32_t *out, hipStream_t s);

// Fused whole-run engines and the per-graph state they cache (released by
// gi_graph_destroy).
int pr_fused_run(gi_graph *g, int32_t iters, float damping, float *d_rank, hipStream_t s);
void release_pr(const gi_graph *g);
// sampled: the VERTEX_BASED schedule (neighbour sampling, then rounds over
// the vertices outside the most frequent label); else full EDGE_ONLY sweeps.
// epilogue (optional): queued after each replay of rounds, before the host
// reads whether the run has ended.
int cc_fused_run(gi_graph *g, int32_t *d_labels, int32_t *rounds, bool sampled, hipStream_t s,
                 const std::function<int(hipStream_t)> *epilogue = nullptr);
void release_cc(const gi_graph *g);
// counts (optional): relax steps, pull steps, windows entered, recorded slots.
int sssp_fused_run(gi_graph *g, int32_t source, int32_t delta, int32_t *d_dist, int32_t *counts,
                   hipStream_t s, int32_t max_cta = 0);
void release_sssp(const gi_graph *g);
// The light/heavy split of g's out-lists for delta, shared with the fused
// SSSP engine (built on first use of a delta): each list reordered light
// edges (w < delta) first, lend[v] the end of v's light part. All null when
// no edge is heavy. The arrays stay valid until the split is rebuilt for
// another delta or the graph is released.
struct SsSplitView {
    const int32_t *nbr = nullptr, *wt = nullptr, *lend = nullptr;
};
int sssp_light_split(gi_graph *g, int32_t delta, SsSplitView *out, hipStream_t s);

// Degree-ordered layout (gi_layout.hip): the view the fused engines run on,
// source mapping, two V-sized scratch arrays in view order, and the
// mappings of results back to the original ids.
void release_layout(gi_graph *g);
int build_layout(gi_graph *g);  // degree-ordered copy (with the weights, if any)
gi_graph *layout_view(const gi_graph *g);
int32_t layout_map_source(const gi_graph *g, int32_t v);
int32_t *layout_tmp(const gi_graph *g, int i);
// First vertex of the view without edges (V when all have edges).
int32_t layout_iso_begin(const gi_graph *g);
int layout_map_back(const gi_graph *g, const int32_t *in_new, int32_t *out_old, bool ids, hipStream_t s);
// BFS depths of a single-source run back to original ids; on a symmetric
// graph the edgeless tail of the view is filled without reading depth_new.
int layout_depth_back(const gi_graph *g, const int32_t *depth_new, int32_t source_old, int32_t *out_old,
                      hipStream_t s);
// hint (optional, device): the label most vertices carry (e.g. the giant
// component's), folded with one atomic per workgroup.
int layout_labels_back(const gi_graph *g, const int32_t *labels_new, int32_t *out_old, hipStream_t s,
                       const int32_t *hint = nullptr);
// The device address of the giant label the sampled CC run on g found (-1 there: none).
const int32_t *cc_giant_label(gi_graph *g);
}  // namespace gi
