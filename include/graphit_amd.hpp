// graphit_amd.hpp — C++ host mirror of the GraphIt operator and runtime API
// over the C ABI of graphit_amd.h.
//
// Names, argument meaning and chaining follow the reference's DSL
// (include/graphit/{graphit_types,schedule,operators,runtime}.h) so an app
// body reads like apps/{bfs,sssp,cc,pr,bc}.cpp. The difference is the UDF:
// the reference stages a C++ lambda into CUDA through BuildIt, which a C ABI
// cannot carry, so a UDF here is one of the compiled operators of the
// runtime (udf::*, filter::*), and vertex data live in device memory owned by
// VertexData<T>. Errors throw graphit::Error with gi_last_error()'s text
// where the reference asserts or exits.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "graphit_amd.h"

namespace graphit {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != GI_OK) throw Error(rc, gi_last_error());
}

// ---------------------------------------------------------------- schedules
// SimpleGPUSchedule (include/graphit/schedule.h:26-118).
struct Schedule {
    virtual ~Schedule() = default;
};

struct SimpleGPUSchedule : Schedule {
    enum class direction_type { PUSH = GI_PUSH, PULL = GI_PULL };
    enum class pull_frontier_rep_type { BITMAP = GI_PULL_BITMAP, BOOLMAP = GI_PULL_BOOLMAP };
    enum class frontier_creation_type { SPARSE = GI_FC_SPARSE, BITMAP = GI_FC_BITMAP, BOOLMAP = GI_FC_BOOLMAP };
    enum class deduplication_type { DISABLED = GI_DEDUP_DISABLED, ENABLED = GI_DEDUP_ENABLED };
    enum class load_balancing_type {
        VERTEX_BASED = GI_LB_VERTEX_BASED, TWC = GI_LB_TWC, TWCE = GI_LB_TWCE, WM = GI_LB_WM,
        CM = GI_LB_CM, STRICT = GI_LB_STRICT, EDGE_ONLY = GI_LB_EDGE_ONLY
    };
    enum class edge_blocking_type { BLOCKED = GI_BLOCKED, UNBLOCKED = GI_UNBLOCKED };
    enum class kernel_fusion_type { DISABLED = GI_FUSION_DISABLED, ENABLED = GI_FUSION_ENABLED };

    static int default_max_cta;
    static int default_cta_size;

    gi_schedule s;

    SimpleGPUSchedule() {
        gi_schedule_init(&s);
        s.max_cta = default_max_cta;
        s.cta_size = default_cta_size;
    }
    void configDirection(direction_type d, pull_frontier_rep_type pf = pull_frontier_rep_type::BOOLMAP) {
        s.direction = (int32_t)d;
        if (d == direction_type::PULL) s.pull_frontier_rep = (int32_t)pf;
    }
    void configFrontierCreation(frontier_creation_type fc) { s.frontier_creation = (int32_t)fc; }
    void configDeduplication(deduplication_type d) { s.deduplication = (int32_t)d; }
    void configLoadBalancing(load_balancing_type lb, edge_blocking_type eb = edge_blocking_type::UNBLOCKED,
                             int bs = -1) {
        s.load_balancing = (int32_t)lb;
        if (lb == load_balancing_type::EDGE_ONLY) {
            s.edge_blocking = (int32_t)eb;
            s.block_size = bs;
        }
    }
    void configKernelFusion(kernel_fusion_type f) { s.kernel_fusion = (int32_t)f; }
};

inline int SimpleGPUSchedule::default_max_cta = 0;     // 0: size the grid to the device
inline int SimpleGPUSchedule::default_cta_size = 512;  // schedule.cpp:5

// HybridGPUSchedule (include/graphit/schedule.h:122-139).
struct HybridGPUSchedule : Schedule {
    gi_hybrid_schedule h;
    HybridGPUSchedule(SimpleGPUSchedule &s1, SimpleGPUSchedule &s2) {
        h.s1 = s1.s;
        h.s2 = s2.s;
        h.threshold = 0.f;
        h.selector = GI_SELECT_HEURISTIC;
    }
    void configThreshold(float t) { h.threshold = t; }
    void bindThreshold(float t) { h.threshold = t; }  // runtime value (schedule.cpp:9-12)
    void configSelector(int32_t mode) { h.selector = mode; }
};

// ---------------------------------------------------------------- data
template <class T>
struct VertexData {
    T *ptr = nullptr;
    int32_t n = 0;
    VertexData() = default;
    explicit VertexData(int32_t count) { allocate(count); }
    VertexData(const VertexData &) = delete;
    VertexData &operator=(const VertexData &) = delete;
    ~VertexData() { gi_free(ptr); }
    void allocate(int32_t count) {  // VertexData::allocate (graphit_types.h:346-406)
        gi_free(ptr);
        n = count;
        void *p = nullptr;
        check(gi_malloc(&p, (uint64_t)count * sizeof(T)));
        ptr = (T *)p;
    }
    std::vector<T> to_host() const {
        std::vector<T> out((size_t)n);
        check(gi_memcpy_d2h(out.data(), ptr, sizeof(T) * (size_t)n, nullptr));
        return out;
    }
    void from_host(const std::vector<T> &v) {
        check(gi_memcpy_h2d(ptr, v.data(), sizeof(T) * v.size(), nullptr));
    }
    void set(int32_t i, T value) {  // host-side element write, e.g. parent[src] = src
        check(gi_memcpy_h2d(ptr + i, &value, sizeof(T), nullptr));
    }
};

struct GraphT {
    gi_graph *g = nullptr;
    bool owner = true;
    int32_t num_vertices = 0;
    int64_t num_edges = 0;
    GraphT() = default;
    explicit GraphT(gi_graph *h, bool own = true) : g(h), owner(own) {
        check(gi_graph_num_vertices(g, &num_vertices));
        check(gi_graph_num_edges(g, &num_edges));
    }
    GraphT(GraphT &&o) noexcept { *this = std::move(o); }
    GraphT &operator=(GraphT &&o) noexcept {
        g = o.g; owner = o.owner; num_vertices = o.num_vertices; num_edges = o.num_edges;
        o.g = nullptr;
        return *this;
    }
    GraphT(const GraphT &) = delete;
    ~GraphT() {
        if (g && owner) gi_graph_destroy(g);
    }
    GraphT get_transposed_graph() {
        gi_graph *t = nullptr;
        check(gi_graph_transpose(g, &t));
        return GraphT(t, false);
    }
};

struct VertexSubset {  // VertexFrontier
    gi_frontier *f = nullptr;
    explicit VertexSubset(int32_t n, bool all = false) { check(gi_frontier_create(n, all ? 1 : 0, &f)); }
    VertexSubset(const VertexSubset &) = delete;
    ~VertexSubset() { gi_frontier_destroy(f); }
    void addVertex(int32_t v) { check(gi_frontier_add_vertex(f, v, nullptr)); }
    int32_t size() {
        int32_t n = 0;
        check(gi_frontier_size(f, &n, nullptr));
        return n;
    }
};

struct FrontierList {  // VertexFrontierList (runtime/infra_gpu/list.h)
    gi_frontier_list *l = nullptr;
    explicit FrontierList(int32_t max_elems) { check(gi_frontier_list_create(max_elems, &l)); }
    ~FrontierList() { gi_frontier_list_destroy(l); }
    void insert(VertexSubset &s) { check(gi_frontier_list_insert(l, s.f, nullptr)); }
    void retrieve(VertexSubset &s) { check(gi_frontier_list_retrieve(l, s.f, nullptr)); }
};

// ---------------------------------------------------------------- UDFs
// The compiled counterparts of the apps' UDFs (operator id + vertex data).
struct Udf {
    int32_t op;
    gi_op_args args;
};

namespace udf {
// apps/bfs.cpp:23-25 (with the CAS patch-up of graphit.cpp:55-122)
inline Udf bfs_parent(VertexData<int32_t> &parent) {
    Udf u{GI_OP_BFS_PARENT, {}};
    u.args.data0 = parent.ptr;
    return u;
}
// apps/sssp.cpp:24-27
inline Udf sssp_relax(VertexData<int32_t> &SP, int32_t cutoff) {
    Udf u{GI_OP_SSSP_RELAX, {}};
    u.args.data0 = SP.ptr;
    u.args.i0 = cutoff;
    return u;
}
// apps/cc.cpp:23-38
inline Udf cc_hook(VertexData<int32_t> &ID, VertexData<int32_t> &updates) {
    Udf u{GI_OP_CC_HOOK, {}};
    u.args.data0 = ID.ptr;
    u.args.data1 = updates.ptr;
    return u;
}
// apps/pr.cpp:32-34
inline Udf pr_update_edge(VertexData<float> &new_rank, VertexData<float> &contrib) {
    Udf u{GI_OP_PR_PULL, {}};
    u.args.data0 = new_rank.ptr;
    u.args.data1 = contrib.ptr;
    return u;
}
}  // namespace udf

// ---------------------------------------------------------------- operators
// edgeset_apply (include/graphit/operators.h:38-125).
struct edgeset_apply {
    const gi_schedule *simple = nullptr;
    const gi_hybrid_schedule *hybrid = nullptr;
    gi_schedule def;
    gi_frontier *from_set = nullptr;
    edgeset_apply() {
        gi_schedule_init(&def);
        simple = &def;
    }
    explicit edgeset_apply(SimpleGPUSchedule &s) : simple(&s.s) {}
    explicit edgeset_apply(HybridGPUSchedule &h) : hybrid(&h.h) {}
    edgeset_apply &from(VertexSubset &f) {
        from_set = f.f;
        return *this;
    }
    // the `to` filter is part of the compiled operator (its dst filter)
    edgeset_apply &to_filter_of_udf() { return *this; }
    void apply(GraphT &g, Udf u) { run(g, nullptr, u); }
    void apply_modified(GraphT &g, VertexSubset &to, Udf u) { run(g, to.f, u); }

  private:
    void run(GraphT &g, gi_frontier *to, Udf &u) {
        if (hybrid) check(gi_edgeset_apply_hybrid(g.g, hybrid, from_set, to, u.op, &u.args, nullptr));
        else check(gi_edgeset_apply(g.g, simple, from_set, to, u.op, &u.args, nullptr));
    }
};

// vertexset_apply (operators.cpp:10-65) with the apps' vertex UDFs.
inline void vertexset_apply(GraphT &g, int32_t vop, gi_op_args a) {
    check(gi_vertexset_apply(g.g, nullptr, vop, &a, nullptr));
}
inline void vertexset_apply(GraphT &g, VertexSubset &f, int32_t vop, gi_op_args a) {
    check(gi_vertexset_apply(g.g, f.f, vop, &a, nullptr));
}

namespace runtime {
inline GraphT load_rmat(int scale, int edge_factor, uint64_t seed, bool symmetrize) {
    gi_graph *g = nullptr;
    check(gi_graph_rmat(scale, edge_factor, seed, symmetrize ? 1 : 0, &g));
    return GraphT(g);
}
inline GraphT load_csr(const std::vector<int32_t> &off, const std::vector<int32_t> &nbr,
                       bool directed) {
    gi_graph *g = nullptr;
    check(gi_graph_from_csr((int32_t)off.size() - 1, (int64_t)nbr.size(), off.data(), nbr.data(),
                            nullptr, directed ? 1 : 0, &g));
    return GraphT(g);
}
inline float start_timer() { check(gi_timer_start(nullptr)); return 0.f; }
inline float stop_timer() {
    float ms = 0.f;
    check(gi_timer_stop(nullptr, &ms));
    return ms / 1000.f;  // seconds, as gpu_intrinsics.h:72-77
}
}  // namespace runtime

}  // namespace graphit
