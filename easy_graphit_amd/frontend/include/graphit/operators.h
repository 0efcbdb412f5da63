// GraphIt's operators as the apps call them (include/graphit/operators.h:
// 20-178 of the reference): vertexset_apply, edgeset_apply(s).from(f).to(filter)
// .apply / .apply_modified / .apply_priority, fuse_kernel. Each call stages
// its UDFs in device context into a device functor (filter(v), apply(src,
// dst, w) returning whether dst enters the output frontier, or a vertex
// op(v)) and emits one call of the easy_graphit_amd runtime's templated
// launcher with that functor and the schedule's gi_schedule descriptor;
// direction, frontier representations, load balancing and blocking are the
// runtime's (libgraphit_amd).
#pragma once

#include <functional>

#include "graphit/graphit_types.h"

namespace graphit {

struct tracking {
    const void *data;  // the VertexData whose writes enqueue
};

using vertex_udf = std::function<void(Vertex)>;
using edge_udf = std::function<void(Vertex, Vertex)>;
using weighted_edge_udf = std::function<void(Vertex, Vertex, dyn_var<int>)>;
using filter_udf = std::function<dyn_var<int>(Vertex)>;

void vertexset_apply(GraphT &g, vertex_udf f);
void vertexset_apply(VertexSubset &s, vertex_udf f);
void vertexset_apply(GraphT &g, void (*f)(Vertex));
void vertexset_apply(VertexSubset &s, void (*f)(Vertex));

class edgeset_apply {
  public:
    explicit edgeset_apply(Schedule &s) : sched_(&s) {}
    edgeset_apply &from(VertexSubset &f) {
        from_ = &f;
        return *this;
    }
    edgeset_apply &to(filter_udf f) {
        to_ = std::move(f);
        return *this;
    }
    void apply(GraphT &g, void (*udf)(Vertex, Vertex));
    void apply(GraphT &g, void (*udf)(Vertex, Vertex, dyn_var<int>));
    template <class T>
    void apply_modified(GraphT &g, VertexSubset &out, VertexData<T> &tracked, void (*udf)(Vertex, Vertex),
                        bool allow_dupes = true) {
        run(g, &out, &tracked, wrap(udf), nullptr, allow_dupes);
    }
    template <class T>
    void apply_modified(GraphT &g, VertexSubset &out, VertexData<T> &tracked,
                        void (*udf)(Vertex, Vertex, dyn_var<int>), bool allow_dupes = true) {
        run(g, &out, &tracked, wrap(udf), nullptr, allow_dupes);
    }
    void apply_priority(GraphT &g, void (*udf)(Vertex, Vertex, dyn_var<int>), PriorityQueue &pq);

  private:
    Schedule *sched_;
    VertexSubset *from_ = nullptr;
    filter_udf to_;
    static weighted_edge_udf wrap(void (*udf)(Vertex, Vertex));
    static weighted_edge_udf wrap(void (*udf)(Vertex, Vertex, dyn_var<int>));
    void run(GraphT &g, VertexSubset *out, const void *tracked, weighted_edge_udf udf, PriorityQueue *pq,
             bool allow_dupes);
};

// fuse_kernel (operators.cpp:651-672): the region is staged as it stands;
// its loop runs as host-driven launches of the generated kernels.
void fuse_kernel(bool to_fuse, std::function<void()> fn);

}  // namespace graphit
