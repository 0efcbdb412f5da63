// GraphIt's staged data types as the apps use them (the reference's
// include/graphit/graphit_types.h:37-444), lowered onto the easy_graphit_amd
// runtime: GraphT is a gi_gen::GraphT (a device CSR handle with its sizes
// and arrays), VertexSubset a gi_frontier*, FrontierList a gi_frontier_list*,
// PriorityQueue a gi_gen::PrioQueue, VertexData<T> a device array T*.
//
// Write semantics of VertexData (graphit_types.h:243-302): in device code
// (an operator's UDF) `x[i] = e` is a plain store, annotated needs_atomics
// unless the index is the vertex the thread owns, `x[i].min(e)` a writeMin
// and `x[i] += e` an atomic add (plain read-modify-write on an owned
// vertex); a write into the vertex data an apply_modified tracks adds the
// written vertex to the output frontier. In host code an element access is
// a device-to-host or host-to-device copy of that element.
#pragma once

#include <string>

#include "builder/dyn_var.h"
#include "graphit/schedule.h"

namespace graphit {

enum class context_type { HOST = 0, DEVICE = 1 };
extern context_type current_context;

#define NEEDS_ATOMICS_ANNOTATION "needs_atomics"

extern const char graph_t_name[];
extern const char vertexsubset_t_name[];
extern const char prioqueue_t_name[];
extern const char frontierlist_t_name[];

// Common base of the staged handle types: a variable (or alias) of the
// generated program.
template <const char *N>
struct handle : public builder::staged_base {
    block::ptr ref;
    handle() { declare(nullptr); }
    explicit handle(const char *global_name) {
        ref = block::make(block::kind::var, global_name);
        ref->global = true;
        ref->type = N;
    }
    handle(const builder::builder &b) { declare(b.take()); }
    handle(const handle &o) { declare(o.ref); }
    handle(builder::alias_t, block::ptr lvalue) : ref(std::move(lvalue)) {}
    block::ptr take() const { return ref; }
    operator builder::builder() const { return builder::builder(ref); }
    void assign(block::ptr rhs) {
        builder::current()->commit(block::make(block::kind::expr_stmt, "",
                                               {block::make(block::kind::assign, "", {ref, rhs})}));
    }

  private:
    void declare(block::ptr init) {
        builder::context *c = builder::current();
        ref = block::make(block::kind::var, c->fresh_var());
        ref->type = N;
        block::ptr d = block::make(block::kind::decl, ref->text);
        d->type = N;
        if (init) d->kids.push_back(init);
        c->commit(d);
    }
};

inline builder::builder to_b(const builder::builder &b) { return builder::to_b(b); }
// any operand as a staged expression (arithmetic values, dyn_var, builder,
// and this namespace's staged types through their to_b)
template <class U>
builder::builder stage(const U &x) {
    using builder::to_b;
    return to_b(x);
}

// edges.get_transposed_graph()[0] / edges.get_blocked_graph(bs)[0]: the
// reference's calls return the graph object indexed once
struct graph_query_result {
    builder::builder value;
    builder::builder operator[](int) const { return builder::builder(value.take()); }
    operator builder::builder() const { return builder::builder(value.take()); }
};

struct GraphT : public handle<graph_t_name> {
    using handle::handle;
    GraphT() : handle() { bind(); }
    GraphT(const GraphT &o) : handle(o) { bind(); }
    GraphT(const builder::builder &b) : handle(b) { bind(); }
    explicit GraphT(const char *name) : handle(name) { bind(); }
    GraphT(const graph_query_result &r) : handle(builder::builder(r.value.take())) { bind(); }
    GraphT &operator=(const builder::builder &b) {
        assign(b.take());
        return *this;
    }
    GraphT &operator=(const GraphT &o) {
        assign(o.ref);
        return *this;
    }
    GraphT *addr() { return this; }

    dyn_var<int> num_vertices{builder::alias, nullptr};
    dyn_var<int> num_edges{builder::alias, nullptr};
    dyn_var<int *> d_src_offsets{builder::alias, nullptr};
    dyn_var<int *> d_edge_dst{builder::alias, nullptr};
    dyn_var<int *> d_edge_weight{builder::alias, nullptr};
    dyn_var<int *> out_degrees{builder::alias, nullptr};

    struct transposed_fn {
        GraphT *self;
        graph_query_result operator()() const;
    } get_transposed_graph{this};
    struct blocked_fn {
        GraphT *self;
        graph_query_result operator()(const builder::builder &bs) const;
    } get_blocked_graph{this};

  private:
    void bind() {
        num_vertices.ref = block::make(block::kind::member, "num_vertices", {ref});
        num_edges.ref = block::make(block::kind::member, "num_edges", {ref});
        d_src_offsets.ref = block::make(block::kind::member, "d_src_offsets", {ref});
        d_edge_dst.ref = block::make(block::kind::member, "d_edge_dst", {ref});
        d_edge_weight.ref = block::make(block::kind::member, "d_edge_weight", {ref});
        out_degrees.ref = block::make(block::kind::member, "out_degrees", {ref});
    }
};

struct VertexSubset : public handle<vertexsubset_t_name> {
    using handle::handle;
    VertexSubset(const VertexSubset &o) : handle(o) {}
    VertexSubset(const builder::builder &b) : handle(b) {}
    VertexSubset &operator=(const builder::builder &b) {
        assign(b.take());
        return *this;
    }
    VertexSubset &operator=(const VertexSubset &o) {
        assign(o.ref);
        return *this;
    }
    VertexSubset *addr() { return this; }
    void addVertex(const builder::builder &v);
    builder::builder size();
};

struct FrontierList : public handle<frontierlist_t_name> {
    using handle::handle;
    FrontierList(const FrontierList &o) : handle(o) {}
    FrontierList(const builder::builder &b) : handle(b) {}
    FrontierList *addr() { return this; }
    void insert(VertexSubset &v);
    void retrieve(VertexSubset &v);
};

// A vertex argument of a UDF and how the running thread may access it.
struct Vertex : public builder::staged_base {
    enum class access_type { INDEPENDENT = 0, SHARED = 1, CONSTANT = 2 };
    builder::builder vid;
    access_type current_access;
    Vertex(const dyn_var<int> &v) : vid(v.take()), current_access(access_type::SHARED) {}
    Vertex(const builder::builder &v, access_type a) : vid(v.take()), current_access(a) {}
    Vertex(const int &x) : vid(x), current_access(access_type::CONSTANT) {}
    Vertex(const Vertex &o) : vid(o.vid.take()), current_access(o.current_access) {}
    operator dyn_var<int>() const { return dyn_var<int>(builder::builder(vid.take())); }
    operator builder::builder() const { return builder::builder(vid.take()); }
};
inline builder::builder to_b(const Vertex &v) { return builder::builder(v.vid.take()); }

struct tracking;  // apply_modified's output of the current UDF (operators.h)
extern tracking *current_tracking;
// the tracked write of `data` at `index` happened: the UDF enqueues
void note_tracked_write(const void *data, const builder::builder &index);

template <typename T>
struct VertexData;

template <typename T>
struct VertexDataIndex : public builder::staged_base {
    const VertexData<T> *owner;
    builder::builder index;
    Vertex::access_type current_access;

    VertexDataIndex(const VertexData<T> *o, const builder::builder &i, Vertex::access_type a)
        : owner(o), index(i.take()), current_access(a) {}
    VertexDataIndex(const VertexDataIndex &o)
        : owner(o.owner), index(o.index.take()), current_access(o.current_access) {}

    block::ptr element() const {
        return block::make(block::kind::index, "", {owner->data.take(), index.take()});
    }
    builder::builder value() const {
        if (current_context == context_type::DEVICE) return builder::builder(element());
        return builder::call("gi_gen::load", {builder::builder(owner->data.take()), builder::builder(index.take())});
    }
    operator dyn_var<T>() const { return dyn_var<T>(value()); }
    operator builder::builder() const { return value(); }

    void operator=(const builder::builder &e) {
        if (current_context == context_type::DEVICE) {
            if (current_access != Vertex::access_type::INDEPENDENT) builder::annotate(NEEDS_ATOMICS_ANNOTATION);
            builder::builder(element()).assign(e);
            note_tracked_write(owner, index);
        } else {
            (void)builder::call("gi_gen::store", {builder::builder(owner->data.take()),
                                                  builder::builder(index.take()), e});
        }
    }
    void operator=(const VertexDataIndex &o) { (*this) = o.value(); }
    template <class U, std::enable_if_t<std::is_arithmetic<U>::value, int> = 0>
    void operator=(U v) {
        (*this) = builder::builder(v);
    }
    void operator=(const Vertex &v) { (*this) = builder::builder(v.vid.take()); }
    template <class U>
    void operator=(const dyn_var<U> &v) {
        (*this) = builder::builder(v.take());
    }

    // writeMin (graphit_types.h:262-280)
    void min(const builder::builder &e) {
        if (current_access == Vertex::access_type::INDEPENDENT) {
            block::ptr cond = (builder::builder(element()) > builder::builder(e.take())).take();
            block::ptr rhs = e.take();
            auto body = builder::capture([&]() {
                builder::builder(element()).assign(builder::builder(rhs));
                note_tracked_write(owner, index);
            });
            block::ptr s = block::make(block::kind::if_stmt, "", {cond});
            s->then_s = body;
            builder::current()->commit(s);
        } else {
            builder::builder w = builder::call("gi_gen::write_min",
                                               {builder::addr_of(builder::builder(element())), e});
            if (is_tracked()) {
                block::ptr cond = w.take();
                auto body = builder::capture([&]() { note_tracked_write(owner, index); });
                block::ptr s = block::make(block::kind::if_stmt, "", {cond});
                s->then_s = body;
                builder::current()->commit(s);
            }
        }
    }
    // writeSum (graphit_types.h:281-295)
    void operator+=(const builder::builder &e) {
        if (current_access == Vertex::access_type::INDEPENDENT) {
            builder::builder(element()).assign(builder::builder(element()) + e);
        } else {
            (void)builder::call("gi_gen::write_sum", {builder::addr_of(builder::builder(element())), e});
        }
        note_tracked_write(owner, index);
    }
    template <class U>
    void operator+=(const U &e) {
        (*this) += stage(e);
    }
    bool is_tracked() const;
};

template <class T>
builder::builder to_b(const VertexDataIndex<T> &x) {
    return x.value();
}

template <typename T>
struct VertexData {
    dyn_var<T *> data;
    explicit VertexData(std::string name) : data(strdup_name(name)) {}

    VertexDataIndex<T> operator[](const dyn_var<int> &i) const {
        return VertexDataIndex<T>(this, builder::builder(i.take()), Vertex::access_type::SHARED);
    }
    VertexDataIndex<T> operator[](const builder::builder &i) const {
        return VertexDataIndex<T>(this, i, Vertex::access_type::SHARED);
    }
    VertexDataIndex<T> operator[](const Vertex &v) const {
        return VertexDataIndex<T>(this, builder::builder(v.vid.take()), v.current_access);
    }
    VertexDataIndex<T> operator[](const int &i) const {
        return VertexDataIndex<T>(this, builder::builder(i), Vertex::access_type::CONSTANT);
    }
    void allocate(const builder::builder &size) {
        data = builder::call("gi_gen::alloc<" + builder::type_name<T>::get() + ">", {size});
    }

  private:
    static const char *strdup_name(const std::string &s) {
        char *p = new char[s.size() + 1];
        for (size_t i = 0; i <= s.size(); ++i) p[i] = s.c_str()[i];
        return p;  // lives as long as the program: the name of a global
    }
};

struct PriorityQueue {
    handle<prioqueue_t_name> pq;
    const VertexData<int> *p = nullptr;
    explicit PriorityQueue(std::string name) : pq(strdup_pq(name)) {}

    // init (graphit_types.h:414-418): the queue over p_, delta-wide windows
    void init(GraphT &edges, VertexData<int> &p_, const builder::builder &init_p, const builder::builder &delta,
              const builder::builder &src) {
        p = &p_;
        (void)builder::call("gi_gen::pq_init", {builder::builder(pq.take()), builder::builder(edges.take()),
                                                builder::builder(p_.data.take()), init_p, delta, src});
    }
    // updatePriorityMin (graphit_types.h:426-430): writeMin, and enqueue
    // below the current window's cutoff
    void updatePriorityMin(const Vertex &v, const builder::builder &np);
    builder::builder finished() { return builder::call("gi_gen::pq_finished", {builder::builder(pq.take())}); }
    builder::builder dequeue() { return builder::call("gi_gen::pq_dequeue", {builder::builder(pq.take())}); }

  private:
    static const char *strdup_pq(const std::string &s) {
        char *q = new char[s.size() + 1];
        for (size_t i = 0; i <= s.size(); ++i) q[i] = s.c_str()[i];
        return q;
    }
};

}  // namespace graphit
