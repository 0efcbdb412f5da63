// GraphIt operators, runtime calls and the HIP pipeline of the front end.
//
// Operators stage each UDF in device context (a nested extraction with the
// UDF's vertex arguments as parameters arg0, arg1 and the edge weight arg2)
// and register the result as a device functor; the host program gets one
// call of the gi_gen runtime per operator. run_graphit_pipeline prints the
// program's globals (one struct G_t, copied by value into every functor),
// the functors, and the staged function with C linkage.
#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

#include "builder/builder_context.h"
#include "graphit/operators.h"
#include "graphit/runtime.h"
#include "pipeline/graphit.h"

namespace graphit {

context_type current_context = context_type::HOST;
tracking *current_tracking = nullptr;

const char graph_t_name[] = "gi_gen::GraphT";
const char vertexsubset_t_name[] = "gi_frontier*";
const char prioqueue_t_name[] = "gi_gen::PrioQueue";
const char frontierlist_t_name[] = "gi_frontier_list*";

void note_tracked_write(const void *data, const builder::builder &index) {
    (void)index;
    if (current_context == context_type::DEVICE && current_tracking && current_tracking->data == data)
        builder::current()->commit(block::make(block::kind::raw, "enq = true;"));
}

template <>
bool VertexDataIndex<int>::is_tracked() const {
    return current_tracking && current_tracking->data == owner;
}
template <>
bool VertexDataIndex<float>::is_tracked() const {
    return current_tracking && current_tracking->data == owner;
}
template <>
bool VertexDataIndex<double>::is_tracked() const {
    return current_tracking && current_tracking->data == owner;
}
template <>
bool VertexDataIndex<unsigned char>::is_tracked() const {
    return current_tracking && current_tracking->data == owner;
}

void VertexSubset::addVertex(const builder::builder &v) {
    (void)builder::call("gi_gen::add_vertex", {builder::builder(ref), v});
}
builder::builder VertexSubset::size() { return builder::call("gi_gen::frontier_size", {builder::builder(ref)}); }
void FrontierList::insert(VertexSubset &v) {
    (void)builder::call("gi_gen::frontier_list_insert", {builder::builder(ref), builder::builder(v.take())});
}
void FrontierList::retrieve(VertexSubset &v) {
    (void)builder::call("gi_gen::frontier_list_retrieve", {builder::builder(ref), builder::builder(v.take())});
}
graph_query_result GraphT::transposed_fn::operator()() const {
    return graph_query_result{builder::builder(builder::call("gi_gen::transposed", {builder::builder(self->ref)}).take())};
}
graph_query_result GraphT::blocked_fn::operator()(const builder::builder &bs) const {
    return graph_query_result{
        builder::builder(builder::call("gi_gen::blocked", {builder::builder(self->ref), bs}).take())};
}

void PriorityQueue::updatePriorityMin(const Vertex &v, const builder::builder &np) {
    if (current_context != context_type::DEVICE || !p)
        throw std::runtime_error("updatePriorityMin outside an apply_priority UDF");
    block::ptr cond = builder::call("gi_gen::update_priority_min",
                                    {builder::builder(p->data.take()),
                                     builder::builder(block::make(block::kind::member, "d_cutoff", {pq.take()})),
                                     builder::builder(v.vid.take()), np})
                          .take();
    block::ptr s = block::make(block::kind::if_stmt, "", {cond});
    s->then_s = {block::make(block::kind::raw, "enq = true;")};
    builder::current()->commit(s);
}

std::string SimpleGPUSchedule::descriptor() const {
    std::ostringstream o;
    o << "gi_gen::schedule(" << (int)direction << ", " << (int)pull_frontier_rep << ", " << (int)frontier_creation
      << ", " << (int)deduplication << ", " << (int)load_balancing << ", " << (int)edge_blocking << ", "
      << block_size << ", " << (int)kernel_fusion << ", " << max_cta << ", " << cta_size << ")";
    return o.str();
}

namespace runtime {
builder::builder load_graph(const builder::builder &f) { return builder::call("gi_gen::load_graph", {f}); }
builder::builder new_vertex_subset(const builder::builder &n) { return builder::call("gi_gen::new_vertex_subset", {n}); }
builder::builder new_frontier_list(const builder::builder &n) { return builder::call("gi_gen::new_frontier_list", {n}); }
builder::builder start_timer() { return builder::call("gi_gen::start_timer", {}); }
builder::builder stop_timer() { return builder::call("gi_gen::stop_timer", {}); }
builder::builder print_time(const builder::builder &t) { return builder::call("gi_gen::print_time", {t}); }
}  // namespace runtime

}  // namespace graphit

// ================================================================ printing
namespace pipeline {
namespace {

struct printer {
    bool device = false;
    std::string expr(const block::ptr &e) const {
        using block::kind;
        switch (e->k) {
            case kind::var:
                return e->global ? std::string(device ? "g." : "G.") + e->text : e->text;
            case kind::lit:
            case kind::raw:
                return e->text;
            case kind::binary:
                return "(" + expr(e->kids[0]) + " " + e->text + " " + expr(e->kids[1]) + ")";
            case kind::unary:
                return "(" + e->text + expr(e->kids[0]) + ")";
            case kind::call: {
                std::string s = e->text + "(";
                for (size_t i = 0; i < e->kids.size(); ++i) s += (i ? ", " : "") + expr(e->kids[i]);
                return s + ")";
            }
            case kind::member:
                return expr(e->kids[0]) + "." + e->text;
            case kind::index:
                return expr(e->kids[0]) + "[" + expr(e->kids[1]) + "]";
            case kind::assign:
                return expr(e->kids[0]) + " = " + expr(e->kids[1]);
            case kind::addr_of:
                return "(&" + expr(e->kids[0]) + ")";
            case kind::deref:
                return "(*" + expr(e->kids[0]) + ")";
            default:
                throw std::runtime_error("printer: a statement in expression position");
        }
    }
    void stmts(const std::vector<block::ptr> &ss, std::ostream &o, int ind) const {
        for (auto &s : ss) stmt(s, o, ind);
    }
    void stmt(const block::ptr &s, std::ostream &o, int ind) const {
        using block::kind;
        const std::string pad(ind * 4, ' ');
        switch (s->k) {
            case kind::decl:
                o << pad << s->type << " " << s->text;
                if (!s->kids.empty()) o << " = " << expr(s->kids[0]);
                o << ";\n";
                break;
            case kind::expr_stmt:
                o << pad << expr(s->kids[0]) << ";\n";
                break;
            case kind::raw:
                o << pad << s->text << "\n";
                break;
            case kind::if_stmt:
                o << pad << "if (" << expr(s->kids[0]) << ") {\n";
                stmts(s->then_s, o, ind + 1);
                if (!s->else_s.empty()) {
                    o << pad << "} else {\n";
                    stmts(s->else_s, o, ind + 1);
                }
                o << pad << "}\n";
                break;
            case kind::while_stmt:
                o << pad << "while (" << expr(s->kids[0]) << ") {\n";
                stmts(s->then_s, o, ind + 1);
                o << pad << "}\n";
                break;
            case kind::break_stmt:
                o << pad << "break;\n";
                break;
            case kind::return_stmt:
                o << pad << "return";
                if (!s->kids.empty()) o << " " << expr(s->kids[0]);
                o << ";\n";
                break;
            default:
                throw std::runtime_error("printer: unexpected statement");
        }
    }
};

void gather_globals(const block::ptr &n, std::map<std::string, std::string> &out) {
    if (!n) return;
    if (n->k == block::kind::var && n->global) out.emplace(n->text, n->type);
    for (auto &k : n->kids) gather_globals(k, out);
    for (auto &k : n->then_s) gather_globals(k, out);
    for (auto &k : n->else_s) gather_globals(k, out);
}

}  // namespace

// ---------------------------------------------------------------- functors
struct functor {
    std::string name;
    std::string text;  // the struct, with `NAME` for its name
    std::vector<block::ptr> roots;
};
static std::vector<functor> g_functors;

static std::string register_functor(const std::string &text, const std::vector<block::ptr> &roots, char kind) {
    for (auto &f : g_functors)
        if (f.text == text) return f.name;
    functor f;
    f.name = std::string(kind == 'v' ? "VOp" : "EOp") + std::to_string(g_functors.size());
    f.text = text;
    f.roots = roots;
    g_functors.push_back(f);
    return f.name;
}

// The reference's atomic patch-up (graphit.cpp:55-122): a needs_atomics
// store `x = v` directly inside `if (x == c)` becomes `if (cas(&x, c, v))`.
static void patch_guarded(std::vector<block::ptr> &ss, const printer &pr) {
    for (auto &s : ss) {
        patch_guarded(s->then_s, pr);
        patch_guarded(s->else_s, pr);
        if (s->k != block::kind::if_stmt) continue;
        const block::ptr &c = s->kids[0];
        if (c->k != block::kind::binary || c->text != "==") continue;
        for (size_t i = 0; i < s->then_s.size(); ++i) {
            const block::ptr &t = s->then_s[i];
            if (t->annotation != NEEDS_ATOMICS_ANNOTATION || t->k != block::kind::expr_stmt) continue;
            const block::ptr &a = t->kids[0];
            if (a->k != block::kind::assign || pr.expr(a->kids[0]) != pr.expr(c->kids[0])) continue;
            s->kids[0] = block::make(block::kind::call, "gi_gen::cas",
                                     {block::make(block::kind::addr_of, "", {a->kids[0]}), c->kids[1], a->kids[1]});
            s->then_s.erase(s->then_s.begin() + i);
            break;
        }
    }
}

// The same patch-up across the operator's filter: the filter `return x[v]
// == c` guards the UDF's store `x[dst] = e` (the apply runs after the filter
// passed), which becomes `if (!cas(&x[dst], c, e)) return false;`.
static void patch_filtered(std::vector<block::ptr> &apply, const std::vector<block::ptr> &filter, const printer &pr) {
    if (filter.empty() || filter.back()->k != block::kind::return_stmt || filter.back()->kids.empty()) return;
    block::ptr r = filter.back()->kids[0];
    for (int hops = 0; r->k == block::kind::var && hops < 16; ++hops) {  // through the copies of the value
        block::ptr init;
        for (auto &s : filter)
            if (s->k == block::kind::decl && s->text == r->text && !s->kids.empty()) init = s->kids[0];
        if (!init) break;
        r = init;
    }
    if (r->k != block::kind::binary || r->text != "==") return;
    const std::string guarded = pr.expr(r->kids[0]);  // spelled with the filter's vertex, arg0
    for (auto &s : apply) {
        if (s->annotation != NEEDS_ATOMICS_ANNOTATION || s->k != block::kind::expr_stmt) continue;
        const block::ptr &a = s->kids[0];
        if (a->k != block::kind::assign) continue;
        // the store's target with the destination (arg1) spelled as the filter's vertex
        std::string target = pr.expr(a->kids[0]);
        const std::string dst = "[arg1]";
        const size_t at = target.rfind(dst);
        if (at == std::string::npos || at + dst.size() != target.size()) continue;
        target.replace(at, dst.size(), "[arg0]");
        if (target != guarded) continue;
        std::ostringstream o;
        o << "if (!gi_gen::cas(&" << pr.expr(a->kids[0]) << ", " << pr.expr(r->kids[1]) << ", "
          << pr.expr(a->kids[1]) << ")) return false;";
        s = block::make(block::kind::raw, o.str());
    }
}

}  // namespace pipeline

namespace graphit {

namespace {

struct device_scope {
    context_type saved;
    tracking *saved_track;
    device_scope(tracking *t) : saved(current_context), saved_track(current_tracking) {
        current_context = context_type::DEVICE;
        current_tracking = t;
    }
    ~device_scope() {
        current_context = saved;
        current_tracking = saved_track;
    }
};

block::ptr param(const char *name) {
    block::ptr v = block::make(block::kind::var, name);
    v->type = "int";
    return v;
}

std::vector<block::ptr> stage_nested(const std::function<void()> &fn) {
    builder::context ctx;
    builder::push_context(&ctx);
    std::vector<block::ptr> body;
    try {
        body = builder::extract_body(fn);
    } catch (...) {
        builder::pop_context();
        throw;
    }
    builder::pop_context();
    return body;
}

std::string vertex_functor(std::function<void(Vertex)> f) {
    std::vector<block::ptr> body;
    {
        device_scope d(nullptr);
        body = stage_nested([&]() { f(Vertex(builder::builder(param("arg0")), Vertex::access_type::INDEPENDENT)); });
    }
    pipeline::printer pr;
    pr.device = true;
    pipeline::patch_guarded(body, pr);
    std::ostringstream o;
    o << "struct NAME {\n    G_t g;\n    __device__ void operator()(int arg0) const {\n";
    pr.stmts(body, o, 2);
    o << "    }\n};\n";
    return pipeline::register_functor(o.str(), body, 'v');
}

std::string edge_functor(const weighted_edge_udf &udf, const filter_udf &to, const void *tracked) {
    tracking t{tracked};
    std::vector<block::ptr> apply, filter;
    {
        device_scope d(tracked ? &t : nullptr);
        apply = stage_nested([&]() {
            udf(Vertex(builder::builder(param("arg0")), Vertex::access_type::SHARED),
                Vertex(builder::builder(param("arg1")), Vertex::access_type::SHARED),
                dyn_var<int>(builder::alias, param("arg2")));
        });
        if (to) {
            device_scope d2(nullptr);
            filter = stage_nested([&]() {
                dyn_var<int> r = to(Vertex(builder::builder(param("arg0")), Vertex::access_type::SHARED));
                builder::current()->commit(block::make(block::kind::return_stmt, "", {r.take()}));
            });
        }
    }
    pipeline::printer pr;
    pr.device = true;
    pipeline::patch_guarded(apply, pr);
    pipeline::patch_guarded(filter, pr);
    pipeline::patch_filtered(apply, filter, pr);
    std::ostringstream o;
    o << "struct NAME {\n    G_t g;\n    __device__ bool filter(int arg0) const {\n";
    if (to) pr.stmts(filter, o, 2);
    else o << "        return true;\n";
    o << "    }\n    __device__ bool apply(int arg0, int arg1, int arg2) const {\n        bool enq = false;\n";
    pr.stmts(apply, o, 2);
    o << "        return enq;\n    }\n};\n";
    std::vector<block::ptr> roots = apply;
    roots.insert(roots.end(), filter.begin(), filter.end());
    return pipeline::register_functor(o.str(), roots, 'e');
}

std::string schedule_text(Schedule *s, bool dedup, std::string *threshold_expr) {
    auto simple = [&](Schedule *x) {
        SimpleGPUSchedule *p = s_to<SimpleGPUSchedule>(x);
        if (!p) throw std::runtime_error("a hybrid schedule nested in a hybrid schedule");
        SimpleGPUSchedule c = *p;
        if (dedup) c.deduplication = SimpleGPUSchedule::deduplication_type::ENABLED;
        return c.descriptor();
    };
    if (HybridGPUSchedule *h = s_to<HybridGPUSchedule>(s)) {
        pipeline::printer pr;
        std::string t = h->bound ? pr.expr(h->threshold) : builder::float_literal(h->static_threshold, true);
        if (threshold_expr) *threshold_expr = t;
        return "gi_gen::hybrid(" + simple(h->s1) + ", " + simple(h->s2) + ", " + t + ")";
    }
    return simple(s);
}

void emit_call(const std::string &callee, std::vector<builder::builder> args) {
    (void)builder::call(callee, std::move(args));
}

}  // namespace

void vertexset_apply(GraphT &g, vertex_udf f) {
    const std::string op = vertex_functor(f);
    emit_call("gi_gen::vertexset_apply", {builder::builder(g.take()), builder::raw(op + "{G}")});
}
void vertexset_apply(VertexSubset &s, vertex_udf f) {
    const std::string op = vertex_functor(f);
    emit_call("gi_gen::vertexset_apply", {builder::builder(s.take()), builder::raw(op + "{G}")});
}
void vertexset_apply(GraphT &g, void (*f)(Vertex)) { vertexset_apply(g, vertex_udf(f)); }
void vertexset_apply(VertexSubset &s, void (*f)(Vertex)) { vertexset_apply(s, vertex_udf(f)); }

weighted_edge_udf edgeset_apply::wrap(void (*udf)(Vertex, Vertex)) {
    return [udf](Vertex s, Vertex d, dyn_var<int>) { udf(s, d); };
}
weighted_edge_udf edgeset_apply::wrap(void (*udf)(Vertex, Vertex, dyn_var<int>)) {
    return [udf](Vertex s, Vertex d, dyn_var<int> w) { udf(s, d, w); };
}

void edgeset_apply::apply(GraphT &g, void (*udf)(Vertex, Vertex)) { run(g, nullptr, nullptr, wrap(udf), nullptr, true); }
void edgeset_apply::apply(GraphT &g, void (*udf)(Vertex, Vertex, dyn_var<int>)) {
    run(g, nullptr, nullptr, wrap(udf), nullptr, true);
}
void edgeset_apply::apply_priority(GraphT &g, void (*udf)(Vertex, Vertex, dyn_var<int>), PriorityQueue &pq) {
    run(g, nullptr, nullptr, wrap(udf), &pq, true);
}

void edgeset_apply::run(GraphT &g, VertexSubset *out, const void *tracked, weighted_edge_udf udf,
                        PriorityQueue *pq, bool allow_dupes) {
    const std::string op = edge_functor(udf, to_, tracked);
    const std::string sched = schedule_text(sched_, !allow_dupes, nullptr);
    builder::builder from = from_ ? builder::builder(from_->take()) : builder::raw("nullptr");
    if (pq) {
        if (!from_) throw std::runtime_error("apply_priority needs a frontier (from)");
        emit_call("gi_gen::edgeset_apply_priority", {builder::builder(g.take()), builder::raw(sched), from,
                                                     builder::builder(pq->pq.take()), builder::raw(op + "{G}")});
    } else {
        builder::builder to = out ? builder::builder(out->take()) : builder::raw("nullptr");
        emit_call("gi_gen::edgeset_apply",
                  {builder::builder(g.take()), builder::raw(sched), from, to, builder::raw(op + "{G}")});
    }
}

void fuse_kernel(bool, std::function<void()> fn) { fn(); }

}  // namespace graphit

// ================================================================ pipeline
namespace pipeline {

void run_graphit_pipeline(block::block::Ptr f, std::ostream &oss) {
    std::map<std::string, std::string> globals;
    for (auto &s : f->body) gather_globals(s, globals);
    for (auto &fn : g_functors)
        for (auto &r : fn.roots) gather_globals(r, globals);
    oss << "// Generated by the easy_graphit_amd front end from the staged GraphIt program " << f->name << ".\n"
        << "// HIP for gfx950 over the easy_graphit_amd runtime (gi_gen.h, libgraphit_amd.so).\n"
        << "#include \"gi_gen.h\"\n\n";
    oss << "struct G_t {\n";
    for (auto &kv : globals) oss << "    " << kv.second << " " << kv.first << "{};\n";
    oss << "};\nstatic G_t G;\n\n";
    for (auto &fn : g_functors) {
        std::string t = fn.text;
        t.replace(t.find("NAME"), 4, fn.name);
        oss << t << "\n";
    }
    printer pr;
    std::ostringstream body;
    pr.stmts(f->body, body, 2);
    oss << "static std::string g_error;\n\n";
    oss << "// " << f->name << ": the staged function; 0 on success, -1 on a runtime error\n";
    oss << "extern \"C\" int " << f->name << "(";
    for (size_t i = 0; i < f->params.size(); ++i) oss << (i ? ", " : "") << f->params[i].first << " " << f->params[i].second;
    oss << ") {\n    try {\n" << body.str() << "    } catch (const std::exception &e) {\n"
        << "        g_error = e.what();\n        return -1;\n    }\n    return 0;\n}\n\n";
    oss << "extern \"C\" const char *gi_gen_error() { return g_error.c_str(); }\n\n";
    oss << "// a global of the program: device arrays by value, other handles by address\n";
    oss << "extern \"C\" void *gi_gen_global(const char *name) {\n";
    for (auto &kv : globals) {
        const bool ptr = !kv.second.empty() && kv.second.back() == '*';
        oss << "    if (!strcmp(name, \"" << kv.first << "\")) return (void *)" << (ptr ? "G." : "&G.") << kv.first
            << ";\n";
    }
    oss << "    return nullptr;\n}\n";
}

}  // namespace pipeline
