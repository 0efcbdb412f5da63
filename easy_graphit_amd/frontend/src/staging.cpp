// The staging engine: passes, forks, loop recovery (see builder/dyn_var.h).
//
// extract_body runs the staged function once per path: the first run with
// no forced branch outcome stops at its first dyn_var condition (a fork);
// the function is re-run with that outcome forced true and forced false, and
// so on. Every committed statement and every condition carries a tag: the
// return addresses of the call stack at that point, plus the indices of
// library loops (static_iteration). When a tag repeats within one run, the
// run has closed a loop: it records a jump back to the first occurrence and
// stops. The tree of runs (shared prefixes, one if per fork, jumps at the
// ends) is then turned into while loops: a sequence ending in an `if` whose
// taken branch always ends by jumping back to a statement p of the same
// sequence is `while (true) { statements from p; if (!cond) break; branch }`
// followed by the other branch (`while (cond) { branch }` when p is the `if`).
#include <execinfo.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "builder/dyn_var.h"

namespace builder {

static std::vector<context *> g_stack;

context *current() {
    if (g_stack.empty()) throw std::logic_error("staged code ran outside extract_function_ast");
    return g_stack.back();
}
void push_context(context *c) { g_stack.push_back(c); }
void pop_context() { g_stack.pop_back(); }

std::vector<void *> context::tag() const {
    void *frames[128];
    int n = backtrace(frames, 128);
    std::vector<void *> t(frames, frames + n);
    for (int s : static_scope) t.push_back(reinterpret_cast<void *>(static_cast<intptr_t>(s) + 1));
    return t;
}

void context::commit(block::ptr stmt) {
    if (!next_annotation.empty()) {
        stmt->annotation = next_annotation;
        next_annotation.clear();
    }
    if (!captures.empty()) {
        captures.back()->push_back(stmt);
        return;
    }
    std::vector<void *> t = tag();
    auto it = seen.find(t);
    if (it != seen.end()) throw loop_signal{it->second};
    const int pos = (int)trace.size();
    stmt->id = pos;
    seen.emplace(std::move(t), pos);
    trace.push_back(stmt);
}

bool context::branch(block::ptr cond) {
    if (!captures.empty()) throw std::logic_error("a staged condition inside a library-built statement");
    std::vector<void *> t = tag();
    auto it = seen.find(t);
    if (it != seen.end()) throw loop_signal{it->second};
    const int pos = (int)trace.size();
    block::ptr placeholder = block::make(block::kind::if_stmt, "", {cond});
    placeholder->id = pos;
    seen.emplace(std::move(t), pos);
    trace.push_back(placeholder);
    if (next_decision < decisions.size()) return decisions[next_decision++];
    throw fork_signal{cond, pos};
}

std::string float_literal(double v, bool single) {
    char buf[64];
    if (std::isinf(v)) return v > 0 ? (single ? "__builtin_inff()" : "__builtin_inf()")
                                     : (single ? "(-__builtin_inff())" : "(-__builtin_inf())");
    snprintf(buf, sizeof buf, single ? "%.9g" : "%.17g", v);
    std::string s = buf;
    if (s.find_first_of(".en") == std::string::npos) s += ".0";
    if (single) s += "f";
    return s;
}

builder::builder(const char *s) {
    std::string q = "\"";
    for (const char *c = s; *c; ++c) {
        if (*c == '"' || *c == '\\') q += '\\';
        q += *c;
    }
    e = block::make(block::kind::lit, q + "\"");
}

builder::~builder() noexcept(false) {
    if (pending && e && std::uncaught_exceptions() == 0) {
        pending = false;
        current()->commit(block::make(block::kind::expr_stmt, "", {e}));
    }
}

builder builder::assign(const builder &rhs) const {
    current()->commit(block::make(block::kind::expr_stmt, "", {block::make(block::kind::assign, "", {take(), rhs.take()})}));
    return builder(e);
}

builder call(const std::string &callee, std::vector<builder> args) {
    std::vector<block::ptr> k;
    for (auto &a : args) k.push_back(a.take());
    return builder(block::make(block::kind::call, callee, std::move(k)), true);
}

builder raw(const std::string &text) { return builder(block::make(block::kind::raw, text)); }
builder addr_of(const builder &x) { return builder(block::make(block::kind::addr_of, "", {x.take()})); }
builder deref(const builder &x) { return builder(block::make(block::kind::deref, "", {x.take()})); }

void annotate(const std::string &a) { current()->next_annotation = a; }

std::vector<block::ptr> capture(const std::function<void()> &fn) {
    std::vector<block::ptr> out;
    context *c = current();
    c->captures.push_back(&out);
    try {
        fn();
    } catch (...) {
        c->captures.pop_back();
        throw;
    }
    c->captures.pop_back();
    return out;
}

namespace {

constexpr long kMaxPasses = 200000;

struct explorer {
    const std::function<void()> &fn;
    context &ctx;
    long passes = 0;

    std::vector<block::ptr> explore(const std::vector<bool> &decisions, int start) {
        if (++passes > kMaxPasses) throw std::runtime_error("staging: too many re-runs (unbounded branching?)");
        ctx.reset_pass(decisions);
        enum { complete, forked, looped } how = complete;
        block::ptr cond;
        int at = -1;
        try {
            fn();
        } catch (fork_signal &f) {
            how = forked;
            cond = f.cond;
            at = f.at;
        } catch (loop_signal &l) {
            how = looped;
            at = l.target;
        }
        std::vector<block::ptr> trace = ctx.trace;  // the recursion below resets the context
        std::vector<block::ptr> seq;
        const int end = how == forked ? at : (int)trace.size();
        for (int i = start; i < end; ++i) seq.push_back(trace[i]);
        if (how == looped) {
            block::ptr g = block::make(block::kind::goto_stmt);
            g->id = at;
            seq.push_back(g);
        } else if (how == forked) {
            block::ptr s = block::make(block::kind::if_stmt, "", {cond});
            s->id = at;
            std::vector<bool> d = decisions;
            d.push_back(true);
            s->then_s = explore(d, at + 1);
            d.back() = false;
            s->else_s = explore(d, at + 1);
            seq.push_back(s);
        }
        return seq;
    }
};

// Jump targets reached at the ends of a sequence (through trailing ifs).
void tail_gotos(const std::vector<block::ptr> &seq, std::vector<int> &out, bool &other) {
    if (seq.empty()) {
        other = true;
        return;
    }
    const block::ptr &last = seq.back();
    if (last->k == block::kind::goto_stmt) {
        out.push_back(last->id);
    } else if (last->k == block::kind::if_stmt) {
        tail_gotos(last->then_s, out, other);
        tail_gotos(last->else_s, out, other);
    } else {
        other = true;
    }
}

// Removes the trailing jumps to `target` (each path then falls off the end).
void strip_tail(std::vector<block::ptr> &seq, int target) {
    if (seq.empty()) return;
    block::ptr last = seq.back();
    if (last->k == block::kind::goto_stmt && last->id == target) {
        seq.pop_back();
    } else if (last->k == block::kind::if_stmt) {
        strip_tail(last->then_s, target);
        strip_tail(last->else_s, target);
    }
}

bool mentions_goto(const std::vector<block::ptr> &seq, int target) {
    for (auto &s : seq) {
        if (s->k == block::kind::goto_stmt && s->id == target) return true;
        if (mentions_goto(s->then_s, target) || mentions_goto(s->else_s, target)) return true;
    }
    return false;
}

void recover(std::vector<block::ptr> &seq) {
    for (auto &s : seq) {
        if (s->k == block::kind::if_stmt) {
            recover(s->then_s);
            recover(s->else_s);
        } else if (s->k == block::kind::while_stmt) {
            recover(s->then_s);
        }
    }
    for (;;) {
        if (seq.empty() || seq.back()->k != block::kind::if_stmt) return;
        block::ptr ifs = seq.back();
        std::vector<int> targets;
        bool other = false;
        tail_gotos(ifs->then_s, targets, other);
        if (other || targets.empty()) return;
        const int label = targets[0];
        for (int t : targets)
            if (t != label) return;
        int p = -1;
        for (int i = 0; i < (int)seq.size(); ++i)
            if (seq[i]->id == label) p = i;
        if (p < 0) return;  // the loop head is further out
        std::vector<block::ptr> body = ifs->then_s;
        strip_tail(body, label);
        if (mentions_goto(body, label)) throw std::runtime_error("staging: a loop with an early continue");
        std::vector<block::ptr> rest = ifs->else_s;
        block::ptr loop = block::make(block::kind::while_stmt);
        if (p == (int)seq.size() - 1) {  // while (cond) { body }
            loop->kids = ifs->kids;
            loop->then_s = body;
        } else {  // while (true) { head; if (!cond) break; body }
            loop->kids = {block::make(block::kind::lit, "true")};
            for (int i = p; i < (int)seq.size() - 1; ++i) loop->then_s.push_back(seq[i]);
            block::ptr exit = block::make(block::kind::if_stmt, "",
                                          {block::make(block::kind::unary, "!", ifs->kids)});
            exit->then_s = {block::make(block::kind::break_stmt)};
            loop->then_s.push_back(exit);
            for (auto &b : body) loop->then_s.push_back(b);
        }
        loop->id = seq[p]->id;
        seq.resize(p);
        seq.push_back(loop);
        for (auto &r : rest) seq.push_back(r);
    }
}

void check_no_goto(const std::vector<block::ptr> &seq) {
    for (auto &s : seq) {
        if (s->k == block::kind::goto_stmt) throw std::runtime_error("staging: control flow that is not a loop");
        check_no_goto(s->then_s);
        check_no_goto(s->else_s);
    }
}

}  // namespace

std::vector<block::ptr> extract_body(const std::function<void()> &fn) {
    context &ctx = *current();
    explorer ex{fn, ctx};
    std::vector<block::ptr> body = ex.explore({}, 0);
    recover(body);
    check_no_goto(body);
    ctx.reset_pass({});
    return body;
}

}  // namespace builder
