// edgeset_apply dispatcher (src/graphit/operators.cpp:513-650 and the
// apply_modified / apply_priority wrappers of include/graphit/operators.h:
// 73-173): resolves direction, graph view (transposed / blocked), enumerator,
// input representation and output frontier, then launches the load-balanced
// kernel of the schedule.
#include "gi_lb.h"

namespace gi {

static bool op_may_duplicate(int32_t op) {
    return op == GI_OP_SSSP_RELAX || op == GI_OP_BC_FORWARD;
}

int validate_schedule(const gi_schedule *s) {
    if (!s) return fail(GI_ERR_INVALID, "null schedule");
    if (s->direction != GI_PUSH && s->direction != GI_PULL)
        return fail(GI_ERR_INVALID, "bad direction %d", s->direction);
    if (s->load_balancing < GI_LB_VERTEX_BASED || s->load_balancing > GI_LB_EDGE_ONLY)
        return fail(GI_ERR_INVALID, "bad load balancing %d", s->load_balancing);
    if (s->frontier_creation < GI_FC_SPARSE || s->frontier_creation > GI_FC_BOOLMAP)
        return fail(GI_ERR_INVALID, "bad frontier creation %d", s->frontier_creation);
    if (s->pull_frontier_rep != GI_PULL_BITMAP && s->pull_frontier_rep != GI_PULL_BOOLMAP)
        return fail(GI_ERR_INVALID, "bad pull frontier representation %d", s->pull_frontier_rep);
    if (s->cta_size < 64 || s->cta_size > 1024 || s->cta_size % 64)
        return fail(GI_ERR_INVALID, "cta_size %d must be a multiple of 64 in [64,1024]", s->cta_size);
    if (s->load_balancing == GI_LB_EDGE_ONLY && s->edge_blocking == GI_BLOCKED && s->block_size <= 0)
        return fail(GI_ERR_INVALID, "EDGE_ONLY + BLOCKED needs block_size > 0");
    return GI_OK;
}

int plan_grid(const gi_schedule *s) {
    return s->max_cta > 0 ? s->max_cta : persistent_grid(s->cta_size);
}

int edgeset_plan(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to, bool may_duplicate,
                 hipStream_t st, hipEvent_t *events, ApplyPlan *out, int32_t *out_fmt) {
    GI_TRY(validate_schedule(s));
    if (from && from->V != g->V) return fail(GI_ERR_INVALID, "frontier size does not match graph");
    if (to && to->V != g->V) return fail(GI_ERR_INVALID, "output frontier size does not match graph");
    ApplyPlan plan{};
    plan.pull = s->direction == GI_PULL;
    gi_graph *tg = g;
    if (plan.pull) GI_TRY(gi_graph_transpose(g, &tg));
    bool edge_only = s->load_balancing == GI_LB_EDGE_ONLY;
    if (edge_only && s->edge_blocking == GI_BLOCKED) GI_TRY(gi_graph_blocked(tg, s->block_size, &tg));
    if (edge_only) GI_TRY(ensure_coo(tg, st));
    plan.g = tg->view();
    plan.en.V = g->V;
    plan.in = make_member(nullptr);
    plan.lp.stream = st;
    plan.lp.block = s->cta_size;
    plan.lp.grid = plan_grid(s);
    plan.lp.lb = s->load_balancing;
    plan.lp.graph = tg;
    plan.lp.src_member = make_member(nullptr);
    plan.lp.ev0 = events ? events[0] : nullptr;
    plan.lp.ev1 = events ? events[1] : nullptr;
    if (!plan.pull) {
        if (from) {
            if (edge_only) {
                GI_TRY(frontier_convert(from, GI_FMT_BITMAP, st));
                plan.lp.src_member = make_member(from);
            } else {
                GI_TRY(frontier_convert(from, GI_FMT_SPARSE, st));
                plan.en.q = from->q[from->side];
                plan.en.count = from->count_in();
            }
        }
    } else if (from) {
        int32_t rep = s->pull_frontier_rep == GI_PULL_BITMAP ? GI_FMT_BITMAP : GI_FMT_BOOLMAP;
        GI_TRY(frontier_convert(from, rep, st));
        plan.in = make_member(from);
        if (!edge_only) {
            // reverse frontier into graph-owned scratch (the reference
            // overwrites the caller's queue, operators.cpp:603-609)
            GI_TRY(ensure_scratch(g, scratch_total_len(g->V)));
            plan.reverse = true;
            plan.rev_count = g->d_scratch + scratch_lb_len(g->V);
            plan.rev_q = plan.rev_count + 64;
        }
    }
    int32_t fmt = GI_FMT_SPARSE;
    if (to) {
        fmt = s->frontier_creation;
        GI_TRY(frontier_clear_out(to, fmt, st));
        plan.out = make_sink(to, fmt, may_duplicate || s->deduplication == GI_DEDUP_ENABLED);
    } else {
        plan.out = make_sink(nullptr, 0, false);
    }
    *out = plan;
    *out_fmt = fmt;
    return GI_OK;
}

int edgeset_finish(gi_frontier *to, int32_t fmt, hipStream_t st) {
    if (to) GI_TRY(gi_frontier_swap(to, fmt, st));
    return GI_OK;
}

int edgeset_apply(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to,
                  int32_t op, const gi_op_args &a, hipStream_t st, const OpBfsVisit *visit,
                  hipEvent_t *events) {
    ApplyPlan plan{};
    int32_t fmt = GI_FMT_SPARSE;
    GI_TRY(edgeset_plan(g, s, from, to, op_may_duplicate(op), st, events, &plan, &fmt));
    int rc;
    if (visit) rc = run_bfs_visit(*visit, plan);
    else if (op == GI_OP_BFS_PARENT || op == GI_OP_BFS_DEPTH) rc = run_bfs_family(op, a, plan);
    else rc = run_other_family(op, a, plan);
    if (rc != GI_OK) return rc;
    return edgeset_finish(to, fmt, st);
}

}  // namespace gi

using namespace gi;

extern "C" {

int gi_edgeset_apply(gi_graph *g, const gi_schedule *s, gi_frontier *from, gi_frontier *to,
                     int32_t op, const gi_op_args *args, void *stream) {
    if (!g || !args) return fail(GI_ERR_INVALID, "null argument");
    if (op < GI_OP_BFS_PARENT || op > GI_OP_BC_BACKWARD) return fail(GI_ERR_INVALID, "unknown operator %d", op);
    if (!g->d_off) return fail(GI_ERR_UNSUPPORTED, "pass the base graph, not a blocked view");
    if (op == GI_OP_SSSP_RELAX && !g->d_w) return fail(GI_ERR_INVALID, "SSSP needs edge weights");
    return edgeset_apply(g, s, from, to, op, *args, (hipStream_t)stream, nullptr);
}

int gi_edgeset_apply_hybrid(gi_graph *g, const gi_hybrid_schedule *h, gi_frontier *from,
                            gi_frontier *to, int32_t op, const gi_op_args *args,
                            void *stream) {
    if (!g || !h || !args) return fail(GI_ERR_INVALID, "null argument");
    const gi_schedule *pick = &h->s1;
    if (h->selector == GI_SELECT_S2) {
        pick = &h->s2;
    } else if (h->selector != GI_SELECT_S1 && from) {
        // operators.cpp:526: |from| <= threshold * V runs s1, else s2
        int32_t n = 0;
        GI_TRY(gi_frontier_size(from, &n, stream));
        pick = ((double)n <= (double)h->threshold * (double)g->V) ? &h->s1 : &h->s2;
    }
    return gi_edgeset_apply(g, pick, from, to, op, args, stream);
}

}  // extern "C"
