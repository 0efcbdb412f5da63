// Instantiations of the load-balanced traversal for the BFS operators.
#include "gi_lb.h"

namespace gi {

int run_bfs_family(int32_t op, const gi_op_args &a, const ApplyPlan &plan) {
    switch (op) {
        case GI_OP_BFS_PARENT: {
            OpBfsParent o{(int32_t *)a.data0};
            return run_plan(o, plan);
        }
        case GI_OP_BFS_DEPTH: {
            OpBfsDepth o{(int32_t *)a.data0, a.i0 + 1};
            return run_plan(o, plan);
        }
        default:
            return fail(GI_ERR_INVALID, "not a BFS operator: %d", op);
    }
}

int run_bfs_visit(const OpBfsVisit &op, const ApplyPlan &plan) { return run_plan(op, plan); }

}  // namespace gi
