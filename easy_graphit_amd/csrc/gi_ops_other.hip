// Instantiations of the load-balanced traversal for the SSSP, CC, PageRank
// and BC operators.
#include "gi_lb.h"

namespace gi {

int run_other_family(int32_t op, const gi_op_args &a, const ApplyPlan &plan) {
    switch (op) {
        case GI_OP_SSSP_RELAX: {
            OpSsspRelax o{(int32_t *)a.data0, a.i0};
            return run_plan(o, plan);
        }
        case GI_OP_CC_HOOK: {
            OpCcHook o{(int32_t *)a.data0, (int32_t *)a.data1};
            return run_plan(o, plan);
        }
        case GI_OP_PR_PULL: {
            OpPrAdd o{(float *)a.data0, (const float *)a.data1};
            return run_plan(o, plan);
        }
        case GI_OP_BC_FORWARD: {
            OpBcForward o{(double *)a.data0, (const uint8_t *)a.data1};
            return run_plan(o, plan);
        }
        case GI_OP_BC_BACKWARD: {
            OpBcBackward o{(float *)a.data0, (const uint8_t *)a.data1};
            return run_plan(o, plan);
        }
        default:
            return fail(GI_ERR_INVALID, "unknown operator %d", op);
    }
}

}  // namespace gi
