// GraphIt's GPU scheduling language as the apps spell it
// (include/graphit/schedule.h:12-139 of the reference): SimpleGPUSchedule
// with its direction / frontier / deduplication / load-balancing /
// blocking / fusion knobs, HybridGPUSchedule with a run-time threshold, and
// the s_isa / s_to casts. The enums keep the reference's order, which is the
// order of the gi_schedule fields of include/graphit_amd.h, so a schedule
// maps field for field onto the runtime's descriptor.
#pragma once

#include <string>

#include "builder/dyn_var.h"
#include "builder/static_var.h"

using builder::as_member_of;
using builder::dyn_var;

namespace graphit {

struct Schedule {
    virtual ~Schedule() = default;
};

template <typename T>
bool s_isa(Schedule *s) {
    return dynamic_cast<T *>(s) != nullptr;
}
template <typename T>
T *s_to(Schedule *s) {
    return dynamic_cast<T *>(s);
}

struct SimpleGPUSchedule : public Schedule {
    inline static int default_max_cta = 0;    // 0: the runtime sizes a persistent grid
    inline static int default_cta_size = 256;

    enum class direction_type { PUSH, PULL };
    enum class pull_frontier_rep_type { BITMAP, BOOLMAP };
    enum class frontier_creation_type { SPARSE, BITMAP, BOOLMAP };
    enum class deduplication_type { DISABLED, ENABLED };
    enum class deduplication_strategy_type { FUSED, UNFUSED };
    enum class load_balancing_type { VERTEX_BASED, TWC, TWCE, WM, CM, STRICT, EDGE_ONLY };
    enum class edge_blocking_type { BLOCKED, UNBLOCKED };
    enum class kernel_fusion_type { DISABLED, ENABLED };

    direction_type direction = direction_type::PUSH;
    pull_frontier_rep_type pull_frontier_rep = pull_frontier_rep_type::BOOLMAP;
    frontier_creation_type frontier_creation = frontier_creation_type::SPARSE;
    deduplication_type deduplication = deduplication_type::DISABLED;
    deduplication_strategy_type deduplication_strategy = deduplication_strategy_type::FUSED;
    load_balancing_type load_balancing = load_balancing_type::VERTEX_BASED;
    edge_blocking_type edge_blocking = edge_blocking_type::UNBLOCKED;
    int block_size = -1;
    kernel_fusion_type kernel_fusion = kernel_fusion_type::DISABLED;
    int max_cta;
    int cta_size;

    SimpleGPUSchedule() : max_cta(default_max_cta), cta_size(default_cta_size) {}

    void configDirection(direction_type dir, pull_frontier_rep_type pf = pull_frontier_rep_type::BOOLMAP) {
        direction = dir;
        if (direction == direction_type::PULL) pull_frontier_rep = pf;
    }
    void configFrontierCreation(frontier_creation_type fc) { frontier_creation = fc; }
    void configDeduplication(deduplication_type dedup) { deduplication = dedup; }
    void configLoadBalancing(load_balancing_type lb, edge_blocking_type eb = edge_blocking_type::UNBLOCKED,
                             int bs = -1) {
        load_balancing = lb;
        if (lb == load_balancing_type::EDGE_ONLY) {
            edge_blocking = eb;
            block_size = bs;
        }
    }
    void configKernelFusion(kernel_fusion_type kf) { kernel_fusion = kf; }

    // the runtime's gi_schedule initializer, field for field
    std::string descriptor() const;
};

struct HybridGPUSchedule : public Schedule {
    Schedule *s1;
    Schedule *s2;
    // a run-time threshold (bindThreshold) or a generation-time one
    bool bound = false;
    block::ptr threshold;  // the bound dyn_var
    float static_threshold = 0.f;

    HybridGPUSchedule(Schedule &a, Schedule &b) : s1(&a), s2(&b) {}
    void configThreshold(float t) {
        bound = false;
        static_threshold = t;
    }
    void bindThreshold(dyn_var<float> &t) {
        bound = true;
        threshold = t.take();
    }
};

}  // namespace graphit
