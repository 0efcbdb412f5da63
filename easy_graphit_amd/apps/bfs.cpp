// apps/bfs.cpp of the reference, written against the C++ mirror API
// (include/graphit_amd.hpp). Same UDF (parent[dst] = src under the
// parent[v] == -1 filter), same schedules: push TWCE, or in "power" mode the
// hybrid of push TWCE and pull bitmap vertex-based (apps/bfs.cpp:66-95).
// Usage: bfs <scale> <edge_factor> <src> [power <threshold>]
// Prints the parent array checksum and the per-trial times.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "graphit_amd.hpp"

using namespace graphit;

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s scale edge_factor src [power threshold]\n", argv[0]);
        return 2;
    }
    int scale = std::atoi(argv[1]), ef = std::atoi(argv[2]), src = std::atoi(argv[3]);
    bool power = argc > 4 && std::string(argv[4]) == "power";
    float t = argc > 5 ? (float)std::atof(argv[5]) : 0.05f;
    try {
        SimpleGPUSchedule s1;
        s1.configDirection(SimpleGPUSchedule::direction_type::PUSH);
        s1.configLoadBalancing(SimpleGPUSchedule::load_balancing_type::TWCE);
        SimpleGPUSchedule s2;
        s2.configDirection(SimpleGPUSchedule::direction_type::PULL,
                           SimpleGPUSchedule::pull_frontier_rep_type::BITMAP);
        s2.configLoadBalancing(SimpleGPUSchedule::load_balancing_type::VERTEX_BASED);
        s2.configFrontierCreation(SimpleGPUSchedule::frontier_creation_type::BITMAP);
        HybridGPUSchedule h1(s1, s2);
        h1.bindThreshold(t);

        GraphT edges = runtime::load_rmat(scale, ef, 27491095, true);
        VertexData<int32_t> parent(edges.num_vertices);
        for (int trial = 0; trial < 4; ++trial) {
            VertexSubset frontier(edges.num_vertices);
            runtime::start_timer();
            gi_op_args reset{};
            reset.data0 = parent.ptr;
            reset.i0 = -1;
            vertexset_apply(edges, GI_VOP_FILL_I32, reset);  // reset(v): parent[v] = -1
            parent.set(src, src);
            frontier.addVertex(src);
            while (frontier.size() != 0) {
                if (power) edgeset_apply(h1).from(frontier).apply_modified(edges, frontier, udf::bfs_parent(parent));
                else edgeset_apply(s1).from(frontier).apply_modified(edges, frontier, udf::bfs_parent(parent));
            }
            float secs = runtime::stop_timer();
            std::printf("trial %d: %f\n", trial, secs);
        }
        std::vector<int32_t> p = parent.to_host();
        long long reached = 0;
        for (int32_t x : p) reached += x >= 0;
        std::printf("reached %lld\n", reached);
        FILE *f = argc > 6 ? std::fopen(argv[6], "wb") : nullptr;
        if (f) {
            std::fwrite(p.data(), sizeof(int32_t), p.size(), f);
            std::fclose(f);
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "graphit error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
