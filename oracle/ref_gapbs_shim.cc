// Driver that exposes the reference's own GAPBS graph construction (compiled
// from the reference sources where they lie; nothing is copied) through a C
// ABI, so the oracle's CSR builder can be checked against it.
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, loaded only by tests.
//
// Reference entry points used:
//   BuilderBase::MakeGraphFromEL  runtime/infra_gapbs/builder.h:205-226
//   BuilderBase::SquishGraph      runtime/infra_gapbs/builder.h:165-179
//   Generator::GenerateEL         runtime/infra_gapbs/generator.h:109-120
#define IGNORE_JULIENNE_TYPES
#include <cstdint>
#include <string>
#include "benchmark.h"

namespace {
// CLBase only carries options; this subclass sets the symmetrize flag the
// builder reads (command_line.h:39, builder.h:48-50).
struct ShimCL : public CLBase {
  explicit ShimCL(bool sym) : CLBase(std::string("")) { symmetrize_ = sym; }
};
}  // namespace

extern "C" int64_t ref_build_csr(int32_t num_vertices, int64_t num_pairs,
                                 const int32_t *src, const int32_t *dst, int symmetrize,
                                 int32_t *offsets, int32_t *nbrs) {
  ShimCL cli(symmetrize != 0);
  Builder b(cli);
  pvector<EdgePair<NodeID, NodeID>> el(num_pairs);
  for (int64_t i = 0; i < num_pairs; ++i) el[i] = EdgePair<NodeID, NodeID>(src[i], dst[i]);
  CSRGraph<NodeID> g = b.MakeGraphFromEL(el);
  CSRGraph<NodeID> sq = b.SquishGraph(g);
  int64_t out = 0;
  offsets[0] = 0;
  for (int32_t v = 0; v < num_vertices; ++v) {
    if (v < sq.num_nodes())
      for (NodeID u : sq.out_neigh(v)) nbrs[out++] = u;
    offsets[v + 1] = (int32_t)out;
  }
  return out;
}

extern "C" int64_t ref_rmat_edges(int scale, int degree, int32_t *src, int32_t *dst) {
  Generator<NodeID> gen(scale, degree);
  pvector<EdgePair<NodeID, NodeID>> el = gen.GenerateEL(false);
  int64_t n = (int64_t)el.size();
  for (int64_t i = 0; i < n; ++i) { src[i] = el[i].u; dst[i] = el[i].v; }
  return n;
}
