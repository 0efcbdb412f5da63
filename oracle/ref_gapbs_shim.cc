// Driver that exposes the reference's own GAPBS graph construction (compiled
// from the reference sources where they lie; nothing is copied) through a C
// ABI, so the oracle's CSR builder can be checked against it.
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, loaded only by tests.
//
// Reference entry points used:
//   BuilderBase::MakeGraphFromEL  runtime/infra_gapbs/builder.h:205-226
//   BuilderBase::SquishGraph      runtime/infra_gapbs/builder.h:165-179
//   BuilderBase::MakeGraph        runtime/infra_gapbs/builder.h:228-246 (Reader, reader.h)
//   Generator::GenerateEL         runtime/infra_gapbs/generator.h:109-120
//   WriterBase::WriteGraph        runtime/infra_gapbs/writer.h
#define IGNORE_JULIENNE_TYPES
#include <cstdint>
#include <string>
#include "benchmark.h"
#include "writer.h"

namespace {
// CLBase only carries options; this subclass sets the symmetrize flag the
// builder reads (command_line.h:39, builder.h:48-50).
struct ShimCL : public CLBase {
  explicit ShimCL(bool sym) : CLBase(std::string("")) { symmetrize_ = sym; }
};

// A file to read, with the symmetrize flag (the reference's load_graph uses
// CLBase(filename), i.e. symmetrize off: graph.h:273).
struct FileCL : public CLBase {
  FileCL(const char *path, bool sym) : CLBase(std::string(path)) { symmetrize_ = sym; }
};

bool is_sg(const std::string &p) { return p.size() >= 3 && p.compare(p.size() - 3, 3, ".sg") == 0; }

// The last graph ref_read_graph built.
CSRGraph<NodeID, WNode> g_wgraph;
CSRGraph<NodeID> g_ugraph;
bool g_weighted = false;

template <class G>
int64_t out_entries(const G &g) {
  int64_t n = 0;
  for (NodeID v = 0; v < g.num_nodes(); ++v) n += g.out_degree(v);
  return n;
}
}  // namespace

// Reads `path` as the reference's load_graph does (graph.h:273-276): GAPBS's
// WeightedBuilder for every format but .sg, which only the unweighted Builder
// accepts (reader.h:332-352). Returns V; *edges is the out-CSR's entry count.
extern "C" int64_t ref_read_graph(const char *path, int symmetrize, int64_t *edges, int *directed) {
  FileCL cli(path, symmetrize != 0);
  if (is_sg(path)) {
    Builder b(cli);
    g_ugraph = b.MakeGraph();
    g_weighted = false;
    *edges = out_entries(g_ugraph);
    *directed = g_ugraph.directed() ? 1 : 0;
    return g_ugraph.num_nodes();
  }
  WeightedBuilder b(cli);
  g_wgraph = b.MakeGraph();
  g_weighted = true;
  *edges = out_entries(g_wgraph);
  *directed = g_wgraph.directed() ? 1 : 0;
  return g_wgraph.num_nodes();
}

// The out-CSR of the graph ref_read_graph built (weights left alone for .sg).
extern "C" void ref_read_graph_copy(int32_t *offsets, int32_t *nbrs, int32_t *weights) {
  int64_t out = 0;
  offsets[0] = 0;
  if (g_weighted) {
    for (NodeID v = 0; v < g_wgraph.num_nodes(); ++v) {
      for (WNode x : g_wgraph.out_neigh(v)) {
        nbrs[out] = x.v;
        weights[out] = x.w;
        ++out;
      }
      offsets[v + 1] = (int32_t)out;
    }
  } else {
    for (NodeID v = 0; v < g_ugraph.num_nodes(); ++v) {
      for (NodeID x : g_ugraph.out_neigh(v)) nbrs[out++] = x;
      offsets[v + 1] = (int32_t)out;
    }
  }
}

// Builds a graph from a text file with GAPBS and writes it serialized
// (.wsg when weighted, else .sg) with GAPBS's writer.
extern "C" void ref_write_serialized(const char *in_path, int symmetrize, int weighted, const char *out_path) {
  FileCL cli(in_path, symmetrize != 0);
  if (weighted) {
    WeightedBuilder b(cli);
    CSRGraph<NodeID, WNode> g = b.MakeGraph();
    WeightedWriter w(g);
    w.WriteGraph(out_path, true);
  } else {
    Builder b(cli);
    CSRGraph<NodeID> g = b.MakeGraph();
    Writer w(g);
    w.WriteGraph(out_path, true);
  }
}

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
