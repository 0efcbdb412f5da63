// Runtime calls of GraphIt programs (include/graphit/runtime.h of the
// reference), each a call of the gi_gen runtime in the generated code.
#pragma once

#include "graphit/graphit_types.h"

namespace graphit {
namespace runtime {

builder::builder load_graph(const builder::builder &file_name);
builder::builder new_vertex_subset(const builder::builder &num_vertices);
builder::builder new_frontier_list(const builder::builder &num_vertices);
builder::builder start_timer();
builder::builder stop_timer();
builder::builder print_time(const builder::builder &t);

}  // namespace runtime
}  // namespace graphit
