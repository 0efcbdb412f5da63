// run_graphit_pipeline (src/pipeline/graphit.cpp:124-151 of the reference):
// prints the staged program as one HIP translation unit for gfx950 that
// includes the easy_graphit_amd runtime header gi_gen.h: the program's
// globals, the device functors of its UDFs (with the reference's atomic
// patch-up: a needs_atomics store guarded by `x == c` becomes a
// compare-and-swap, graphit.cpp:55-122), and the staged function with C
// linkage.
#pragma once

#include <ostream>

#include "blocks/c_code_generator.h"
#include "blocks/rce.h"
#include "builder/builder_context.h"
#include "builder/dyn_var.h"
#include "graphit/operators.h"
#include "graphit/runtime.h"
#include "pipeline/extract_cuda.h"

namespace pipeline {
void run_graphit_pipeline(block::block::Ptr ast, std::ostream &oss);
}  // namespace pipeline
