// Annotation names of the staged operators (the reference's
// include/pipeline/extract_cuda.h). Kernel extraction itself does not
// exist here: operators stage their UDFs into device functors directly.
#pragma once

#include <algorithm>
#include <iostream>
#include <vector>

#include "builder/dyn_var.h"

#define CUDA_ANNOTATION_STRING "run_on_device"
#define CUDA_COOP_ANNOTATION_STRING "coop_run_on_device"
#ifndef NEEDS_ATOMICS_ANNOTATION
#define NEEDS_ATOMICS_ANNOTATION "needs_atomics"
#endif
