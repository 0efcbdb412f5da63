// The AST printer lives in the pipeline (pipeline/graphit.h prints HIP);
// this header keeps the include spelling of staged programs.
#pragma once

#include "builder/dyn_var.h"
