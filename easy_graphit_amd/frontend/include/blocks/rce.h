// Redundant-copy elimination is left to the HIP compiler (the generated
// code declares one variable per staged value); this header keeps the
// include spelling of staged programs.
#pragma once

#include "builder/dyn_var.h"
