// static_var: a generation-time value (BuildIt's builder::static_var). The
// GraphIt programs of this front end keep generation-time state in plain
// C++ values, so this is the plain value with the same spelling.
#pragma once

#include "builder/dyn_var.h"

namespace builder {
template <class T>
using static_var = T;
}  // namespace builder
