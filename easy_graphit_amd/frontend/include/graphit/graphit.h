#pragma once

#include "graphit/graphit_types.h"
#include "graphit/operators.h"
#include "graphit/runtime.h"
#include "graphit/schedule.h"
