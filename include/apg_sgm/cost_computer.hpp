// CostComputer (reference include/apg_sgm/cost_computer.hpp:10-32).
// The stage classes share one declaration header; this path is the
// reference's, so code written against it compiles unchanged.
#pragma once

#include "apg_sgm/stages.hpp"
