// SgmOptimizer and can_use_u16_sgm_accumulator (reference include/apg_sgm/sgm_optimizer.hpp:9-61).
// The stage classes share one declaration header; this path is the
// reference's, so code written against it compiles unchanged.
#pragma once

#include "apg_sgm/stages.hpp"
