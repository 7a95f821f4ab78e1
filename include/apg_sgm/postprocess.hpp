// PostProcessor (reference include/apg_sgm/postprocess.hpp:8-13).
// The stage classes share one declaration header; this path is the
// reference's, so code written against it compiles unchanged.
#pragma once

#include "apg_sgm/stages.hpp"
