// StereoWorkspace (reference include/apg_sgm/workspace.hpp:10-75).
// The session types share one declaration header; this path is the
// reference's, so code written against it compiles unchanged.
#pragma once

#include "apg_sgm/pipeline.hpp"
