// StereoSession (reference include/apg_sgm/session.hpp:17-110).
// The session types share one declaration header; this path is the
// reference's, so code written against it compiles unchanged.
#pragma once

#include "apg_sgm/pipeline.hpp"
