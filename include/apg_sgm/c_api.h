/* Source-compatible include path of the reference (apg_sgm/c_api.h). */
#include "../apg_sgm_c_api.h"
