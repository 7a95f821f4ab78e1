// The benchmark's ablation configurations A-G, with the reference's
// interface and settings (reference include/apg_sgm/ablation.hpp:8-26,
// src/ablation.cpp:8-150): each adds one stage to the previous one.
#pragma once

#include <string>
#include <vector>

#include "apg_sgm/config.hpp"

namespace apg {

enum class AblationId { A_Census4, B_MultiCost4, C_Cross4, D_Adaptive4, E_Prior4, F_Refine4, G_HQ8 };

PipelineConfig make_ablation_config(AblationId id, int max_disparity);
std::string ablation_id_to_string(AblationId id);
std::string ablation_id_description(AblationId id);
bool parse_ablation_id(const std::string& str, AblationId& out);
std::vector<AblationId> all_ablation_ids();

}  // namespace apg
