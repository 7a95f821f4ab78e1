// Ablation configurations of the benchmark (settings of reference
// src/ablation.cpp:8-150).
#include "apg_sgm/ablation.hpp"

#include <cctype>

namespace apg {

namespace {
struct AblationRow {
    AblationId id;
    const char* name;
    const char* description;
    bool multi_cost, cross, adaptive, prior, refine, paths8;
};
// cost terms, Cross, adaptive P2, prior, refiner (3 iterations), 8 paths
constexpr AblationRow kRows[] = {
    {AblationId::A_Census4, "A", "Census + 4SGM", false, false, false, false, false, false},
    {AblationId::B_MultiCost4, "B", "Census+AD+Grad + 4SGM", true, false, false, false, false, false},
    {AblationId::C_Cross4, "C", "Census+AD+Grad + Cross + 4SGM", true, true, false, false, false, false},
    {AblationId::D_Adaptive4, "D", "Census+AD+Grad + Cross + AdaptiveP2 + 4SGM", true, true, true, false, false, false},
    {AblationId::E_Prior4, "E", "AdaptiveP2 + Prior + 4SGM", true, true, true, true, false, false},
    {AblationId::F_Refine4, "F", "Prior + Refiner + 4SGM", true, true, true, true, true, false},
    {AblationId::G_HQ8, "G", "Prior + Refiner + 8SGM (HQ)", true, true, true, true, true, true},
};
const AblationRow& row(AblationId id) {
    for (const AblationRow& r : kRows)
        if (r.id == id) return r;
    return kRows[0];
}
}  // namespace

PipelineConfig make_ablation_config(AblationId id, int max_disparity) {
    const AblationRow& r = row(id);
    PipelineConfig c;
    c.min_disparity = 0;
    c.max_disparity = max_disparity;
    c.post.subpixel = true;
    c.post.lr_check = true;
    c.post.lr_max_diff = 1;
    c.post.median_radius = 1;
    c.post.max_fill_gap = 32;
    c.cost.use_ad = c.cost.use_grad = r.multi_cost;
    c.aggregation.enable = r.cross;
    if (r.cross) c.aggregation.iterations = 1;
    c.sgm.adaptive_penalties = r.adaptive;
    c.sgm.paths = r.paths8 ? PathType::Path8 : PathType::Path4;
    c.prior.enable = r.prior;
    c.refine.enable = r.refine;
    if (r.refine) c.refine.iterations = 3;
    return c;
}

std::string ablation_id_to_string(AblationId id) { return row(id).name; }
std::string ablation_id_description(AblationId id) { return row(id).description; }

bool parse_ablation_id(const std::string& str, AblationId& out) {
    std::string s = str;
    for (char& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    static const char* const kLong[] = {"A_CENSUS4", "B_MULTICOST4", "C_CROSS4", "D_ADAPTIVE4",
                                        "E_PRIOR4", "F_REFINE4", "G_HQ8"};
    for (int i = 0; i < 7; ++i)
        if (s == kRows[i].name || s == kLong[i]) {
            out = kRows[i].id;
            return true;
        }
    return false;
}

std::vector<AblationId> all_ablation_ids() {
    std::vector<AblationId> v;
    for (const AblationRow& r : kRows) v.push_back(r.id);
    return v;
}

}  // namespace apg
