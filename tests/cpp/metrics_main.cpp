// Test driver of the accuracy metrics (apg::evaluate_stereo): reads a
// disparity PFM, a ground-truth PFM, the left image and an optional visibility
// mask, and prints the metric columns of the benchmark CSV as one JSON object,
// formatted as the benchmark app writes them (fixed, 4 decimals).
//   metrics_main disp.pfm gt.pfm left.pgm [vis.pgm]
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "apg_sgm/image.hpp"
#include "apg_sgm/metrics.hpp"

namespace {
std::ostringstream js;
bool first = true;
template <typename T>
void put(const char* k, T v) {
    std::ostringstream s;
    s << std::fixed << std::setprecision(4) << v;
    js << (first ? "{" : ", ") << "\"" << k << "\": \"" << s.str() << "\"";
    first = false;
}
void region(const std::string& p, const char* total_name, const apg::RegionMetrics& r) {
    put(total_name, r.total_pixels);
    put((p + "_valid_est_pixels").c_str(), r.valid_est_pixels);
    put((p + "_valid_ratio").c_str(), r.valid_ratio);
    put((p + "_epe").c_str(), r.epe);
    put((p + "_bad05").c_str(), r.bad_0_5);
    put((p + "_bad10").c_str(), r.bad_1_0);
    put((p + "_bad20").c_str(), r.bad_2_0);
    put((p + "_bad30").c_str(), r.bad_3_0);
    put((p + "_d1").c_str(), r.d1);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: metrics_main disp.pfm gt.pfm left.pgm [vis.pgm]\n");
        return 2;
    }
    apg::Image32f disp, gt;
    apg::Image8 left, vis;
    if (!apg::load_pfm(argv[1], disp) || !apg::load_pfm(argv[2], gt) || !apg::load_image(argv[3], left)) {
        std::fprintf(stderr, "cannot read inputs\n");
        return 1;
    }
    const bool has_vis = argc > 4 && apg::load_image(argv[4], vis);
    const apg::StereoMetrics m =
        apg::evaluate_stereo(disp, gt, nullptr, nullptr, nullptr, has_vis ? &vis : nullptr, 1e5f, nullptr, &left);
    put("epe", m.epe);
    put("bad_0_5", m.bad_0_5);
    put("bad_1_0", m.bad_1_0);
    put("bad_2_0", m.bad_2_0);
    put("bad_3_0", m.bad_3_0);
    put("kitti_d1_all", m.kitti_d1_all);
    put("kitti_d1_noc", m.kitti_d1_noc);
    put("valid_ratio", m.valid_ratio);
    put("lr_fail_ratio", m.lr_fail_ratio);
    put("edge_epe", m.edge_epe);
    put("nonedge_epe", m.nonedge_epe);
    region("all", "all_gt_pixels", m.reg_all);
    region("noc", "noc_pixels", m.reg_noc);
    region("occ", "occ_pixels", m.reg_occ);
    region("bnd", "bnd_pixels", m.reg_boundary);
    region("nonbnd", "nonbnd_pixels", m.reg_non_boundary);
    region("lowtex", "lowtex_pixels", m.reg_low_texture);
    region("hightex", "hightex_pixels", m.reg_high_texture);
    std::cout << js.str() << "}\n";
    return 0;
}
