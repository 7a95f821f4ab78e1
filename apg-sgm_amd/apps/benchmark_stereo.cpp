// apg-benchmark for the MI355X build: the reference's benchmark app
// (reference apps/benchmark_stereo.cpp:30-937) on the device pipeline.  Same
// manifest format, options, ablation configurations A-G, timing (warmup, then
// the median of `repeat` runs), accuracy metrics and CSV schema (header
// :441-470, one row per case and configuration), so its CSVs compare column
// for column with the CPU reference's.
//
// Not in this build: the hierarchical range estimator, the non-legacy subpixel
// and inpainting methods and their telemetry (the options are refused); their
// CSV columns hold what the reference writes when they are off.
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "apg_sgm/ablation.hpp"
#include "apg_sgm/image.hpp"
#include "apg_sgm/metrics.hpp"
#include "apg_sgm/pipeline.hpp"

namespace fs = std::filesystem;

namespace {

struct Case {
    std::string name, left, right, gt;
    int dmax = 128, dmin = 0;
    std::string vis;
};

// "name left right gt [dmax [dmin [vis]]]" per line, '#' comments, paths
// relative to the manifest's directory
std::vector<Case> read_manifest(const std::string& path) {
    std::vector<Case> out;
    std::ifstream f(path);
    if (!f) {
        std::cerr << "Error: cannot open manifest file: " << path << std::endl;
        return out;
    }
    const fs::path base = fs::path(path).parent_path();
    auto rel = [&](std::string& p) {
        if (!p.empty() && fs::path(p).is_relative()) p = (base / p).lexically_normal().string();
    };
    std::string line;
    while (std::getline(f, line)) {
        const size_t b = line.find_first_not_of(" \t\r\n");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream is(line.substr(b));
        Case c;
        if (!(is >> c.name >> c.left >> c.right >> c.gt)) continue;
        is >> c.dmax;
        is >> c.dmin;
        is >> c.vis;
        rel(c.left);
        rel(c.right);
        rel(c.gt);
        rel(c.vis);
        out.push_back(c);
    }
    return out;
}

template <typename P>
bool same_bits(const char* what, const P& a, const P& b) {
    if (a.width() != b.width() || a.height() != b.height()) {
        std::cerr << "Verification FAILED: " << what << " dimensions differ\n";
        return false;
    }
    size_t bad = 0;
    for (int y = 0; y < a.height(); ++y)
        for (int x = 0; x < a.width(); ++x) {
            const auto va = a.at(x, y), vb = b.at(x, y);
            if (std::memcmp(&va, &vb, sizeof(va)) != 0) {
                if (bad < 5) std::cerr << "Mismatch in " << what << " at (" << x << ", " << y << ")\n";
                ++bad;
            }
        }
    if (bad) std::cerr << "Total mismatches in " << what << ": " << bad << "\n";
    return bad == 0;
}

const char* kHeader =
    "backend,threads,effective_threads,warmup,repeat,build_type,"
    "case,ablation_id,ablation_name,cost,cross,p2,prior,refine,paths,w,h,dmax,"
    "epe,bad_0_5,bad_1_0,bad_2_0,bad_3_0,kitti_d1_all,kitti_d1_noc,valid_ratio,lr_fail_ratio,edge_epe,nonedge_epe,"
    "recall_all,recall_matchable,recall_visible,range_vis_eval_px,range_vis_in_px,range_mat_eval_px,range_mat_in_px,"
    "prior_miss_vis_count,prior_miss_edge_pct,prior_miss_nonedge_pct,"
    "prior_supports,support_p05,support_p1,support_p2,support_p1_vis,support_pvis_05,support_pvis_1,support_pvis_2,"
    "support_gt_valid,support_visible,support_correct_05,support_correct_1,support_correct_2,"
    "support_visible_correct_05,support_visible_correct_1,support_visible_correct_2,"
    "support_visible_gt_cells,support_correct_visible_cells,"
    "support_mae,support_mae_vis,grid_recall_1,"
    "mean_search_width,mean_geom_width,geom_reduction_ratio,prior_incremental_reduction,total_reduction_ratio,"
    "cost_bytes,cost32_bytes,peak_bytes,"
    "reliable_ratio,unreliable_ratio,refine_changed_ratio,refine_epe_delta,post_epe_delta,total_epe_delta,"
    "time_total_ms,time_aux_ms,time_prior_ms,time_cost_ms,time_right_wta_ms,time_cross_ms,time_sgm_ms,time_wta_ms,"
    "time_confidence_ms,time_refine_ms,time_post_ms,"
    "all_gt_pixels,all_valid_est_pixels,all_valid_ratio,all_epe,all_bad05,all_bad10,all_bad20,all_bad30,all_d1,"
    "noc_pixels,noc_valid_est_pixels,noc_valid_ratio,noc_epe,noc_bad05,noc_bad10,noc_bad20,noc_bad30,noc_d1,"
    "occ_pixels,occ_valid_est_pixels,occ_valid_ratio,occ_epe,occ_bad05,occ_bad10,occ_bad20,occ_bad30,occ_d1,"
    "bnd_pixels,bnd_valid_est_pixels,bnd_valid_ratio,bnd_epe,bnd_bad05,bnd_bad10,bnd_bad20,bnd_bad30,bnd_d1,"
    "nonbnd_pixels,nonbnd_valid_est_pixels,nonbnd_valid_ratio,nonbnd_epe,nonbnd_bad05,nonbnd_bad10,nonbnd_bad20,"
    "nonbnd_bad30,nonbnd_d1,"
    "lowtex_pixels,lowtex_valid_est_pixels,lowtex_valid_ratio,lowtex_epe,lowtex_bad05,lowtex_bad10,lowtex_bad20,"
    "lowtex_bad30,lowtex_d1,"
    "hightex_pixels,hightex_valid_est_pixels,hightex_valid_ratio,hightex_epe,hightex_bad05,hightex_bad10,"
    "hightex_bad20,hightex_bad30,hightex_d1,"
    "peak_workspace_bytes,retained_workspace_bytes,"
    "avg_states_per_pixel,coarse_valid_ratio,fallback_pixel_ratio,full_range_pixel_ratio,"
    "subpixel_method,subpixel_param,subpixel_candidate_px,subpixel_applied_px,subpixel_applied_ratio,"
    "subpixel_degenerate_count,subpixel_clamped_count,subpixel_clamped_ratio,subpixel_neg_curv_count,"
    "subpixel_mean_abs_delta,subpixel_p50_abs_delta,subpixel_p95_abs_delta,"
    "inpaint_method,inpaint_gap,inpaint_total_holes,inpaint_filled_holes,inpaint_fill_ratio,"
    "inpaint_occ_two_sided,inpaint_occ_left_only,inpaint_occ_right_only,inpaint_occ_unfilled,"
    "inpaint_mism_two_sided,inpaint_mism_left_only,inpaint_mism_right_only,inpaint_mism_unfilled,"
    "inpaint_reclassified_shadow,inpaint_hit_max_gap\n";

void region(std::ostream& o, const apg::RegionMetrics& r) {
    o << r.total_pixels << "," << r.valid_est_pixels << "," << std::setprecision(4) << r.valid_ratio << "," << r.epe
      << "," << r.bad_0_5 << "," << r.bad_1_0 << "," << r.bad_2_0 << "," << r.bad_3_0 << "," << r.d1 << ",";
}

void usage() {
    std::cout << "Usage: apg-benchmark [options]\n"
              << "  --manifest <path>     dataset manifest (required)\n"
              << "  --ablation, --modes <id,...|all>  configurations A-G (default: all)\n"
              << "  --output <path>       CSV file (default: results/ablation.csv)\n"
              << "  --save_disp <dir>     also save disparity maps (PFM + PGM preview)\n"
              << "  --warmup <N>          untimed runs first (default: 2)\n"
              << "  --repeat <N>          timed runs, median reported (default: 5)\n"
              << "  --threads <N>         host threads of the HighQuality host steps (0 = auto)\n"
              << "  --backend <auto|dense|packed>, --packed   volume backend policy\n"
              << "  --verify-backends     check dense and packed backends give identical buffers\n"
              << "  --inpaint-gap <N>     hole-filling gap (default: 32)\n"
              << "  --subpixel-method legacy, --inpaint-method legacy   (the only methods of this build)\n";
}

}  // namespace

int main(int argc, char** argv) {
    std::string manifest, modes = "all", output = "results/ablation.csv", save_dir, backend = "auto";
    int warmup = 2, repeat = 5, threads = 0, inpaint_gap = 32;
    bool verify = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string { return i + 1 < argc ? argv[++i] : std::string(); };
        const bool has = i + 1 < argc;
        if (a == "--manifest" && has) manifest = next();
        else if ((a == "--ablation" || a == "--modes") && has) modes = next();
        else if (a == "--output" && has) output = next();
        else if (a == "--save_disp" && has) save_dir = next();
        else if (a == "--warmup" && has) warmup = std::max(0, std::stoi(next()));
        else if (a == "--repeat" && has) repeat = std::max(1, std::stoi(next()));
        else if (a == "--threads" && has) threads = std::max(0, std::stoi(next()));
        else if (a == "--backend" && has) backend = next();
        else if (a == "--packed") backend = "packed";
        else if (a == "--verify-backends") verify = true;
        else if (a == "--inpaint-gap" && has) inpaint_gap = std::max(1, std::stoi(next()));
        else if ((a == "--subpixel-method" || a == "--inpaint-method") && has) {
            const std::string m = next();
            if (m != "legacy" && m != "A" && m != "a") {
                std::cerr << "Error: " << a << " '" << m << "' is not part of this device build (legacy only)\n";
                return 1;
            }
        } else if (a.rfind("--hierarchical", 0) == 0 || a == "--subpixel-param" || a == "--subpixel-telem" ||
                   a == "--inpaint-telem") {
            std::cerr << "Error: " << a << " is not part of this device build\n";
            return 1;
        } else if (a == "--help" || a == "-h") {
            usage();
            return 0;
        } else {
            std::cerr << "Error: unknown or incomplete argument '" << a << "'. Use --help for usage.\n";
            return 1;
        }
    }
    apg::VolumeBackend vb;
    if (backend == "packed" || backend == "PACKED") vb = apg::VolumeBackend::Packed;
    else if (backend == "dense" || backend == "DENSE") vb = apg::VolumeBackend::Dense;
    else if (backend == "auto" || backend == "AUTO") vb = apg::VolumeBackend::Auto;
    else {
        std::cerr << "Error: unrecognized --backend '" << backend << "'. Options: auto, dense, packed.\n";
        return 1;
    }
    if (manifest.empty()) {
        std::cerr << "Error: --manifest is required. Run with --help for details.\n";
        return 1;
    }
    const std::vector<Case> cases = read_manifest(manifest);
    if (cases.empty()) {
        std::cerr << "Error: no valid test cases found in manifest: " << manifest << "\n";
        return 1;
    }
    std::vector<apg::AblationId> ids;
    if (modes == "all" || modes == "ALL") {
        ids = apg::all_ablation_ids();
    } else {
        std::stringstream ss(modes);
        std::string tok;
        while (std::getline(ss, tok, ',')) {
            tok.erase(0, tok.find_first_not_of(" \t"));
            tok.erase(tok.find_last_not_of(" \t") + 1);
            apg::AblationId id;
            if (tok.empty()) continue;
            if (apg::parse_ablation_id(tok, id)) ids.push_back(id);
            else std::cerr << "Warning: unrecognized ablation ID: " << tok << "\n";
        }
    }
    if (ids.empty()) {
        std::cerr << "Error: no valid ablation configurations selected.\n";
        return 1;
    }
    auto config = [&](apg::AblationId id, const Case& c) {
        apg::PipelineConfig cfg = apg::make_ablation_config(id, c.dmax);
        cfg.min_disparity = c.dmin;
        cfg.max_disparity = c.dmax;
        cfg.volume_backend = vb;
        if (threads > 0) cfg.num_threads = threads;
        cfg.post.max_fill_gap = inpaint_gap;
        return cfg;
    };

    if (verify) {
        bool all = true;
        for (const Case& c : cases) {
            apg::Image8 L, R;
            if (!apg::load_image(c.left, L) || !apg::load_image(c.right, R)) {
                std::cerr << "Error loading input images for case: " << c.name << "\n";
                return 1;
            }
            for (apg::AblationId id : ids) {
                std::cout << "Testing case " << c.name << " [" << apg::ablation_id_to_string(id) << "] ... " << std::flush;
                apg::PipelineConfig cd = config(id, c), cp = cd;
                cd.volume_backend = apg::VolumeBackend::Dense;
                cp.volume_backend = apg::VolumeBackend::Packed;
                apg::PipelineBuffers bd, bp;
                const apg::StereoMatcher md(cd), mp(cp);
                if (!md.compute(L, R, bd, nullptr) || !mp.compute(L, R, bp, nullptr)) {
                    std::cerr << "compute failed on case " << c.name << ": " << md.last_error() << mp.last_error() << "\n";
                    return 1;
                }
                bool ok = same_bits("disparity", bd.disparity, bp.disparity);
                ok &= same_bits("disparity_right", bd.disparity_right, bp.disparity_right);
                ok &= same_bits("d_before_refine", bd.d_before_refine, bp.d_before_refine);
                ok &= same_bits("d_after_refine", bd.d_after_refine, bp.d_after_refine);
                ok &= same_bits("confidence", bd.confidence, bp.confidence);
                ok &= same_bits("reliable_mask", bd.reliable_mask, bp.reliable_mask);
                ok &= same_bits("invalid_reason", bd.invalid_reason, bp.invalid_reason);
                std::cout << (ok ? "PASS (100% bit-exact parity across all buffers)\n" : "FAILED!\n");
                all &= ok;
            }
        }
        if (!all) {
            std::cerr << "\nVerification FAILED on one or more cases!\n";
            return 1;
        }
        std::cout << "\nAll test cases and ablation modes passed bit-exact dual backend parity!\n";
        return 0;
    }

    const fs::path out_path(output);
    if (out_path.has_parent_path()) fs::create_directories(out_path.parent_path());
    if (!save_dir.empty()) fs::create_directories(save_dir);
    std::ofstream csv(output);
    if (!csv) {
        std::cerr << "Error: cannot open output CSV file: " << output << "\n";
        return 1;
    }
    // as the reference's OpenMP default: OMP_NUM_THREADS, else the cores
    int eff_threads = threads;
    if (eff_threads <= 0) {
        const char* env = std::getenv("OMP_NUM_THREADS");
        eff_threads = env && std::atoi(env) > 0 ? std::atoi(env)
                                                : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    }
    csv << kHeader;
    std::cout << "Starting benchmark: " << cases.size() << " cases, " << ids.size() << " ablation configs, "
              << "backend_policy=" << backend << ", warmup=" << warmup << ", repeat=" << repeat << "\n"
              << "Output CSV: " << output << "\n";

    for (const Case& c : cases) {
        std::cout << "\n=== Dataset Case: " << c.name << " (dmax=" << c.dmax << ", dmin=" << c.dmin << ") ===\n";
        apg::Image8 L, R, vis;
        if (!apg::load_image(c.left, L)) {
            std::cerr << "Failed to load left image: " << c.left << "\n";
            continue;
        }
        if (!apg::load_image(c.right, R)) {
            std::cerr << "Failed to load right image: " << c.right << "\n";
            continue;
        }
        apg::Image32f gt;
        const bool has_gt = apg::load_pfm(c.gt, gt);
        if (!has_gt) std::cout << "Note: could not load GT PFM (" << c.gt << "), will report metrics as N/A\n";
        const bool has_vis = !c.vis.empty() && apg::load_image(c.vis, vis);
        const int w = L.width(), h = L.height();

        for (apg::AblationId id : ids) {
            const apg::PipelineConfig cfg = config(id, c);
            const apg::StereoMatcher matcher(cfg);
            apg::PipelineBuffers bufs;
            for (int r = 0; r < warmup; ++r)
                if (!matcher.compute(L, R, bufs, nullptr)) {
                    std::cerr << "StereoMatcher::compute failed during warmup: " << matcher.last_error() << "\n";
                    return 1;
                }
            struct Sample {
                double ms = 0.0;
                apg::PipelineStats stats;
            };
            std::vector<Sample> runs(repeat);
            for (Sample& s : runs) {
                const auto t0 = std::chrono::high_resolution_clock::now();
                if (!matcher.compute(L, R, bufs, &s.stats)) {
                    std::cerr << "StereoMatcher::compute failed during benchmark: " << matcher.last_error() << "\n";
                    return 1;
                }
                s.ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
            }
            std::sort(runs.begin(), runs.end(), [](const Sample& a, const Sample& b) { return a.ms < b.ms; });
            const Sample& med = runs[repeat / 2];
            const apg::PipelineStats& st = med.stats;

            apg::StereoMetrics m;
            if (has_gt)
                m = apg::evaluate_stereo(bufs.disparity, gt, cfg.prior.enable ? &bufs.range : nullptr,
                                         cfg.refine.enable ? &bufs.d_before_refine : nullptr,
                                         cfg.refine.enable ? &bufs.d_after_refine : nullptr, has_vis ? &vis : nullptr,
                                         1e5f, cfg.prior.enable ? &bufs.supports : nullptr, &L);
            const std::string ids_ = apg::ablation_id_to_string(id), desc = apg::ablation_id_description(id);
            const int paths = cfg.sgm.paths == apg::PathType::Path8 ? 8 : 4;

            std::cout << "[" << ids_ << "] " << std::setw(32) << std::left << desc << std::right << " | EPE: " << std::fixed
                      << std::setprecision(3) << (has_gt ? m.epe : -1.f) << " | Bad2.0: " << std::setprecision(1)
                      << (has_gt ? m.bad_2_0 : -1.f) << "% | D1-all: " << (has_gt ? m.kitti_d1_all : -1.f)
                      << "% | D_bar: " << st.mean_search_width << " | Time: " << std::setprecision(2) << med.ms
                      << " ms (SGM: " << st.timing.sgm_ms << " ms)" << std::endl;

            csv.unsetf(std::ios::floatfield);
            csv << std::setprecision(6);
            csv << (cfg.resolves_to_packed() ? "packed" : "dense") << "," << threads << "," << eff_threads << ","
                << warmup << "," << repeat << ",Release," << c.name << "," << ids_ << ",\"" << desc << "\","
                << ((cfg.cost.use_ad || cfg.cost.use_grad) ? "Census+AD+Grad" : "Census") << ","
                << (cfg.aggregation.enable ? "On(" + std::to_string(cfg.aggregation.iterations) + ")" : std::string("Off"))
                << "," << (cfg.sgm.adaptive_penalties ? "Adaptive" : "Fixed") << "," << (cfg.prior.enable ? "On" : "Off")
                << "," << (cfg.refine.enable ? "On" : "Off") << "," << paths << "," << w << "," << h << "," << c.dmax
                << ",";
            csv << std::fixed << std::setprecision(4) << (has_gt ? m.epe : 0.f) << "," << (has_gt ? m.bad_0_5 : 0.f)
                << "," << (has_gt ? m.bad_1_0 : 0.f) << "," << (has_gt ? m.bad_2_0 : 0.f) << ","
                << (has_gt ? m.bad_3_0 : 0.f) << "," << (has_gt ? m.kitti_d1_all : 0.f) << ","
                << (has_gt ? m.kitti_d1_noc : 0.f) << "," << (has_gt ? m.valid_ratio : 0.f) << ","
                << (has_gt ? m.lr_fail_ratio : 0.f) << "," << (has_gt ? m.edge_epe : 0.f) << ","
                << (has_gt ? m.nonedge_epe : 0.f) << ",";
            if (cfg.prior.enable && has_gt && m.has_range)
                csv << m.range_recall_all << "," << m.range_recall_matchable << "," << m.range_recall_visible << ","
                    << m.range_visible_eval_pixels << "," << m.range_visible_in_pixels << ","
                    << m.range_matchable_eval_pixels << "," << m.range_matchable_in_pixels << ","
                    << m.prior_miss_visible_count << "," << std::setprecision(2) << (m.prior_miss_edge_ratio * 100.0f)
                    << "," << (m.prior_miss_nonedge_ratio * 100.0f) << ",";
            else
                csv << "N/A,N/A,N/A,0,0,0,0,0,0.00,0.00,";
            if (cfg.prior.enable && has_gt && m.has_supports) {
                const apg::SupportMetrics& s = m.support_metrics;
                csv << st.prior_support_count << "," << std::setprecision(4) << s.precision_all_05 << ","
                    << s.precision_all_1 << "," << s.precision_all_2 << "," << s.precision_vis_1 << ","
                    << s.precision_vis_05 << "," << s.precision_vis_1 << "," << s.precision_vis_2 << "," << s.gt_valid
                    << "," << s.visible << "," << s.correct_05 << "," << s.correct_1 << "," << s.correct_2 << ","
                    << s.visible_correct_05 << "," << s.visible_correct_1 << "," << s.visible_correct_2 << ","
                    << s.visible_gt_cells << "," << s.correct_visible_cells << "," << s.mae_all << ","
                    << s.mae_visible << "," << s.grid_recall_1 << ",";
            } else {
                csv << st.prior_support_count << ",0.0,0.0,0.0,0.0,0.0,0.0,0.0,0,0,0,0,0,0,0,0,0,0,0.0,0.0,0.0,";
            }
            csv << std::setprecision(2) << st.mean_search_width << "," << st.mean_geometry_width << ","
                << std::setprecision(4) << st.geometry_reduction_ratio << ","
                << (cfg.prior.enable ? st.prior_incremental_reduction_ratio : 0.0) << "," << st.search_reduction_ratio
                << "," << st.cost_bytes << "," << st.aggregated_cost_bytes << "," << st.estimated_peak_bytes << ","
                << st.reliable_ratio << "," << st.unreliable_ratio << ","
                << (m.has_refine_stats ? m.refine_changed_ratio : 0.f) << ","
                << (m.has_refine_stats ? m.refine_epe_delta : 0.f) << ","
                << (m.has_refine_stats ? m.post_epe_delta : 0.f) << ","
                << (m.has_refine_stats ? m.total_epe_delta : 0.f) << "," << std::setprecision(2) << med.ms << ","
                << st.timing.aux_ms << "," << st.timing.prior_ms << "," << st.timing.cost_ms << ","
                << st.timing.right_wta_ms << "," << st.timing.cross_ms << "," << st.timing.sgm_ms << ","
                << st.timing.wta_ms << "," << st.timing.confidence_ms << "," << st.timing.refine_ms << ","
                << st.timing.post_ms << ",";
            csv << m.reg_all.total_pixels << "," << m.reg_all.valid_est_pixels << "," << std::setprecision(4)
                << m.reg_all.valid_ratio << "," << m.reg_all.epe << "," << m.reg_all.bad_0_5 << ","
                << m.reg_all.bad_1_0 << "," << m.reg_all.bad_2_0 << "," << m.reg_all.bad_3_0 << "," << m.reg_all.d1
                << ",";
            for (const apg::RegionMetrics* r : {&m.reg_noc, &m.reg_occ, &m.reg_boundary, &m.reg_non_boundary,
                                               &m.reg_low_texture, &m.reg_high_texture})
                region(csv, *r);
            // telemetry of the options this build does not have: what the
            // reference writes with them off
            csv << (st.peak_workspace_bytes > 0 ? st.peak_workspace_bytes : st.estimated_peak_bytes) << ","
                << st.retained_workspace_bytes << "," << std::fixed << std::setprecision(4)
                << st.mean_search_width << "," << 0.f << "," << 0.f << "," << 0.f << ",legacy,"
                << 0.f << ",0,0," << 0.0 << ",0,0," << 0.0 << ",0," << 0.0 << "," << 0.f << "," << 0.f << ",legacy,"
                << inpaint_gap << ",0,0," << 0.0 << ",0,0,0,0,0,0,0,0,0,0\n";
            csv.flush();

            if (!save_dir.empty()) {
                const std::string pre = save_dir + "/" + c.name + "_" + ids_;
                apg::save_pfm(pre + "_disp.pfm", bufs.disparity);
                apg::save_disparity_preview(pre + "_disp.pgm", bufs.disparity, static_cast<float>(c.dmax));
                if (cfg.prior.enable && has_gt && m.has_range) {
                    apg::Image8 miss(w, h, 1);
                    for (int y = 0; y < h; ++y)
                        for (int x = 0; x < w; ++x) {
                            const float g = gt.at(x, y);
                            if (g < 0.f || !std::isfinite(g)) continue;
                            const int xr = x - static_cast<int>(std::round(g));
                            const bool v = has_vis ? vis.at(x, y) >= 200 : (xr >= 0 && xr < w);
                            if (v && (g < bufs.range.dmin.at(x, y) || g >= bufs.range.dmax.at(x, y))) miss.at(x, y) = 255;
                        }
                    apg::save_pgm(pre + "_miss.pgm", miss);
                }
            }
        }
    }
    std::cout << "\nBenchmark complete. Results written to: " << output << std::endl;
    return 0;
}
