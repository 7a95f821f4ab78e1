// Drives the C++ stage classes (include/apg_sgm/stages.hpp) and StereoMatcher
// the way the reference's StereoMatcher::compute_impl does (reference
// src/pipeline.cpp:476-605, dense branch) and prints one JSON object of
// FNV-1a-64 hashes per stage output.  tests/test_gpu_cpp_api.py compares the
// hashes with the CPU oracle on the same input.
//   usage: stage_classes_main <left.raw> <right.raw> <w> <h> <mode> <max_disp> [min_disp]
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "apg_sgm/pipeline.hpp"
#include "apg_sgm/stages.hpp"

static uint64_t fnv(const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    uint64_t h = 14695981039346656037ULL;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

static apg::Image8 load_raw(const char* path, int w, int h) {
    apg::Image8 img(w, h, 1);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(img.data()), static_cast<std::streamsize>(w) * h);
    return img;
}

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s left right w h mode max_disp [min_disp]\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), mode = std::atoi(argv[5]), dmax = std::atoi(argv[6]);
    const int dmin = argc > 7 ? std::atoi(argv[7]) : 0;
    apg::Image8 L = load_raw(argv[1], w, h), R = load_raw(argv[2], w, h);
    apg::PipelineConfig cfg = apg::PipelineConfig::from_mode(static_cast<apg::QualityMode>(mode), dmax);
    cfg.min_disparity = dmin;
    const size_t n = static_cast<size_t>(w) * h;
    try {
        apg::PipelineBuffers buf;
        apg::CostComputer cost;
        apg::PriorEstimator prior;
        apg::CostAggregator agg;
        apg::SgmOptimizer sgm;
        apg::ConfidenceEstimator conf;
        apg::PostProcessor post;
        cost.compute_aux(L, R, cfg, buf);
        prior.estimate(cfg, buf);
        cost.compute_volume(cfg, buf);
        const uint64_t h_cost = fnv(buf.cost.data(), buf.cost.bytes());
        // right disparity from the raw volume is part of the pipeline only;
        // take it from StereoMatcher's legacy buffer API below.
        agg.aggregate(cfg, buf);
        const uint64_t h_agg = fnv(buf.cost.data(), buf.cost.bytes());
        sgm.optimize(cfg, buf);
        const uint64_t h_sgm = fnv(buf.cost32.data(), buf.cost32.bytes());
        apg::Image32f best, second;
        sgm.winner_take_all(cfg, buf.cost32, buf.range, buf.disparity, &best, &second);
        const uint64_t h_wta = fnv(buf.disparity.data(), n * 4), h_best = fnv(best.data(), n * 4),
                       h_second = fnv(second.data(), n * 4);
        apg::PipelineBuffers full;
        apg::StereoMatcher m(cfg);
        if (!m.compute(L, R, full)) {
            std::fprintf(stderr, "compute failed: %s\n", m.last_error().c_str());
            return 1;
        }
        buf.disparity_right = full.disparity_right;
        conf.estimate(cfg, buf, best, second);
        const uint64_t h_conf = fnv(buf.confidence.data(), n * 4);
        post.left_right_check(cfg, buf);
        post.fill_holes(cfg, buf);
        post.median(cfg, buf);
        // 1-path class API: left-to-right path of the raw volume into a zeroed u32 accumulator
        apg::CostVolume32 acc(w, h, cfg.min_disparity, cfg.max_disparity, 0);
        apg::PipelineBuffers b2;
        cost.compute_aux(L, R, cfg, b2);
        prior.estimate(cfg, b2);
        cost.compute_volume(cfg, b2);
        sgm.aggregate_path(cfg, b2.left_gray, b2.cost, acc, 1, 0);
        const uint64_t h_path = fnv(acc.data(), acc.bytes());
        std::printf(
            "{\"cost\": \"0x%llx\", \"cost_agg\": \"0x%llx\", \"sgm\": \"0x%llx\", \"disp_wta\": \"0x%llx\", "
            "\"best\": \"0x%llx\", \"second\": \"0x%llx\", \"confidence_stage\": \"0x%llx\", "
            "\"disparity_stages\": \"0x%llx\", \"disparity\": \"0x%llx\", \"confidence\": \"0x%llx\", "
            "\"invalid_reason\": \"0x%llx\", \"reliable\": \"0x%llx\", \"disp_right\": \"0x%llx\", "
            "\"path_lr_raw\": \"0x%llx\"}\n",
            (unsigned long long)h_cost, (unsigned long long)h_agg, (unsigned long long)h_sgm,
            (unsigned long long)h_wta, (unsigned long long)h_best, (unsigned long long)h_second,
            (unsigned long long)h_conf, (unsigned long long)fnv(buf.disparity.data(), n * 4),
            (unsigned long long)fnv(full.disparity.data(), n * 4), (unsigned long long)fnv(full.confidence.data(), n * 4),
            (unsigned long long)fnv(full.invalid_reason.data(), n), (unsigned long long)fnv(full.reliable_mask.data(), n),
            (unsigned long long)fnv(full.disparity_right.data(), n * 4), (unsigned long long)h_path);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 1;
    }
    return 0;
}
