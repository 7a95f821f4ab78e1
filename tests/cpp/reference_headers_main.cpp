// A client written against the reference's own header paths: every header a
// reference test or app includes, each by its reference path, then the
// session, the matcher with a workspace and the HighQuality stage chain
// (prior, 8-path SGM, confidence, refiner) through the stage classes.
// Prints one JSON object of FNV-1a-64 hashes; tests/test_gpu_cpp_api.py
// compares them with the CPU oracle on the same input.
//   usage: reference_headers_main <left.raw> <right.raw> <w> <h> <max_disp>
#include "apg_sgm/buffers.hpp"
#include "apg_sgm/confidence_estimator.hpp"
#include "apg_sgm/config.hpp"
#include "apg_sgm/cost_aggregator.hpp"
#include "apg_sgm/cost_computer.hpp"
#include "apg_sgm/execution.hpp"
#include "apg_sgm/image.hpp"
#include "apg_sgm/image_view.hpp"
#include "apg_sgm/packed_volume.hpp"
#include "apg_sgm/pipeline.hpp"
#include "apg_sgm/pipeline_stats.hpp"
#include "apg_sgm/postprocess.hpp"
#include "apg_sgm/prior_estimator.hpp"
#include "apg_sgm/refiner.hpp"
#include "apg_sgm/result.hpp"
#include "apg_sgm/search_range.hpp"
#include "apg_sgm/session.hpp"
#include "apg_sgm/sgm_optimizer.hpp"
#include "apg_sgm/status.hpp"
#include "apg_sgm/types.hpp"
#include "apg_sgm/workspace.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>

static unsigned long long fnv(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    unsigned long long h = 14695981039346656037ULL;
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
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s left right w h max_disp\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), dmax = std::atoi(argv[5]);
    const apg::Image8 L = load_raw(argv[1], w, h), R = load_raw(argv[2], w, h);
    const size_t n = static_cast<size_t>(w) * h;

    // session.hpp / result.hpp / pipeline_stats.hpp: a Balanced frame
    apg::StereoSession session(apg::PipelineConfig::from_mode(apg::QualityMode::Balanced, dmax));
    apg::StereoResult res;
    apg::PipelineStats stats;
    const apg::Status st = session.process_status(L, R, res, &stats);
    if (!st.ok() || session.frame_count() != 1 || session.allocated_bytes() != 0) {
        std::fprintf(stderr, "session: %s\n", session.last_error().c_str());
        return 1;
    }

    // workspace.hpp / execution.hpp: the matcher on a retained workspace, twice
    apg::ExecutionOptions opt;
    opt.memory_policy = apg::MemoryPolicy::BoundedReuse;
    apg::StereoWorkspace ws(opt);
    apg::StereoMatcher matcher(apg::PipelineConfig::from_mode(apg::QualityMode::Balanced, dmax));
    apg::StereoResult r1, r2;
    if (!matcher.compute_status(L, R, r1, ws).ok() || !matcher.compute_status(L, R, r2, ws).ok() ||
        ws.frame_reuse_count() != 1) {
        std::fprintf(stderr, "matcher: %s\n", matcher.last_error().c_str());
        return 1;
    }

    // the HighQuality stage chain through the class headers
    const apg::PipelineConfig hq = apg::PipelineConfig::from_mode(apg::QualityMode::HighQuality, dmax);
    unsigned long long h_wta = 0, h_conf = 0, h_ref = 0, h_dmin = 0, h_dmax = 0;
    size_t nsup = 0;
    try {
        apg::PipelineBuffers buf;
        apg::CostComputer cost;
        apg::PriorEstimator prior;
        apg::CostAggregator agg;
        apg::SgmOptimizer sgm;
        apg::ConfidenceEstimator conf;
        apg::Refiner refiner;
        cost.compute_aux(L, R, hq, buf);
        prior.estimate(hq, buf);
        nsup = prior.extract_supports(hq, buf).size();
        h_dmin = fnv(buf.range.dmin.data(), n * 2);
        h_dmax = fnv(buf.range.dmax.data(), n * 2);
        cost.compute_volume(hq, buf);
        agg.aggregate(hq, buf);
        sgm.optimize(hq, buf);
        apg::Image32f best, second;
        sgm.winner_take_all(hq, buf.cost32, buf.range, buf.disparity, &best, &second);
        h_wta = fnv(buf.disparity.data(), n * 4);
        // the right-view disparity is a pipeline step, not a class method
        apg::PipelineBuffers full;
        apg::StereoMatcher m(hq);
        if (!m.compute(L, R, full)) {
            std::fprintf(stderr, "hq compute: %s\n", m.last_error().c_str());
            return 1;
        }
        buf.disparity_right = full.disparity_right;
        conf.estimate(hq, buf, best, second);
        h_conf = fnv(buf.confidence.data(), n * 4);
        refiner.refine(hq, buf);
        h_ref = fnv(buf.disparity.data(), n * 4);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 1;
    }
    std::printf(
        "{\"disparity\": \"0x%llx\", \"confidence\": \"0x%llx\", \"invalid_reason\": \"0x%llx\", "
        "\"reliable\": \"0x%llx\", \"matcher_disparity\": \"0x%llx\", \"hq_dmin\": \"0x%llx\", "
        "\"hq_dmax\": \"0x%llx\", \"hq_disp_wta\": \"0x%llx\", \"hq_confidence\": \"0x%llx\", "
        "\"hq_after_refine\": \"0x%llx\", \"hq_supports\": %zu, \"valid_ratio\": %.6f}\n",
        fnv(res.disparity.data(), n * 4), fnv(res.confidence.data(), n * 4), fnv(res.invalid_reason.data(), n),
        fnv(res.reliable_mask.data(), n), fnv(r2.disparity.data(), n * 4), h_dmin, h_dmax, h_wta, h_conf, h_ref,
        nsup, static_cast<double>(stats.valid_disparity_ratio));
    return 0;
}
