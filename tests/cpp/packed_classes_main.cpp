// Drives the ragged (packed) stage-class methods the way the reference's
// StereoMatcher::compute_impl does for the HighQuality preset (reference
// src/pipeline.cpp:513-558): prior ranges -> layout -> packed cost -> packed
// Cross -> packed SGM (u16 and u32) -> packed WTA, plus one packed path.
// Writes each output as raw little-endian arrays into <outdir>;
// tests/test_gpu_cpp_api.py packs the oracle's dense volumes over the same
// ranges and compares.
//   usage: packed_classes_main <left.raw> <right.raw> <w> <h> <max_disp> <outdir>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "apg_sgm/stages.hpp"

static apg::Image8 load_raw(const char* path, int w, int h) {
    apg::Image8 img(w, h, 1);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(img.data()), static_cast<std::streamsize>(w) * h);
    return img;
}

static void save(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    f.write(static_cast<const char*>(p), static_cast<std::streamsize>(bytes));
}

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s left right w h max_disp outdir\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), dmax = std::atoi(argv[5]);
    const std::string out = argv[6];
    apg::Image8 L = load_raw(argv[1], w, h), R = load_raw(argv[2], w, h);
    const apg::PipelineConfig cfg = apg::PipelineConfig::from_mode(apg::QualityMode::HighQuality, dmax);
    const size_t n = static_cast<size_t>(w) * h;
    try {
        apg::PipelineBuffers buf;
        apg::CostComputer cost;
        apg::PriorEstimator prior;
        apg::CostAggregator agg;
        apg::SgmOptimizer sgm;
        cost.compute_aux(L, R, cfg, buf);
        prior.estimate(cfg, buf);
        save(out + "/dmin.bin", buf.range.dmin.data(), n * 2);
        save(out + "/dmax.bin", buf.range.dmax.data(), n * 2);
        buf.packed_cost.allocate_for_overwrite(apg::PackedVolumeLayout::from_range(buf.range));
        cost.compute_volume_packed(cfg, buf, buf.packed_cost);
        save(out + "/cost.bin", buf.packed_cost.data(), buf.packed_cost.data_bytes());
        apg::PackedCostVolume32 path(buf.packed_cost.layout(), 0u);
        sgm.aggregate_path_packed(cfg, buf.left_gray, buf.packed_cost, path, 1, 0);
        save(out + "/path.bin", path.data(), path.data_bytes());
        agg.aggregate_packed(cfg, buf.left_gray, buf.packed_cost);
        save(out + "/cost_agg.bin", buf.packed_cost.data(), buf.packed_cost.data_bytes());
        sgm.optimize_packed(cfg, buf.left_gray, buf.packed_cost, buf.packed_sgm16);
        sgm.optimize_packed(cfg, buf.left_gray, buf.packed_cost, buf.packed_cost32);
        save(out + "/sgm16.bin", buf.packed_sgm16.data(), buf.packed_sgm16.data_bytes());
        save(out + "/sgm32.bin", buf.packed_cost32.data(), buf.packed_cost32.data_bytes());
        apg::Image32f disp, best, second, disp16;
        sgm.winner_take_all_packed(cfg, buf.packed_cost32, disp, &best, &second);
        sgm.winner_take_all_packed(cfg, buf.packed_sgm16, disp16);
        save(out + "/disp.bin", disp.data(), n * 4);
        save(out + "/best.bin", best.data(), n * 4);
        save(out + "/second.bin", second.data(), n * 4);
        save(out + "/disp16.bin", disp16.data(), n * 4);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 1;
    }
    return 0;
}
