// StereoMatcher / StereoWorkspace / StereoSession with the reference's
// interface (reference include/apg_sgm/pipeline.hpp:24-99, workspace.hpp:10-75,
// session.hpp:17-110, result.hpp, execution.hpp, pipeline_stats.hpp).  The
// compute path runs on the MI355X: inputs are uploaded once, every stage from
// gray/census to the median filter runs on device, and only the four result
// planes come back.  Intermediate volumes live in a device workspace owned by
// the StereoWorkspace.
#pragma once

#include <cstddef>
#include <string>

#include "apg_sgm/buffers.hpp"
#include "apg_sgm/config.hpp"
#include "apg_sgm/image_view.hpp"
#include "apg_sgm/status.hpp"

struct apg_hip_ctx_s;

namespace apg {

enum class MemoryPolicy { ReleasePerFrame = 0, BoundedReuse = 1 };

struct ExecutionOptions {
    MemoryPolicy memory_policy = MemoryPolicy::ReleasePerFrame;
    size_t max_retained_bytes = 0;
    int num_threads = 0;   // host threads; the device path ignores it
    int device = 0;        // HIP device of the workspace
};

struct StereoResult {
    Image32f disparity;
    Image32f confidence;
    Image8u1 invalid_reason;
    Image8u1 reliable_mask;
    void clear() { *this = StereoResult(); }
    bool empty() const { return disparity.empty(); }
    int width() const { return disparity.width(); }
    int height() const { return disparity.height(); }
};

struct StageTiming {
    double aux_ms = 0, prior_ms = 0, cost_ms = 0, cross_ms = 0, right_wta_ms = 0, sgm_ms = 0, wta_ms = 0,
           confidence_ms = 0, refine_ms = 0, post_ms = 0, total_ms = 0;
};

struct PipelineStats {
    StageTiming timing;
    size_t cost_bytes = 0;
    size_t aggregated_cost_bytes = 0;
    size_t estimated_peak_bytes = 0;
    size_t peak_workspace_bytes = 0;
    size_t retained_workspace_bytes = 0;
    size_t frame_allocation_count = 0;
    size_t frame_reuse_count = 0;
    size_t allocation_count = 0;
    size_t reuse_count = 0;
    int requested_threads = 0;
    int effective_threads = 0;
    std::string requested_backend = "Auto";
    std::string resolved_backend = "Dense";
    std::string memory_policy = "ReleasePerFrame";
    // search ranges: mean width of the frame's ranges and of the dense
    // geometric init [min, min(max, x+1)), and the reductions they give
    size_t prior_support_count = 0;
    double mean_search_width = 0.0;
    double mean_geometry_width = 0.0;
    double geometry_reduction_ratio = 0.0;
    double prior_incremental_reduction_ratio = 0.0;
    double search_reduction_ratio = 0.0;
    size_t valid_pixels = 0;
    size_t reliable_pixels = 0;
    size_t unreliable_pixels = 0;
    size_t refine_changed_pixels = 0;
    size_t lr_fail_pixels = 0;
    float valid_disparity_ratio = 0.f;
    float reliable_ratio = 0.f;
    float unreliable_ratio = 0.f;
    float refine_changed_ratio = 0.f;
    float lr_fail_ratio = 0.f;
    std::string to_json() const;
};

Status validate_config(const PipelineConfig& cfg);
Status validate_inputs(const ImageView8& left, const ImageView8& right);
inline Status validate_inputs(const Image8& l, const Image8& r) { return validate_inputs(ImageView8(l), ImageView8(r)); }
inline Status validate_inputs(const StereoInputView& in) { return validate_inputs(in.left, in.right); }

// Same estimate as the reference's early budget gate (reference
// src/pipeline.cpp:246-333), so that a given max_retained_bytes rejects the
// same frame sizes.
size_t estimate_workspace_bytes(int width, int height, const PipelineConfig& cfg);

// Device workspace + frame counters (reference include/apg_sgm/workspace.hpp).
class StereoWorkspace {
public:
    explicit StereoWorkspace(ExecutionOptions options = {});
    ~StereoWorkspace();
    StereoWorkspace(const StereoWorkspace&) = delete;
    StereoWorkspace& operator=(const StereoWorkspace&) = delete;

    const ExecutionOptions& options() const { return options_; }
    void set_options(const ExecutionOptions& o) { options_ = o; }
    size_t allocated_bytes() const;
    size_t frame_allocation_count() const { return frame_allocation_count_; }
    size_t frame_reuse_count() const { return frame_reuse_count_; }
    size_t allocation_count() const { return frame_allocation_count_; }
    size_t reuse_count() const { return frame_reuse_count_; }
    void record_frame(int w, int h, int d_max, bool was_reuse);
    bool can_reuse(int w, int h, int d_max) const;
    void release();
    void reset_counters() { frame_allocation_count_ = frame_reuse_count_ = 0; }
    void trim(size_t max_bytes) {
        if (allocated_bytes() > max_bytes) release();
    }
    // Device context (created on first use).
    apg_hip_ctx_s* device_context();

private:
    ExecutionOptions options_;
    apg_hip_ctx_s* ctx_ = nullptr;
    int last_w_ = 0, last_h_ = 0, last_dmax_ = 0;
    size_t frame_allocation_count_ = 0, frame_reuse_count_ = 0;
    friend class StereoMatcher;
    void* dev_io_ = nullptr;   // device input/output staging
    size_t dev_io_bytes_ = 0;
    void* host_io_ = nullptr;  // pinned host staging of the same layout (retaining policies)
    size_t host_io_bytes_ = 0;
};

class StereoMatcher {
public:
    explicit StereoMatcher(PipelineConfig cfg = PipelineConfig::from_mode(QualityMode::Balanced)) : cfg_(std::move(cfg)) {}
    void set_config(const PipelineConfig& cfg) { cfg_ = cfg; }
    const PipelineConfig& config() const { return cfg_; }

    Status compute_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                          StereoWorkspace& workspace, PipelineStats* stats = nullptr) const;
    Status compute_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                          PipelineStats* stats = nullptr) const;
    Status compute_status(const StereoInputView& in, StereoResult& r, StereoWorkspace& ws,
                          PipelineStats* stats = nullptr) const {
        return compute_status(in.left, in.right, r, ws, stats);
    }
    Status compute_status(const StereoInputView& in, StereoResult& r, PipelineStats* stats = nullptr) const {
        return compute_status(in.left, in.right, r, stats);
    }
    Status compute_status(const Image8& l, const Image8& r, StereoResult& res, StereoWorkspace& ws,
                          PipelineStats* stats = nullptr) const {
        return compute_status(ImageView8(l), ImageView8(r), res, ws, stats);
    }
    Status compute_status(const Image8& l, const Image8& r, StereoResult& res, PipelineStats* stats = nullptr) const {
        return compute_status(ImageView8(l), ImageView8(r), res, stats);
    }
    bool compute(const ImageView8& l, const ImageView8& r, StereoResult& res, StereoWorkspace& ws,
                 PipelineStats* stats = nullptr) const;
    bool compute(const ImageView8& l, const ImageView8& r, StereoResult& res, PipelineStats* stats = nullptr) const;
    bool compute(const StereoInputView& in, StereoResult& res, StereoWorkspace& ws, PipelineStats* stats = nullptr) const {
        return compute(in.left, in.right, res, ws, stats);
    }
    bool compute(const StereoInputView& in, StereoResult& res, PipelineStats* stats = nullptr) const {
        return compute(in.left, in.right, res, stats);
    }
    bool compute(const Image8& l, const Image8& r, StereoResult& res, StereoWorkspace& ws,
                 PipelineStats* stats = nullptr) const {
        return compute(ImageView8(l), ImageView8(r), res, ws, stats);
    }
    bool compute(const Image8& l, const Image8& r, StereoResult& res, PipelineStats* stats = nullptr) const {
        return compute(ImageView8(l), ImageView8(r), res, stats);
    }
    // Legacy API: fills the caller's buffers (raw-volume stages are not
    // materialised on the host; disparity/confidence/masks/right disparity are).
    bool compute(const Image8& left, const Image8& right, PipelineBuffers& out, PipelineStats* stats = nullptr) const;

    const std::string& last_error() const { return last_error_; }

private:
    Status compute_impl(const ImageView8& left, const ImageView8& right, StereoWorkspace& ws, StereoResult& result,
                        PipelineBuffers* extra, PipelineStats* stats) const;
    PipelineConfig cfg_;
    mutable std::string last_error_;
};

class StereoSession {
public:
    explicit StereoSession(PipelineConfig cfg = PipelineConfig::from_mode(QualityMode::Balanced),
                           ExecutionOptions opt = {})
        : cfg_(std::move(cfg)), workspace_(opt), matcher_(cfg_) {}

    Status process_status(const ImageView8& left, const ImageView8& right, StereoResult& result,
                          PipelineStats* stats = nullptr);
    Status process_status(const StereoInputView& in, StereoResult& r, PipelineStats* stats = nullptr) {
        return process_status(in.left, in.right, r, stats);
    }
    Status process_status(const Image8& l, const Image8& r, StereoResult& res, PipelineStats* stats = nullptr) {
        return process_status(ImageView8(l), ImageView8(r), res, stats);
    }
    bool process(const ImageView8& l, const ImageView8& r, StereoResult& res, PipelineStats* stats = nullptr) {
        return process_status(l, r, res, stats).ok();
    }
    bool process(const StereoInputView& in, StereoResult& res, PipelineStats* stats = nullptr) {
        return process(in.left, in.right, res, stats);
    }
    bool process(const Image8& l, const Image8& r, StereoResult& res, PipelineStats* stats = nullptr) {
        return process(ImageView8(l), ImageView8(r), res, stats);
    }
    void reset() {
        workspace_.release();
        frame_count_ = 0;
        total_time_ms_ = 0.0;
        last_error_.clear();
    }
    void set_config(const PipelineConfig& cfg) {
        cfg_ = cfg;
        matcher_.set_config(cfg_);
    }
    const PipelineConfig& config() const { return cfg_; }
    void set_options(const ExecutionOptions& opt) { workspace_.set_options(opt); }
    const ExecutionOptions& options() const { return workspace_.options(); }
    size_t frame_count() const { return frame_count_; }
    double total_time_ms() const { return total_time_ms_; }
    double average_time_ms() const { return frame_count_ ? total_time_ms_ / static_cast<double>(frame_count_) : 0.0; }
    size_t allocated_bytes() const { return workspace_.allocated_bytes(); }
    const std::string& last_error() const { return last_error_; }
    StereoWorkspace& workspace() { return workspace_; }
    const StereoWorkspace& workspace() const { return workspace_; }

private:
    PipelineConfig cfg_;
    StereoWorkspace workspace_;
    StereoMatcher matcher_;
    size_t frame_count_ = 0;
    double total_time_ms_ = 0.0;
    std::string last_error_;
};

}  // namespace apg
