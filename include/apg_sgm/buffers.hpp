// Host-side stage buffers with the reference's layout and names (reference
// include/apg_sgm/buffers.hpp:13-160, search_range.hpp:10-47).  Volumes are
// [y][x][d]: element (x, y, d) at ((y * w + x) * D + d - d0).
#pragma once

#include <algorithm>
#include <vector>

#include "apg_sgm/image.hpp"
#include "apg_sgm/packed_volume.hpp"
#include "apg_sgm/search_range.hpp"
#include "apg_sgm/types.hpp"

namespace apg {

template <typename T>
class CostVolumeT {
public:
    CostVolumeT() = default;
    CostVolumeT(int w, int h, int d0, int d1, T fill = 0) { allocate(w, h, d0, d1, fill); }
    void allocate(int w, int h, int d0, int d1, T fill = 0) {
        w_ = w;
        h_ = h;
        d0_ = d0;
        D_ = d1 - d0;
        data_.assign(static_cast<size_t>(w) * h * D_, fill);
    }
    void release() {
        data_.clear();
        data_.shrink_to_fit();
        w_ = h_ = d0_ = D_ = 0;
    }
    int width() const { return w_; }
    int height() const { return h_; }
    int d0() const { return d0_; }
    int D() const { return D_; }
    bool empty() const { return data_.empty(); }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }
    size_t index(int x, int y, int d) const { return (static_cast<size_t>(y) * w_ + x) * D_ + (d - d0_); }
    T& at(int x, int y, int d) { return data_[index(x, y, d)]; }
    T at(int x, int y, int d) const { return data_[index(x, y, d)]; }
    T* slice(int x, int y) { return data_.data() + (static_cast<size_t>(y) * w_ + x) * D_; }
    const T* slice(int x, int y) const { return data_.data() + (static_cast<size_t>(y) * w_ + x) * D_; }
    size_t bytes() const { return data_.size() * sizeof(T); }
    size_t capacity_bytes() const { return data_.capacity() * sizeof(T); }

private:
    int w_ = 0, h_ = 0, d0_ = 0, D_ = 0;
    std::vector<T> data_;
};

using CostVolume = CostVolumeT<uint16_t>;
using CostVolume32 = CostVolumeT<uint32_t>;

// The reference's PipelineBuffers (reference include/apg_sgm/buffers.hpp:66-160).
struct PipelineBuffers {
    Image8 left, right;
    Image8 left_gray, right_gray;
    Image8 left_gx, left_gy, right_gx, right_gy;
    std::vector<uint32_t> census_left, census_right;
    std::vector<uint64_t> census_left64, census_right64;
    SearchRange range;
    std::vector<SupportMatch> supports;
    size_t support_count = 0;
    Image32f d_prior;
    CostVolume cost;
    CostVolume32 cost32;
    PackedCostVolume16 packed_cost;
    PackedCostVolume16 packed_sgm16;
    PackedCostVolume32 packed_cost32;
    Image32f disparity;
    Image32f d_before_refine;
    Image32f d_after_refine;
    Image32f disparity_right;
    Image32f confidence;
    Image8u1 invalid_reason;
    Image8u1 reliable_mask;

    void release() { *this = PipelineBuffers(); }

    size_t allocated_bytes() const {
        size_t t = left.bytes() + right.bytes() + left_gray.bytes() + right_gray.bytes();
        t += left_gx.bytes() + left_gy.bytes() + right_gx.bytes() + right_gy.bytes();
        t += census_left.capacity() * 4 + census_right.capacity() * 4;
        t += census_left64.capacity() * 8 + census_right64.capacity() * 8;
        t += range.bytes() + supports.capacity() * sizeof(SupportMatch) + d_prior.bytes();
        t += cost.capacity_bytes() + cost32.capacity_bytes();
        t += packed_cost.allocated_bytes() + packed_sgm16.allocated_bytes() + packed_cost32.allocated_bytes();
        t += disparity.bytes() + d_before_refine.bytes() + d_after_refine.bytes();
        t += disparity_right.bytes() + confidence.bytes() + invalid_reason.bytes() + reliable_mask.bytes();
        return t;
    }
};

}  // namespace apg
