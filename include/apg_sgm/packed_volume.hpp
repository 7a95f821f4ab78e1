// Ragged ("packed") cost volumes of the HighQuality path, with the
// reference's interface (reference include/apg_sgm/packed_volume.hpp:15-233):
// pixel p holds the disparities [dmin_p, dmax_p) contiguously at
// offsets[p], offsets being the prefix sum of the per-pixel widths.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <vector>

#include "apg_sgm/search_range.hpp"

namespace apg {

struct PackedVolumeLayout {
    int width = 0;
    int height = 0;
    std::vector<uint32_t> offsets;   // N + 1 prefix sums of the widths
    std::vector<int16_t> dmin;       // N

    static std::shared_ptr<PackedVolumeLayout> from_range(const SearchRange& range) {
        auto L = std::make_shared<PackedVolumeLayout>();
        L->width = range.dmin.width();
        L->height = range.dmin.height();
        const size_t n = static_cast<size_t>(L->width) * L->height;
        L->offsets.assign(n + 1, 0u);
        L->dmin.resize(n);
        uint64_t total = 0;
        for (size_t p = 0; p < n; ++p) {
            const int lo = range.dmin.data()[p], hi = range.dmax.data()[p];
            total += hi > lo ? static_cast<uint64_t>(hi - lo) : 0u;
            if (total > UINT32_MAX)
                throw std::runtime_error("PackedVolumeLayout: total disparity states exceed uint32 capacity");
            L->offsets[p + 1] = static_cast<uint32_t>(total);
            L->dmin[p] = static_cast<int16_t>(lo);
        }
        return L;
    }

    uint32_t state_count() const { return offsets.empty() ? 0u : offsets.back(); }
    int disp_width(size_t p) const { return static_cast<int>(offsets[p + 1] - offsets[p]); }
    int dmax(size_t p) const { return static_cast<int>(dmin[p]) + disp_width(p); }
    size_t bytes() const { return offsets.size() * sizeof(uint32_t) + dmin.size() * sizeof(int16_t); }
    size_t allocated_bytes() const { return offsets.capacity() * sizeof(uint32_t) + dmin.capacity() * sizeof(int16_t); }
};

template <typename T>
class PackedCostVolume {
public:
    PackedCostVolume() = default;
    explicit PackedCostVolume(std::shared_ptr<const PackedVolumeLayout> layout, T fill = 0) {
        allocate(std::move(layout), fill);
    }
    explicit PackedCostVolume(const SearchRange& range, T fill = 0) { allocate(range, fill); }

    void allocate_for_overwrite(std::shared_ptr<const PackedVolumeLayout> layout) {
        layout_ = std::move(layout);
        data_.resize(layout_ ? layout_->state_count() : 0);
    }
    void allocate_for_overwrite(const SearchRange& range) { allocate_for_overwrite(PackedVolumeLayout::from_range(range)); }
    void allocate(std::shared_ptr<const PackedVolumeLayout> layout, T fill = 0) {
        allocate_for_overwrite(std::move(layout));
        std::fill(data_.begin(), data_.end(), fill);
    }
    void allocate(const SearchRange& range, T fill = 0) { allocate(PackedVolumeLayout::from_range(range), fill); }
    void fill(T value) { std::fill(data_.begin(), data_.end(), value); }
    void release() {
        data_ = std::vector<T>();
        layout_.reset();
    }

    size_t capacity() const { return data_.capacity(); }
    size_t capacity_bytes() const { return data_.capacity() * sizeof(T); }
    std::shared_ptr<const PackedVolumeLayout> layout() const { return layout_; }
    int width() const { return layout_ ? layout_->width : 0; }
    int height() const { return layout_ ? layout_->height : 0; }
    bool empty() const { return data_.empty(); }
    size_t total_elements() const { return data_.size(); }

    int dmin(int x, int y) const { return layout_ ? layout_->dmin[pix(x, y)] : 0; }
    int dmax(int x, int y) const { return layout_ ? layout_->dmax(pix(x, y)) : 0; }
    int disp_width(int x, int y) const { return layout_ ? layout_->disp_width(pix(x, y)) : 0; }
    bool contains(int x, int y, int d) const {
        if (!layout_ || x < 0 || x >= layout_->width || y < 0 || y >= layout_->height) return false;
        return d >= dmin(x, y) && d < dmax(x, y);
    }
    uint32_t offset(int x, int y) const { return layout_->offsets[pix(x, y)]; }
    T* slice(int x, int y) { return data_.data() + offset(x, y); }
    const T* slice(int x, int y) const { return data_.data() + offset(x, y); }
    T& at(int x, int y, int d) { return slice(x, y)[d - dmin(x, y)]; }
    const T& at(int x, int y, int d) const { return slice(x, y)[d - dmin(x, y)]; }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }

    size_t data_bytes() const { return data_.size() * sizeof(T); }
    size_t layout_bytes() const { return layout_ ? layout_->bytes() : 0; }
    size_t bytes() const { return data_bytes() + layout_bytes(); }
    size_t allocated_bytes() const { return capacity_bytes() + (layout_ ? layout_->allocated_bytes() : 0); }

private:
    size_t pix(int x, int y) const { return static_cast<size_t>(y) * layout_->width + x; }

    std::shared_ptr<const PackedVolumeLayout> layout_;
    std::vector<T> data_;
};

using PackedCostVolume16 = PackedCostVolume<uint16_t>;
using PackedCostVolume32 = PackedCostVolume<uint32_t>;

}  // namespace apg
