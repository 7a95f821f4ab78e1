// Host image containers with the reference's interface (reference
// include/apg_sgm/image.hpp:10-140, src/image.cpp:8-41): row-major,
// interleaved channels, clamp-to-edge sample(), (77r+150g+29b)>>8 gray().
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "apg_sgm/types.hpp"

namespace apg {

class Image8 {
public:
    Image8() = default;
    Image8(int w, int h, int channels = 1) : w_(w), h_(h), c_(channels), data_(static_cast<size_t>(w) * h * channels, 0) {}

    int width() const { return w_; }
    int height() const { return h_; }
    int channels() const { return c_; }
    Size size() const { return {w_, h_}; }
    bool empty() const { return data_.empty(); }
    uint8_t* data() { return data_.data(); }
    const uint8_t* data() const { return data_.data(); }
    uint8_t& at(int x, int y, int ch = 0) { return data_[(static_cast<size_t>(y) * w_ + x) * c_ + ch]; }
    uint8_t at(int x, int y, int ch = 0) const { return data_[(static_cast<size_t>(y) * w_ + x) * c_ + ch]; }
    uint8_t sample(int x, int y, int ch = 0) const { return at(clampi(x, 0, w_ - 1), clampi(y, 0, h_ - 1), ch); }
    uint8_t gray(int x, int y) const {
        if (c_ == 1) return at(x, y);
        return static_cast<uint8_t>((77 * at(x, y, 0) + 150 * at(x, y, 1) + 29 * at(x, y, 2)) >> 8);
    }
    void fill(uint8_t v) { std::fill(data_.begin(), data_.end(), v); }
    void release() {
        data_.clear();
        data_.shrink_to_fit();
        w_ = h_ = 0;
        c_ = 1;
    }
    size_t bytes() const { return data_.capacity(); }

private:
    int w_ = 0, h_ = 0, c_ = 1;
    std::vector<uint8_t> data_;
};

template <typename T>
class Plane {
public:
    Plane() = default;
    Plane(int w, int h, T fill = T()) : w_(w), h_(h), data_(static_cast<size_t>(w) * h, fill) {}
    int width() const { return w_; }
    int height() const { return h_; }
    Size size() const { return {w_, h_}; }
    bool empty() const { return data_.empty(); }
    T* data() { return data_.data(); }
    const T* data() const { return data_.data(); }
    T& at(int x, int y) { return data_[static_cast<size_t>(y) * w_ + x]; }
    T at(int x, int y) const { return data_[static_cast<size_t>(y) * w_ + x]; }
    T& operator[](int i) { return data_[i]; }
    T operator[](int i) const { return data_[i]; }
    void fill(T v) { std::fill(data_.begin(), data_.end(), v); }
    void release() {
        data_.clear();
        data_.shrink_to_fit();
        w_ = h_ = 0;
    }
    size_t bytes() const { return data_.capacity() * sizeof(T); }

private:
    int w_ = 0, h_ = 0;
    std::vector<T> data_;
};

using Image32f = Plane<float>;
using Image16s = Plane<int16_t>;
using Image8u1 = Plane<uint8_t>;

// File I/O with the reference's formats (reference src/image.cpp:45-205):
// binary 8-bit PGM (P5) and PPM (P6) in, P5 out; PFM (Pf / PF, first channel,
// bottom-up rows, byte order from the scale's sign) in, little-endian Pf out;
// the preview maps d * 255 / dmax to 8 bits (invalid -> 0).
bool load_image(const std::string& path, Image8& out);
bool save_pgm(const std::string& path, const Image8& img);
bool load_pfm(const std::string& path, Image32f& out);
bool save_pfm(const std::string& path, const Image32f& img);
bool save_disparity_preview(const std::string& path, const Image32f& disp, float dmax);

}  // namespace apg
