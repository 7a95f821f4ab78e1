// Host copies split over a small pool of threads (serving form and the
// session path): one stereo request moves its two images in and four result
// planes out, about 10 bytes per pixel, through pinned staging buffers, and
// at serving rates that is more than one core copies.
#pragma once

#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

namespace apgk {

class CopyPool {
public:
    explicit CopyPool(int nthreads) {
        for (int i = 0; i < nthreads; ++i) th_.emplace_back([this] { work(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    // Copies every (dst, src, bytes) of the list in 1 MiB pieces, the calling
    // thread included; returns when all are done.
    void copy(const std::vector<std::tuple<void*, const void*, size_t>>& list) {
        constexpr size_t kPiece = size_t(1) << 20;
        std::vector<std::tuple<void*, const void*, size_t>> pieces;
        for (const auto& c : list) {
            uint8_t* d = (uint8_t*)std::get<0>(c);
            const uint8_t* s = (const uint8_t*)std::get<1>(c);
            const size_t n = std::get<2>(c);
            for (size_t o = 0; o < n; o += kPiece) pieces.emplace_back(d + o, s + o, std::min(kPiece, n - o));
        }
        if (th_.empty() || pieces.size() <= 1) {
            for (const auto& p : pieces) memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            for (const auto& p : pieces) q_.push_back(p);
            pending_ += pieces.size();
        }
        cv_.notify_all();
        for (;;) {
            std::tuple<void*, const void*, size_t> p;
            {
                std::unique_lock<std::mutex> g(m_);
                if (q_.empty()) {
                    done_cv_.wait(g, [this] { return pending_ == 0; });
                    return;
                }
                p = q_.front();
                q_.pop_front();
            }
            memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
            finish_one();
        }
    }
    // The pool size used by default: up to 7 threads besides the caller.
    static int default_threads() {
        const int hw = (int)std::thread::hardware_concurrency();
        return std::max(0, std::min(7, hw / 2 - 1));
    }

private:
    void finish_one() {
        std::lock_guard<std::mutex> g(m_);
        if (--pending_ == 0) done_cv_.notify_all();
    }
    void work() {
        for (;;) {
            std::tuple<void*, const void*, size_t> p;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                p = q_.front();
                q_.pop_front();
            }
            memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
            finish_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::tuple<void*, const void*, size_t>> q_;
    size_t pending_ = 0;
    bool stop_ = false;
};

}  // namespace apgk
