// Python module apg_sgm._core with the reference's Python API (reference
// python/src/binding.cpp:374-563, python/apg_sgm/__init__.py): StereoSession
// over the MI355X library's C++ StereoSession, StereoResult with NumPy views
// of the result planes, StageTiming / PipelineStats and the enums.  Errors
// follow the reference: bad option strings and mismatched images raise
// ValueError, a failed frame raises RuntimeError with the status code.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <cctype>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "apg_sgm/pipeline.hpp"

namespace py = pybind11;

namespace {

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

// A uint8 HxW or HxWxC (C = 1 or 3) buffer as an ImageView8: rows may be
// strided, pixels must be packed; anything else is copied.
struct HeldView {
    apg::ImageView8 view;
    std::vector<uint8_t> copy;
};

HeldView hold(const py::buffer& b, const char* name) {
    const py::buffer_info bi = b.request();
    if (bi.format != py::format_descriptor<uint8_t>::format())
        throw py::value_error(std::string(name) + " image must be uint8");
    if (bi.ndim != 2 && bi.ndim != 3) throw py::value_error(std::string(name) + " image must be HxW or HxWxC");
    const int h = (int)bi.shape[0], w = (int)bi.shape[1], c = bi.ndim == 3 ? (int)bi.shape[2] : 1;
    if (c != 1 && c != 3) throw py::value_error(std::string(name) + " image must have 1 or 3 channels");
    const py::ssize_t sy = bi.strides[0], sx = bi.strides[1], sc = bi.ndim == 3 ? bi.strides[2] : 1;
    const uint8_t* base = static_cast<const uint8_t*>(bi.ptr);
    HeldView hv;
    if (sx == c && sc == 1 && sy >= (py::ssize_t)w * c) {
        hv.view = apg::ImageView8(base, w, h, (int)sy, c);
        return hv;
    }
    hv.copy.resize((size_t)w * h * c);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < c; ++k) hv.copy[((size_t)y * w + x) * c + k] = base[y * sy + x * sx + k * sc];
    hv.view = apg::ImageView8(hv.copy.data(), w, h, w * c, c);
    return hv;
}

class Result {
public:
    Result(apg::StereoResult r, apg::PipelineStats s) : r_(std::move(r)), s_(std::move(s)) {}
    const apg::StereoResult& r() const { return r_; }
    const apg::PipelineStats& stats() const { return s_; }

private:
    apg::StereoResult r_;
    apg::PipelineStats s_;
};

template <typename T, typename Plane>
py::array plane_view(py::object owner, const Plane& p) {
    if (p.empty()) return py::array_t<T>();
    const py::ssize_t h = p.height(), w = p.width();
    return py::array_t<T>({h, w}, {(py::ssize_t)(w * sizeof(T)), (py::ssize_t)sizeof(T)}, p.data(), owner);
}

class Session {
public:
    Session(const py::object& mode, int threads, int max_disparity, int min_disparity, const py::object& backend,
            const py::object& memory_policy, size_t max_retained_bytes, bool hierarchical, int hierarchical_radius,
            const py::object& hierarchical_fallback, bool adaptive_radius, float confidence_threshold,
            int expanded_radius) {
        apg::QualityMode qm = apg::QualityMode::Balanced;
        if (py::isinstance<py::str>(mode)) {
            const std::string m = lower(mode.cast<std::string>());
            if (m == "fast") qm = apg::QualityMode::Fast;
            else if (m == "balanced") qm = apg::QualityMode::Balanced;
            else if (m == "high_quality" || m == "highquality" || m == "hq") qm = apg::QualityMode::HighQuality;
            else throw py::value_error("Unrecognized mode '" + m + "'. Valid: 'fast', 'balanced', 'high_quality'");
        } else if (!mode.is_none()) {
            qm = mode.cast<apg::QualityMode>();
        }
        apg::PipelineConfig cfg = apg::PipelineConfig::from_mode(qm, max_disparity);
        cfg.min_disparity = min_disparity;
        cfg.num_threads = threads;
        if (py::isinstance<py::str>(backend)) {
            const std::string b = lower(backend.cast<std::string>());
            if (b == "auto") cfg.volume_backend = apg::VolumeBackend::Auto;
            else if (b == "dense") cfg.volume_backend = apg::VolumeBackend::Dense;
            else if (b == "packed") cfg.volume_backend = apg::VolumeBackend::Packed;
            else throw py::value_error("Unrecognized backend '" + b + "'. Valid: 'auto', 'dense', 'packed'");
        } else if (!backend.is_none()) {
            cfg.volume_backend = backend.cast<apg::VolumeBackend>();
        }
        cfg.hierarchical.enable = hierarchical;
        cfg.hierarchical.radius = hierarchical_radius;
        cfg.hierarchical.adaptive_radius = adaptive_radius;
        cfg.hierarchical.confidence_threshold = confidence_threshold;
        cfg.hierarchical.expanded_radius = expanded_radius;
        if (py::isinstance<py::str>(hierarchical_fallback)) {
            const std::string f = lower(hierarchical_fallback.cast<std::string>());
            if (f == "nearest_valid") cfg.hierarchical.fallback = apg::HierarchicalFallback::NearestValid;
            else if (f == "full_range") cfg.hierarchical.fallback = apg::HierarchicalFallback::FullRange;
            else if (f == "expanded_radius") cfg.hierarchical.fallback = apg::HierarchicalFallback::ExpandedRadius;
            else
                throw py::value_error("Unrecognized hierarchical_fallback '" + f +
                                      "'. Valid: 'nearest_valid', 'full_range', 'expanded_radius'");
        } else if (!hierarchical_fallback.is_none()) {
            cfg.hierarchical.fallback = hierarchical_fallback.cast<apg::HierarchicalFallback>();
        }
        apg::ExecutionOptions opt;
        opt.num_threads = threads;
        if (py::isinstance<py::str>(memory_policy)) {
            const std::string p = lower(memory_policy.cast<std::string>());
            if (p == "release_per_frame") opt.memory_policy = apg::MemoryPolicy::ReleasePerFrame;
            else if (p == "bounded_reuse") opt.memory_policy = apg::MemoryPolicy::BoundedReuse;
            else
                throw py::value_error("Unrecognized memory_policy '" + p +
                                      "'. Valid: 'bounded_reuse', 'release_per_frame'");
        } else if (!memory_policy.is_none()) {
            opt.memory_policy = memory_policy.cast<apg::MemoryPolicy>();
        }
        if (max_retained_bytes > 0) opt.max_retained_bytes = max_retained_bytes;
        const apg::Status st = apg::validate_config(cfg);
        if (!st.ok()) throw std::runtime_error(st.message);
        s_ = std::make_unique<apg::StereoSession>(cfg, opt);
    }

    Result process(const py::buffer& left, const py::buffer& right) {
        HeldView l = hold(left, "left"), r = hold(right, "right");
        if (l.view.width() != r.view.width() || l.view.height() != r.view.height())
            throw py::value_error("Left and right images must have identical dimensions (" +
                                  std::to_string(l.view.width()) + "x" + std::to_string(l.view.height()) + " vs " +
                                  std::to_string(r.view.width()) + "x" + std::to_string(r.view.height()) + ")");
        if (l.view.channels() != r.view.channels())
            throw py::value_error("Left and right images must have matching channel counts (" +
                                  std::to_string(l.view.channels()) + " vs " + std::to_string(r.view.channels()) +
                                  ")");
        apg::StereoResult res;
        apg::PipelineStats stats;
        apg::Status st;
        {
            py::gil_scoped_release nogil;
            std::lock_guard<std::mutex> g(mu_);
            st = s_->process_status(l.view, r.view, res, &stats);
        }
        if (!st.ok())
            throw std::runtime_error("Stereo processing failed [" + std::to_string((int)st.code) + "]: " + st.message);
        return Result(std::move(res), std::move(stats));
    }

    void reset() {
        py::gil_scoped_release nogil;
        std::lock_guard<std::mutex> g(mu_);
        s_->reset();
    }
    template <typename F>
    auto locked(F f) {
        std::lock_guard<std::mutex> g(mu_);
        return f(*s_);
    }
    std::string repr() {
        return locked([](apg::StereoSession& s) {
            std::ostringstream o;
            o << "<apg_sgm.StereoSession frames=" << s.frame_count() << " avg_time_ms=" << s.average_time_ms()
              << " allocated_bytes=" << s.allocated_bytes();
            if (s.config().hierarchical.enable) o << " hierarchical=True (R=" << s.config().hierarchical.radius << ")";
            o << ">";
            return o.str();
        });
    }

private:
    std::unique_ptr<apg::StereoSession> s_;
    std::mutex mu_;
};

py::dict timing_dict(const apg::StageTiming& t) {
    py::dict d;
    d["aux_ms"] = t.aux_ms;
    d["prior_ms"] = t.prior_ms;
    d["cost_ms"] = t.cost_ms;
    d["right_wta_ms"] = t.right_wta_ms;
    d["cross_ms"] = t.cross_ms;
    d["sgm_ms"] = t.sgm_ms;
    d["wta_ms"] = t.wta_ms;
    d["confidence_ms"] = t.confidence_ms;
    d["refine_ms"] = t.refine_ms;
    d["post_ms"] = t.post_ms;
    d["total_ms"] = t.total_ms;
    return d;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "apg-sgm stereo engine on MI355X (gfx950): StereoSession API";

    py::enum_<apg::QualityMode>(m, "QualityMode")
        .value("Fast", apg::QualityMode::Fast)
        .value("Balanced", apg::QualityMode::Balanced)
        .value("HighQuality", apg::QualityMode::HighQuality)
        .export_values();
    py::enum_<apg::VolumeBackend>(m, "VolumeBackend")
        .value("Auto", apg::VolumeBackend::Auto)
        .value("Dense", apg::VolumeBackend::Dense)
        .value("Packed", apg::VolumeBackend::Packed)
        .export_values();
    py::enum_<apg::MemoryPolicy>(m, "MemoryPolicy")
        .value("ReleasePerFrame", apg::MemoryPolicy::ReleasePerFrame)
        .value("BoundedReuse", apg::MemoryPolicy::BoundedReuse)
        .export_values();
    py::enum_<apg::StatusCode>(m, "StatusCode")
        .value("Ok", apg::StatusCode::Ok)
        .value("EmptyInput", apg::StatusCode::EmptyInput)
        .value("SizeMismatch", apg::StatusCode::SizeMismatch)
        .value("InvalidDisparityRange", apg::StatusCode::InvalidDisparityRange)
        .value("InvalidConfig", apg::StatusCode::InvalidConfig)
        .value("UnsupportedBackend", apg::StatusCode::UnsupportedBackend)
        .value("ResourceLimitExceeded", apg::StatusCode::ResourceLimitExceeded)
        .value("AllocationFailure", apg::StatusCode::AllocationFailure)
        .value("InternalError", apg::StatusCode::InternalError)
        .export_values();
    py::enum_<apg::HierarchicalFallback>(m, "HierarchicalFallback")
        .value("NearestValid", apg::HierarchicalFallback::NearestValid)
        .value("FullRange", apg::HierarchicalFallback::FullRange)
        .value("ExpandedRadius", apg::HierarchicalFallback::ExpandedRadius)
        .export_values();

    py::class_<apg::StageTiming>(m, "StageTiming", "Per-stage times in milliseconds")
        .def_readonly("aux_ms", &apg::StageTiming::aux_ms)
        .def_readonly("prior_ms", &apg::StageTiming::prior_ms)
        .def_readonly("cost_ms", &apg::StageTiming::cost_ms)
        .def_readonly("right_wta_ms", &apg::StageTiming::right_wta_ms)
        .def_readonly("cross_ms", &apg::StageTiming::cross_ms)
        .def_readonly("sgm_ms", &apg::StageTiming::sgm_ms)
        .def_readonly("wta_ms", &apg::StageTiming::wta_ms)
        .def_readonly("confidence_ms", &apg::StageTiming::confidence_ms)
        .def_readonly("refine_ms", &apg::StageTiming::refine_ms)
        .def_readonly("post_ms", &apg::StageTiming::post_ms)
        .def_readonly("total_ms", &apg::StageTiming::total_ms)
        .def("to_dict", &timing_dict)
        .def("__repr__", [](const apg::StageTiming& t) {
            std::ostringstream o;
            o << "<StageTiming total=" << t.total_ms << "ms (sgm=" << t.sgm_ms << "ms, cost=" << t.cost_ms << "ms)>";
            return o.str();
        });

    py::class_<apg::PipelineStats>(m, "PipelineStats", "Per-frame diagnostics")
        .def_readonly("timing", &apg::PipelineStats::timing)
        .def_readonly("peak_workspace_bytes", &apg::PipelineStats::peak_workspace_bytes)
        .def_readonly("retained_workspace_bytes", &apg::PipelineStats::retained_workspace_bytes)
        .def_readonly("cost_bytes", &apg::PipelineStats::cost_bytes)
        .def_readonly("estimated_peak_bytes", &apg::PipelineStats::estimated_peak_bytes)
        .def_readonly("valid_disparity_ratio", &apg::PipelineStats::valid_disparity_ratio)
        .def_readonly("reliable_ratio", &apg::PipelineStats::reliable_ratio)
        .def_readonly("unreliable_ratio", &apg::PipelineStats::unreliable_ratio)
        .def_readonly("lr_fail_ratio", &apg::PipelineStats::lr_fail_ratio)
        .def_readonly("effective_threads", &apg::PipelineStats::effective_threads)
        .def_readonly("requested_threads", &apg::PipelineStats::requested_threads)
        .def_readonly("resolved_backend", &apg::PipelineStats::resolved_backend)
        .def_readonly("requested_backend", &apg::PipelineStats::requested_backend)
        .def_readonly("memory_policy", &apg::PipelineStats::memory_policy)
        .def("to_dict",
             [](const apg::PipelineStats& s) {
                 py::dict d;
                 d["timing"] = timing_dict(s.timing);
                 d["peak_workspace_bytes"] = s.peak_workspace_bytes;
                 d["retained_workspace_bytes"] = s.retained_workspace_bytes;
                 d["valid_disparity_ratio"] = s.valid_disparity_ratio;
                 d["reliable_ratio"] = s.reliable_ratio;
                 d["unreliable_ratio"] = s.unreliable_ratio;
                 d["effective_threads"] = s.effective_threads;
                 d["resolved_backend"] = s.resolved_backend;
                 d["memory_policy"] = s.memory_policy;
                 return d;
             })
        .def("__repr__", [](const apg::PipelineStats& s) {
            std::ostringstream o;
            o << "<PipelineStats total=" << s.timing.total_ms << "ms valid_ratio=" << s.valid_disparity_ratio
              << " peak_ws_mb=" << (s.peak_workspace_bytes / (1024.0 * 1024.0)) << ">";
            return o.str();
        });

    py::class_<Result>(m, "StereoResult", "Disparity, confidence, reliable mask and invalid reasons of one frame")
        .def_property_readonly("width", [](const Result& r) { return r.r().width(); })
        .def_property_readonly("height", [](const Result& r) { return r.r().height(); })
        .def_property_readonly("empty", [](const Result& r) { return r.r().empty(); })
        .def_property_readonly("stats", &Result::stats)
        .def_property_readonly("disparity",
                               [](py::object self) { return plane_view<float>(self, self.cast<Result&>().r().disparity); })
        .def_property_readonly("confidence",
                               [](py::object self) { return plane_view<float>(self, self.cast<Result&>().r().confidence); })
        .def_property_readonly("reliable_mask",
                               [](py::object self) {
                                   return plane_view<uint8_t>(self, self.cast<Result&>().r().reliable_mask);
                               })
        .def_property_readonly("invalid_reason",
                               [](py::object self) {
                                   return plane_view<uint8_t>(self, self.cast<Result&>().r().invalid_reason);
                               })
        .def("__repr__", [](const Result& r) {
            std::ostringstream o;
            o << "<apg_sgm.StereoResult " << r.r().width() << "x" << r.r().height()
              << " valid=" << (r.r().empty() ? "False" : "True") << ">";
            return o.str();
        });

    py::class_<Session>(m, "StereoSession", "Stateful stereo matcher session on one MI355X")
        .def(py::init<py::object, int, int, int, py::object, py::object, size_t, bool, int, py::object, bool, float,
                      int>(),
             py::arg("mode") = "balanced", py::arg("threads") = 0, py::arg("max_disparity") = 128,
             py::arg("min_disparity") = 0, py::arg("backend") = "auto", py::arg("memory_policy") = "release_per_frame",
             py::arg("max_retained_bytes") = 0, py::arg("hierarchical") = false, py::arg("hierarchical_radius") = 4,
             py::arg("hierarchical_fallback") = py::none(), py::arg("adaptive_radius") = false,
             py::arg("confidence_threshold") = 0.5f, py::arg("expanded_radius") = 8)
        .def("process", &Session::process, py::arg("left"), py::arg("right"))
        .def("compute", &Session::process, py::arg("left"), py::arg("right"))
        .def("reset", &Session::reset)
        .def_property_readonly("frame_count",
                               [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.frame_count(); }); })
        .def_property_readonly(
            "total_time_ms", [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.total_time_ms(); }); })
        .def_property_readonly(
            "average_time_ms",
            [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.average_time_ms(); }); })
        .def_property_readonly(
            "allocated_bytes",
            [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.allocated_bytes(); }); })
        .def_property_readonly("last_error",
                               [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.last_error(); }); })
        .def_property_readonly(
            "hierarchical_enabled",
            [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.config().hierarchical.enable; }); })
        .def_property_readonly(
            "hierarchical_radius",
            [](Session& s) { return s.locked([](apg::StereoSession& x) { return x.config().hierarchical.radius; }); })
        .def_property_readonly("adaptive_radius",
                               [](Session& s) {
                                   return s.locked(
                                       [](apg::StereoSession& x) { return x.config().hierarchical.adaptive_radius; });
                               })
        .def("__repr__", &Session::repr);
}
