// Session C ABI (include/apg_sgm_c_api.h) over StereoSession.
#include <exception>
#include <new>
#include <string>

#include "apg_sgm/pipeline.hpp"
#include "apg_sgm_c_api.h"

struct apg_session_s {
    apg::StereoSession session;
    std::string last_error;
    apg_session_s(apg::PipelineConfig cfg, apg::ExecutionOptions opt) : session(std::move(cfg), opt) {}
};

struct apg_result_s {
    apg::StereoResult result;
};

extern "C" {

apg_session_t apg_session_create(int quality_mode, int memory_policy) {
    apg_session_options_t o = {};
    o.struct_size = static_cast<uint32_t>(sizeof(apg_session_options_t));
    o.quality_mode = quality_mode;
    o.memory_policy = memory_policy;
    o.max_disparity = 128;
    return apg_session_create_ex(&o);
}

apg_session_t apg_session_create_ex(const apg_session_options_t* o) {
    if (!o) return nullptr;
    if (o->struct_size != 0 && o->struct_size != sizeof(apg_session_options_t)) return nullptr;
    try {
        apg::QualityMode qm = apg::QualityMode::Balanced;
        if (o->quality_mode == APG_QUALITY_FAST) qm = apg::QualityMode::Fast;
        else if (o->quality_mode == APG_QUALITY_HIGH_QUALITY) qm = apg::QualityMode::HighQuality;
        const int max_disp = o->max_disparity > o->min_disparity ? o->max_disparity : 128;
        apg::PipelineConfig cfg = apg::PipelineConfig::from_mode(qm, max_disp);
        cfg.min_disparity = o->min_disparity;
        cfg.max_disparity = max_disp;
        cfg.volume_backend = o->backend == 1 ? apg::VolumeBackend::Dense
                             : o->backend == 2 ? apg::VolumeBackend::Packed : apg::VolumeBackend::Auto;
        cfg.num_threads = o->num_threads;
        apg::ExecutionOptions ex;
        ex.memory_policy = o->memory_policy == APG_MEMORY_BOUNDED_REUSE ? apg::MemoryPolicy::BoundedReuse
                                                                         : apg::MemoryPolicy::ReleasePerFrame;
        ex.max_retained_bytes = o->max_retained_bytes;
        ex.num_threads = o->num_threads;
        return new apg_session_s(std::move(cfg), ex);
    } catch (...) {
        return nullptr;
    }
}

void apg_session_destroy(apg_session_t s) { delete s; }

void apg_session_reset(apg_session_t s) {
    if (!s) return;
    s->session.reset();
    s->last_error.clear();
}

apg_result_t apg_result_create(void) {
    try {
        return new apg_result_s();
    } catch (...) {
        return nullptr;
    }
}

void apg_result_destroy(apg_result_t r) { delete r; }

int apg_session_process(apg_session_t s, const apg_image_view_t* left, const apg_image_view_t* right,
                        apg_result_t result) {
    if (!s || !result) return APG_STATUS_INTERNAL_ERROR;
    if (!left || !right || !left->data || !right->data) {
        s->last_error = "Null input pointer";
        return APG_STATUS_EMPTY_INPUT;
    }
    try {
        apg::ImageView8 l(left->data, left->width, left->height, left->stride, left->channels);
        apg::ImageView8 r(right->data, right->width, right->height, right->stride, right->channels);
        const apg::Status st = s->session.process_status(l, r, result->result);
        if (!st.ok()) {
            s->last_error = st.message;
            return static_cast<int>(st.code);
        }
        s->last_error.clear();
        return APG_STATUS_OK;
    } catch (const std::bad_alloc&) {
        s->last_error = "Memory allocation failure";
        return APG_STATUS_ALLOCATION_FAILURE;
    } catch (const std::exception& e) {
        s->last_error = e.what();
        return APG_STATUS_INTERNAL_ERROR;
    } catch (...) {
        s->last_error = "Unknown error";
        return APG_STATUS_INTERNAL_ERROR;
    }
}

}  // extern "C"

template <typename T, typename Img>
static const T* plane(apg_result_t r, const Img& img, int* w, int* h) {
    if (!r || img.empty()) {
        if (w) *w = 0;
        if (h) *h = 0;
        return nullptr;
    }
    if (w) *w = r->result.width();
    if (h) *h = r->result.height();
    return img.data();
}

extern "C" {

const float* apg_result_get_disparity(apg_result_t r, int* w, int* h) {
    return r ? plane<float>(r, r->result.disparity, w, h) : plane<float>(r, apg::Image32f(), w, h);
}
const float* apg_result_get_confidence(apg_result_t r, int* w, int* h) {
    return r ? plane<float>(r, r->result.confidence, w, h) : plane<float>(r, apg::Image32f(), w, h);
}
const uint8_t* apg_result_get_invalid_reason(apg_result_t r, int* w, int* h) {
    return r ? plane<uint8_t>(r, r->result.invalid_reason, w, h) : plane<uint8_t>(r, apg::Image8u1(), w, h);
}
const uint8_t* apg_result_get_reliable_mask(apg_result_t r, int* w, int* h) {
    return r ? plane<uint8_t>(r, r->result.reliable_mask, w, h) : plane<uint8_t>(r, apg::Image8u1(), w, h);
}

const char* apg_session_get_last_error(apg_session_t s) { return s ? s->last_error.c_str() : "Invalid session handle"; }
size_t apg_session_get_allocated_bytes(apg_session_t s) { return s ? s->session.allocated_bytes() : 0; }
size_t apg_session_get_frame_count(apg_session_t s) { return s ? s->session.frame_count() : 0; }
double apg_session_get_total_time_ms(apg_session_t s) { return s ? s->session.total_time_ms() : 0.0; }
const char* apg_status_string(int code) { return apg::status_code_string(static_cast<apg::StatusCode>(code)); }
const char* apg_version_string(void) { return "1.0.0"; }

}  // extern "C"
