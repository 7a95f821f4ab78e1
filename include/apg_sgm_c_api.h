/*
 * apg_sgm_c_api.h -- the session C ABI of apg-sgm, exported by
 * libapg_sgm_amd.so with the same symbols, types and values as the reference
 * (reference include/apg_sgm/c_api.h:23-108), so existing C/FFI bindings link
 * unchanged.  Behind it every frame runs on the MI355X: the input views are
 * uploaded once, the dense pipeline runs in HBM and the four result planes
 * are copied back into the result handle.
 *
 * Contract (reference src/c_api.cpp:10-204): opaque handles owned by the
 * library; borrowed strided input views (stride >= width*channels, 1 or 3
 * channels); result arrays valid until the next process/destroy; int status
 * codes, never an exception across the ABI; a failed frame does not count.
 */
#ifndef APG_SGM_C_API_H
#define APG_SGM_C_API_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__) && __GNUC__ >= 4
#define APG_API __attribute__((visibility("default")))
#else
#define APG_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define APG_QUALITY_FAST 0
#define APG_QUALITY_BALANCED 1
#define APG_QUALITY_HIGH_QUALITY 2

#define APG_MEMORY_RELEASE_PER_FRAME 0
#define APG_MEMORY_BOUNDED_REUSE 1

typedef enum {
    APG_STATUS_OK = 0,
    APG_STATUS_EMPTY_INPUT = 1,
    APG_STATUS_SIZE_MISMATCH = 2,
    APG_STATUS_INVALID_DISPARITY_RANGE = 3,
    APG_STATUS_INVALID_CONFIG = 4,
    APG_STATUS_UNSUPPORTED_BACKEND = 5,
    APG_STATUS_RESOURCE_LIMIT_EXCEEDED = 6,
    APG_STATUS_ALLOCATION_FAILURE = 7,
    APG_STATUS_INTERNAL_ERROR = 8
} apg_status_code_t;

typedef struct {
    const uint8_t* data;
    int width;
    int height;
    int stride;
    int channels;
} apg_image_view_t;

typedef struct apg_session_s* apg_session_t;
typedef struct apg_result_s* apg_result_t;

/* Frozen layout (reference include/apg_sgm/c_api.h:63-72); struct_size 0 or sizeof. */
typedef struct {
    uint32_t struct_size;
    int quality_mode;
    int memory_policy;
    size_t max_retained_bytes;
    int num_threads;
    int backend;          /* 0 Auto, 1 Dense, 2 Packed */
    int min_disparity;
    int max_disparity;    /* <= min_disparity selects 128 */
} apg_session_options_t;

/* reference src/c_api.cpp:22-33 */
APG_API apg_session_t apg_session_create(int quality_mode, int memory_policy);
/* reference src/c_api.cpp:35-74 */
APG_API apg_session_t apg_session_create_ex(const apg_session_options_t* options);
/* reference src/c_api.cpp:76-85 */
APG_API void apg_session_destroy(apg_session_t session);
APG_API void apg_session_reset(apg_session_t session);
/* reference src/c_api.cpp:87-97 */
APG_API apg_result_t apg_result_create(void);
APG_API void apg_result_destroy(apg_result_t result);
/* reference src/c_api.cpp:99-132 (StereoSession::process_status) */
APG_API int apg_session_process(apg_session_t session, const apg_image_view_t* left,
                                const apg_image_view_t* right, apg_result_t result);
/* reference src/c_api.cpp:134-176 */
APG_API const float* apg_result_get_disparity(apg_result_t result, int* width, int* height);
APG_API const float* apg_result_get_confidence(apg_result_t result, int* width, int* height);
APG_API const uint8_t* apg_result_get_invalid_reason(apg_result_t result, int* width, int* height);
APG_API const uint8_t* apg_result_get_reliable_mask(apg_result_t result, int* width, int* height);
/* reference src/c_api.cpp:178-204 */
APG_API const char* apg_session_get_last_error(apg_session_t session);
APG_API size_t apg_session_get_allocated_bytes(apg_session_t session);
APG_API size_t apg_session_get_frame_count(apg_session_t session);
APG_API double apg_session_get_total_time_ms(apg_session_t session);
APG_API const char* apg_status_string(int status_code);
APG_API const char* apg_version_string(void);

#ifdef __cplusplus
}
#endif

#endif /* APG_SGM_C_API_H */
