/* Pure C99 client of the session C ABI (include/apg_sgm_c_api.h), in the
 * spirit of the reference's tests/c_api_smoke.c: lifecycle, error codes,
 * struct_size validation, a strided 3-channel frame and result access.
 * Prints the disparity plane as raw floats to argv[1] for the Python side to
 * compare with the oracle. Exit 0 = all contract checks passed. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "apg_sgm_c_api.h"

#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); return 1; } } while (0)

int main(int argc, char** argv) {
    const int w = 64, h = 40, ch = 3, stride = w * ch + 16;
    uint8_t* left = (uint8_t*)calloc((size_t)stride * h, 1);
    uint8_t* right = (uint8_t*)calloc((size_t)stride * h, 1);
    int x, y, c;
    for (y = 0; y < h; ++y)
        for (x = 0; x < w; ++x)
            for (c = 0; c < ch; ++c) {
                const int v = (x * 37 + y * 73 + x * y * 13 + c * 30) % 256;
                left[y * stride + x * ch + c] = (uint8_t)v;
                right[y * stride + x * ch + c] = (uint8_t)(((x + 4) * 37 + y * 73 + (x + 4) * y * 13 + c * 30) % 256);
            }
    apg_session_options_t bad;
    memset(&bad, 0, sizeof(bad));
    bad.struct_size = 3;
    CHECK(apg_session_create_ex(&bad) == NULL);
    CHECK(apg_session_create_ex(NULL) == NULL);

    apg_session_options_t o;
    memset(&o, 0, sizeof(o));
    o.struct_size = (uint32_t)sizeof(o);
    o.quality_mode = APG_QUALITY_BALANCED;
    o.memory_policy = APG_MEMORY_BOUNDED_REUSE;
    o.max_disparity = 24;
    apg_session_t s = apg_session_create_ex(&o);
    apg_result_t r = apg_result_create();
    CHECK(s && r);

    apg_image_view_t lv = {left, w, h, stride, ch}, rv = {right, w, h, stride, ch};
    apg_image_view_t empty = {NULL, w, h, stride, ch};
    CHECK(apg_session_process(s, &empty, &rv, r) == APG_STATUS_EMPTY_INPUT);
    apg_image_view_t badstride = {left, w, h, w, ch};
    CHECK(apg_session_process(s, &badstride, &rv, r) == APG_STATUS_SIZE_MISMATCH);
    CHECK(apg_session_get_frame_count(s) == 0);
    CHECK(strlen(apg_session_get_last_error(s)) > 0);

    int rc = apg_session_process(s, &lv, &rv, r);
    if (rc != APG_STATUS_OK) {
        fprintf(stderr, "process: %s (%s)\n", apg_status_string(rc), apg_session_get_last_error(s));
        return 1;
    }
    CHECK(apg_session_get_frame_count(s) == 1);
    CHECK(apg_session_get_allocated_bytes(s) > 0); /* BoundedReuse keeps the device workspace */
    int ow = 0, oh = 0;
    const float* d = apg_result_get_disparity(r, &ow, &oh);
    CHECK(d && ow == w && oh == h);
    CHECK(apg_result_get_confidence(r, &ow, &oh) != NULL);
    CHECK(apg_result_get_invalid_reason(r, &ow, &oh) != NULL);
    CHECK(apg_result_get_reliable_mask(r, &ow, &oh) != NULL);
    CHECK(apg_result_get_disparity(NULL, &ow, &oh) == NULL && ow == 0 && oh == 0);
    if (argc > 1) {
        FILE* f = fopen(argv[1], "wb");
        CHECK(f);
        fwrite(d, sizeof(float), (size_t)w * h, f);
        fclose(f);
    }
    /* second frame reuses the workspace */
    CHECK(apg_session_process(s, &lv, &rv, r) == APG_STATUS_OK);
    CHECK(apg_session_get_frame_count(s) == 2);
    apg_session_reset(s);
    CHECK(apg_session_get_frame_count(s) == 0 && apg_session_get_allocated_bytes(s) == 0);

    /* ReleasePerFrame leaves nothing allocated after a frame */
    apg_session_t s2 = apg_session_create(APG_QUALITY_FAST, APG_MEMORY_RELEASE_PER_FRAME);
    CHECK(s2);
    CHECK(apg_session_process(s2, &lv, &rv, r) == APG_STATUS_OK);
    CHECK(apg_session_get_allocated_bytes(s2) == 0);
    CHECK(strcmp(apg_version_string(), "1.0.0") == 0);
    CHECK(strcmp(apg_status_string(6), "ResourceLimitExceeded") == 0);

    /* budget gate: a tiny max_retained_bytes rejects before compute */
    o.max_retained_bytes = 1024;
    apg_session_t s3 = apg_session_create_ex(&o);
    CHECK(apg_session_process(s3, &lv, &rv, r) == APG_STATUS_RESOURCE_LIMIT_EXCEEDED);
    CHECK(apg_result_get_disparity(r, &ow, &oh) == NULL);

    apg_session_destroy(s);
    apg_session_destroy(s2);
    apg_session_destroy(s3);
    apg_result_destroy(r);
    free(left);
    free(right);
    printf("c_api_smoke ok\n");
    return 0;
}
