/* The call sequence of the reference's HIP backend arm, replayed in C.
 *
 * The arm (SiftBackend::Hip in src/sift.rs:1249, built on src/lib.rs:18-49) is
 * Rust, which this machine cannot compile; this program makes the same calls
 * in the same order with the same structs:
 *   sift_params_default -> sift_hip_create(&p, 0, &ctx)          (once per process)
 *   per image: sift_params_t from the Sift fields (contrast already / intervals,
 *   border 5), sift_color_image_t { data, w, h, stride 0, format } for Luma8,
 *   Rgb8, Rgba8 and LumaA8 buffers, sift_hip_detect_batch_color(ctx, &img, &p, 1,
 *   &r), then the keypoints and descriptors are read and sift_result_free(&r);
 *   errors are read with sift_hip_last_error.
 * Each result is checked against the grey path (sift_hip_to_luma8, then
 * sift_hip_detect_batch of the grey image), bit for bit, and the argument
 * checks are exercised (null pointers, empty batches, bad formats and params),
 * as is the host-only drawing routine. Built once plainly and once with
 * AddressSanitizer and UndefinedBehaviorSanitizer on the host code of the
 * library and of this program.
 *
 * Exit status 0 and the last line "ARM_REPLAY_OK" on success;
 * "ARM_REPLAY_NO_DEVICE" (status 0) when the checks that need no GPU passed and
 * no GPU is visible; any failure prints "FAIL: ..." and exits 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sift_hip.h"

static int g_failed = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            printf("FAIL: %s:%d: ", __FILE__, __LINE__);   \
            printf(__VA_ARGS__);                           \
            printf("\n");                                  \
            fflush(stdout);                                \
            g_failed = 1;                                  \
        }                                                  \
    } while (0)

/* Deterministic test picture: smooth blobs and a few bars on a gradient. */
static void make_rgb(uint8_t* rgb, int w, int h) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float v = 40.0f + 60.0f * (float)x / (float)w;
            for (int b = 0; b < 24; b++) {
                float cx = (float)((b * 97) % w), cy = (float)((b * 61 + 13) % h);
                float r = 6.0f + (float)(b % 5) * 3.0f;
                float dx = (float)x - cx, dy = (float)y - cy;
                v += (b % 2 ? 90.0f : -70.0f) * expf(-(dx * dx + dy * dy) / (2.0f * r * r));
            }
            if ((x / 37) % 5 == 0 && (y / 53) % 3 == 1) v += 50.0f;
            int base = v < 0.0f ? 0 : (v > 255.0f ? 255 : (int)v);
            uint8_t* p = rgb + 3 * ((size_t)y * w + x);
            p[0] = (uint8_t)base;
            p[1] = (uint8_t)((base * 7 + x) & 255);
            p[2] = (uint8_t)(255 - base);
        }
}

static size_t bpp_of(uint32_t fmt) {
    switch (fmt) {
        case SIFT_PIXEL_GRAY8: return 1;
        case SIFT_PIXEL_RGB8: return 3;
        case SIFT_PIXEL_LUMAA8: return 2;
        default: return 4;
    }
}

/* The pixel buffer the arm would hand over for one DynamicImage variant. */
static uint8_t* make_format(const uint8_t* rgb, const uint8_t* grey, int w, int h, uint32_t fmt) {
    size_t bpp = bpp_of(fmt), n = (size_t)w * h;
    uint8_t* out = (uint8_t*)malloc(n * bpp);
    for (size_t i = 0; i < n; i++) {
        switch (fmt) {
            case SIFT_PIXEL_GRAY8: out[i] = grey[i]; break;
            case SIFT_PIXEL_RGB8: memcpy(out + 3 * i, rgb + 3 * i, 3); break;
            case SIFT_PIXEL_LUMAA8:
                out[2 * i] = grey[i];
                out[2 * i + 1] = (uint8_t)(i * 31);
                break;
            default:
                memcpy(out + 4 * i, rgb + 3 * i, 3);
                out[4 * i + 3] = (uint8_t)(255 - (i & 127));
                break;
        }
    }
    return out;
}

/* The arm's params: the Sift fields, contrast as stored (already / intervals). */
static sift_params_t arm_params(float sigma, uint32_t octaves, uint32_t intervals, float blur,
                                float contrast, float edge) {
    sift_params_t p;
    p.sigma = sigma;
    p.num_octaves = octaves;
    p.num_intervals = intervals;
    p.assumed_blur = blur;
    p.contrast_threshold = contrast / (float)intervals;  /* Sift::new stores it divided */
    p.edge_threshold = edge;
    p.image_border_width = 5;
    return p;
}

static int same_result(const sift_result_t* a, const sift_result_t* b) {
    if (a->count != b->count) return 0;
    if (a->count == 0) return 1;
    return memcmp(a->keypoints, b->keypoints, a->count * sizeof(sift_keypoint_t)) == 0 &&
           memcmp(a->descriptors, b->descriptors, a->count * 128 * sizeof(float)) == 0;
}

static void check_result_sane(const sift_result_t* r, int w, int h) {
    for (size_t i = 0; i < r->count; i++) {
        const sift_keypoint_t* k = &r->keypoints[i];
        CHECK(isfinite(k->x) && isfinite(k->y) && k->x >= 0.0f && k->y >= 0.0f && k->x < (float)w &&
                  k->y < (float)h,
              "keypoint %zu at (%f, %f)", i, k->x, k->y);
        CHECK(k->size > 0.0f && isfinite(k->angle), "keypoint %zu size %f angle %f", i, k->size,
              k->angle);
        double ss = 0.0;
        for (int j = 0; j < 128; j++) {
            float d = r->descriptors[i * 128 + j];
            CHECK(d >= 0.0f && d <= 1.0f, "descriptor %zu[%d] = %f", i, j, d);
            ss += (double)d * d;
        }
        CHECK(ss == 0.0 || fabs(ss - 1.0) < 1e-4, "descriptor %zu squared norm %f", i, ss);
    }
}

/* Checks that need no device: parameter helpers, argument validation before
 * any device work, and the host-only drawing routine. */
static void host_checks(void) {
    sift_params_t p;
    sift_params_default(&p);
    CHECK(p.sigma == 1.6f && p.num_octaves == 4 && p.num_intervals == 3 && p.assumed_blur == 0.5f &&
              p.edge_threshold == 10.0f && p.image_border_width == 5,
          "sift_params_default fields");
    CHECK(p.contrast_threshold == 0.04f / 3.0f, "stored contrast %.9g", p.contrast_threshold);
    sift_params_t q;
    sift_params_new(&q, 1.6f, 5, 4, 0.5f, 0.03f, 12.0f);
    CHECK(q.num_octaves == 5 && q.num_intervals == 4 && q.contrast_threshold == 0.03f / 4.0f &&
              q.edge_threshold == 12.0f,
          "sift_params_new fields");

    sift_hip_ctx* ctx = NULL;
    CHECK(sift_hip_create(NULL, 0, &ctx) == SIFT_ERR_INVALID && ctx == NULL, "create(NULL params)");
    CHECK(strlen(sift_hip_last_error()) > 0, "last_error after a failed create");
    CHECK(sift_hip_create(&p, 0, NULL) == SIFT_ERR_INVALID, "create(NULL out)");
    sift_params_t bad = p;
    bad.num_intervals = 0;
    CHECK(sift_hip_create(&bad, 0, &ctx) == SIFT_ERR_INVALID && ctx == NULL, "create(0 intervals)");
    bad = p;
    bad.sigma = -1.0f;
    CHECK(sift_hip_create(&bad, 0, &ctx) == SIFT_ERR_INVALID && ctx == NULL, "create(sigma < 0)");
    CHECK(sift_hip_detect_batch_color(NULL, NULL, NULL, 1, NULL) == SIFT_ERR_INVALID,
          "detect_batch_color(NULL ctx)");
    sift_result_t empty = {NULL, NULL, 0};
    sift_result_free(&empty);  /* freeing an empty result is a no-op */
    sift_result_free(NULL);

    /* drawing (host code only): a few keypoints on a small RGB image */
    const int w = 64, h = 48;
    uint8_t* rgb = (uint8_t*)malloc((size_t)w * h * 3);
    uint8_t* out = (uint8_t*)malloc((size_t)w * h * 3);
    make_rgb(rgb, w, h);
    sift_color_image_t im = {rgb, (uint32_t)w, (uint32_t)h, 0, SIFT_PIXEL_RGB8};
    sift_keypoint_t kps[3] = {{10.5f, 12.25f, 2.0f, 0.5f, 0.1f, 0, 1},
                              {63.9f, 47.9f, 3.0f, -2.0f, 0.1f, 1, 2},
                              {-5.0f, 100.0f, 1.0f, 3.1f, 0.1f, 0, 1}};  /* partly outside */
    const uint8_t red[3] = {255, 0, 0};
    CHECK(sift_draw_keypoints(&im, kps, 3, red, out) == SIFT_OK, "draw_keypoints");
    size_t changed = 0;
    for (size_t i = 0; i < (size_t)w * h * 3; i += 3)
        if (out[i] == 255 && out[i + 1] == 0 && out[i + 2] == 0) changed++;
    CHECK(changed > 20, "drawing changed %zu pixels", changed);
    CHECK(sift_draw_keypoints(&im, kps, 3, red, NULL) == SIFT_ERR_INVALID, "draw(NULL out)");
    free(rgb);
    free(out);
}

static void device_checks(sift_hip_ctx* ctx) {
    const int w = 640, h = 480;
    uint8_t* rgb = (uint8_t*)malloc((size_t)w * h * 3);
    uint8_t* grey = (uint8_t*)malloc((size_t)w * h);
    make_rgb(rgb, w, h);
    sift_color_image_t rgb_img = {rgb, (uint32_t)w, (uint32_t)h, 0, SIFT_PIXEL_RGB8};
    CHECK(sift_hip_to_luma8(ctx, &rgb_img, grey) == SIFT_OK, "to_luma8: %s", sift_hip_last_error());

    const sift_params_t params[2] = {arm_params(1.6f, 4, 3, 0.5f, 0.04f, 10.0f),
                                     arm_params(1.6f, 5, 4, 0.5f, 0.03f, 10.0f)};
    const uint32_t formats[4] = {SIFT_PIXEL_GRAY8, SIFT_PIXEL_RGB8, SIFT_PIXEL_RGBA8,
                                 SIFT_PIXEL_LUMAA8};
    for (int pi = 0; pi < 2; pi++) {
        /* the grey path */
        sift_image_t gimg = {grey, (uint32_t)w, (uint32_t)h, 0};
        sift_result_t ref = {NULL, NULL, 0};
        int rc = sift_hip_detect_batch(ctx, &gimg, &params[pi], 1, &ref);
        CHECK(rc == SIFT_OK, "detect_batch grey: %d %s", rc, sift_hip_last_error());
        CHECK(ref.count > 50, "grey path found %zu keypoints", ref.count);
        check_result_sane(&ref, w, h);
        for (int f = 0; f < 4; f++) {
            uint8_t* buf = make_format(rgb, grey, w, h, formats[f]);
            /* exactly the arm's call: stride 0, n = 1, the params it builds */
            sift_color_image_t image = {buf, (uint32_t)w, (uint32_t)h, 0, formats[f]};
            sift_result_t r = {NULL, NULL, 0};
            rc = sift_hip_detect_batch_color(ctx, &image, &params[pi], 1, &r);
            CHECK(rc == SIFT_OK, "detect_batch_color format %u: %d %s", formats[f], rc,
                  sift_hip_last_error());
            /* RGBA8 goes through to_luma8 (alpha ignored), LumaA8 keeps L: all
             * equal the grey path of the same luma */
            CHECK(same_result(&r, &ref), "params %d format %u: %zu keypoints vs %zu on the grey path",
                  pi, formats[f], r.count, ref.count);
            sift_result_free(&r);
            CHECK(r.keypoints == NULL && r.descriptors == NULL && r.count == 0,
                  "result_free clears the struct");
            free(buf);
        }
        sift_result_free(&ref);
    }

    /* errors come back as codes with a message; nothing is substituted */
    sift_result_t r = {NULL, NULL, 0};
    sift_color_image_t im = {NULL, (uint32_t)w, (uint32_t)h, 0, SIFT_PIXEL_RGB8};
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 1, &r) == SIFT_ERR_INVALID,
          "null pixels");
    CHECK(strlen(sift_hip_last_error()) > 0, "message for null pixels");
    im.data = rgb;
    im.format = 9;
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 1, &r) == SIFT_ERR_INVALID,
          "unknown pixel format");
    im.format = SIFT_PIXEL_RGB8;
    im.width = 0;
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 1, &r) == SIFT_ERR_INVALID, "width 0");
    im.width = (uint32_t)w;
    im.stride = 5;  /* below width * 3 */
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 1, &r) == SIFT_ERR_INVALID,
          "stride below a row");
    im.stride = 0;
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 0, &r) == SIFT_ERR_INVALID, "n = 0");
    sift_params_t bad = params[0];
    bad.num_octaves = 0;
    CHECK(sift_hip_detect_batch_color(ctx, &im, &bad, 1, &r) == SIFT_ERR_INVALID, "0 octaves");
    CHECK(r.count == 0 && r.keypoints == NULL, "no result on error");
    /* the context still works after the errors */
    CHECK(sift_hip_detect_batch_color(ctx, &im, &params[0], 1, &r) == SIFT_OK, "after errors: %s",
          sift_hip_last_error());
    CHECK(r.count > 50, "after errors: %zu keypoints", r.count);
    sift_result_free(&r);

    /* a tiny image and a 1-pixel-wide one: valid, no keypoints expected */
    uint8_t tiny[3 * 7 * 5];
    memset(tiny, 128, sizeof tiny);
    sift_color_image_t t = {tiny, 7, 5, 0, SIFT_PIXEL_RGB8};
    CHECK(sift_hip_detect_batch_color(ctx, &t, &params[0], 1, &r) == SIFT_OK, "7x5: %s",
          sift_hip_last_error());
    CHECK(r.count == 0, "7x5 flat image: %zu keypoints", r.count);
    sift_result_free(&r);
    free(rgb);
    free(grey);
}

int main(void) {
    host_checks();
    sift_params_t p;
    sift_params_default(&p);
    sift_hip_ctx* ctx = NULL;
    int rc = sift_hip_create(&p, 0, &ctx);
    if (rc == SIFT_ERR_DEVICE) {
        printf("%s\n", g_failed ? "FAIL" : "ARM_REPLAY_NO_DEVICE");
        return g_failed;
    }
    CHECK(rc == SIFT_OK && ctx != NULL, "create: %d %s", rc, sift_hip_last_error());
    if (rc == SIFT_OK) {
        device_checks(ctx);
        sift_hip_destroy(ctx);
    }
    printf("%s\n", g_failed ? "FAIL" : "ARM_REPLAY_OK");
    return g_failed;
}
