/* Runs the oracle (test infrastructure) over small and awkward inputs, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (`make san` in
 * this directory). Every public function is called at least once: the full
 * detect_and_compute_cpu restatement (src/sift.rs:1200-1240) on images from
 * 1x1 up, with and without the stage outputs, serial and threaded; the blur,
 * Lanczos3 resize and pyramid on odd sizes; the colour conversions and drawing
 * with row padding; the single-keypoint pieces near image borders.
 * Prints "SAN_CHECK_OK" and exits 0 when the run completes; the sanitizers
 * abort the process on any finding.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sift_oracle.h"

static void make_image(uint8_t* img, int w, int h, unsigned seed) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float v = 90.0f + 40.0f * sinf(0.21f * (float)x + 0.05f * (float)seed) *
                                  cosf(0.17f * (float)y);
            for (int b = 0; b < 6; b++) {
                float cx = (float)((b * 53 + (int)seed * 7) % (w > 0 ? w : 1));
                float cy = (float)((b * 29 + (int)seed * 3) % (h > 0 ? h : 1));
                float dx = (float)x - cx, dy = (float)y - cy;
                v += (b % 2 ? 80.0f : -60.0f) * expf(-(dx * dx + dy * dy) / 18.0f);
            }
            v += (float)((x * 131 + y * 71 + (int)seed) % 7);  /* a little texture */
            img[(size_t)y * w + x] = (uint8_t)(v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v));
        }
}

static void detect(const oracle_params_t* p, const uint8_t* img, int w, int h, int threads,
                   int stages) {
    oracle_keypoint_t *cand = NULL, *ref = NULL, *kps = NULL;
    size_t nc = 0, nr = 0, nk = 0;
    float* desc = NULL;
    double ms[6];
    oracle_detect_and_compute(p, img, w, h, threads, stages ? &cand : NULL, stages ? &nc : NULL,
                              stages ? &ref : NULL, stages ? &nr : NULL, &kps, &desc, &nk,
                              stages ? ms : NULL);
    for (size_t i = 0; i < nk; i++) {  /* touch every output */
        float s = 0.0f;
        for (int j = 0; j < 128; j++) s += desc[i * 128 + j] * desc[i * 128 + j];
        if (!(s >= 0.0f) || !isfinite(kps[i].x)) abort();
    }
    printf("  %dx%d threads %d: %zu candidates, %zu refined, %zu keypoints\n", w, h, threads, nc, nr,
           nk);
    oracle_free(cand);
    oracle_free(ref);
    oracle_free(kps);
    oracle_free(desc);
}

int main(void) {
    oracle_params_t p;
    oracle_params_new(&p, 1.6f, 4, 3, 0.5f, 0.04f, 10.0f);
    const int sizes[][2] = {{1, 1}, {2, 3}, {7, 5}, {16, 16}, {97, 61}, {333, 45}, {200, 150}};
    for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; i++) {
        int w = sizes[i][0], h = sizes[i][1];
        uint8_t* img = (uint8_t*)malloc((size_t)w * h);
        make_image(img, w, h, (unsigned)i);
        detect(&p, img, w, h, 1, 1);
        detect(&p, img, w, h, 0, 0);
        /* pyramid, blur and resize on the same odd shapes */
        size_t ng = oracle_pyramid_size(&p, w, h), nd = oracle_dog_size(&p, w, h);
        uint8_t* g = (uint8_t*)malloc(ng ? ng : 1);
        float* dog = (float*)malloc((nd ? nd : 1) * sizeof(float));
        oracle_build_pyramid(&p, img, w, h, g, dog);
        oracle_build_pyramid(&p, img, w, h, g, NULL);
        uint8_t* b = (uint8_t*)malloc((size_t)w * h);
        oracle_gaussian_blur(img, w, h, 2.5f, b);
        int nw = w / 2 > 0 ? w / 2 : 1, nh = h / 2 > 0 ? h / 2 : 1;
        uint8_t* r = (uint8_t*)malloc((size_t)nw * nh);
        oracle_resize_lanczos3(img, w, h, r, nw, nh);
        free(g);
        free(dog);
        free(b);
        free(r);
        free(img);
    }
    /* more octaves than the image supports, 1 and 6 intervals */
    {
        const int w = 120, h = 90;
        uint8_t* img = (uint8_t*)malloc((size_t)w * h);
        make_image(img, w, h, 99);
        oracle_params_t q;
        oracle_params_new(&q, 1.6f, 9, 1, 0.5f, 0.02f, 10.0f);
        detect(&q, img, w, h, 1, 1);
        oracle_params_new(&q, 2.0f, 3, 6, 0.5f, 0.02f, 10.0f);
        detect(&q, img, w, h, 1, 1);
        free(img);
    }
    /* tables and level schedules */
    {
        float taps[64], blur[64], s0;
        int n = oracle_gaussian_kernel(3.0f, NULL, 0);
        if (n > 64 || oracle_gaussian_kernel(3.0f, taps, 64) != n) abort();
        int mt = oracle_lanczos3_table(97, 48, NULL, NULL, NULL, 0);
        int* left = (int*)malloc(48 * sizeof(int));
        int* count = (int*)malloc(48 * sizeof(int));
        float* wts = (float*)malloc((size_t)48 * mt * sizeof(float));
        oracle_lanczos3_table(97, 48, left, count, wts, mt);
        free(left);
        free(count);
        free(wts);
        oracle_level_sigmas(&p, blur, &s0);
        int ws[16], hs[16];
        oracle_octave_shapes(&p, 97, 61, ws, hs);
    }
    /* colour conversion and drawing with padded rows, every format */
    {
        const int w = 33, h = 17;
        for (int f = 0; f <= 4; f++) {
            int bpp = f == 0 ? 1 : f == 1 ? 3 : f == 3 ? 2 : 4;
            size_t stride = (size_t)w * bpp + 5;
            uint8_t* src = (uint8_t*)malloc(stride * h);
            for (size_t k = 0; k < stride * h; k++) src[k] = (uint8_t)(k * 37);
            uint8_t* grey = (uint8_t*)malloc((size_t)w * h);
            oracle_to_luma8(src, w, h, stride, f, grey);
            uint8_t* rgb = (uint8_t*)malloc((size_t)w * h * 3);
            oracle_keypoint_t k[3] = {{3.5f, 4.5f, 2.0f, 0.3f, 0.1f, 0, 1},
                                      {32.9f, 16.9f, 1.0f, -3.0f, 0.1f, 0, 1},
                                      {-20.0f, 40.0f, 1.0f, 1.0f, 0.1f, 0, 1}};
            const uint8_t c[3] = {255, 0, 0};
            oracle_draw_keypoints(src, w, h, stride, f, k, 3, c, rgb);
            free(src);
            free(grey);
            free(rgb);
        }
        float d[256];
        uint8_t u[256];
        for (int k = 0; k < 256; k++) d[k] = (float)k / 300.0f;
        oracle_normalize_and_clip_descriptor(d, 128);
        oracle_descriptor_u8(d, 256, u);
    }
    /* the single-keypoint pieces, centred and at the borders */
    {
        const int w = 48, h = 40;
        uint8_t* g = (uint8_t*)malloc((size_t)w * h);
        make_image(g, w, h, 5);
        const float xy[][2] = {{24.3f, 19.7f}, {0.0f, 0.0f}, {47.0f, 39.0f}, {1.5f, 38.2f}};
        for (int i = 0; i < 4; i++) {
            oracle_keypoint_t kp = {xy[i][0], xy[i][1], 3.2f, 0.7f, 0.02f, 0, 1};
            float angles[64], hist[128], desc[128], oh[36];
            oracle_orientation_single(&p, g, w, h, &kp, angles);
            oracle_orientation_histogram_single(&p, g, w, h, &kp, oh);
            oracle_smooth_orientation_histogram(oh);
            oracle_orientation_peaks(oh, angles);
            oracle_descriptor_raw_single(&p, g, w, h, &kp, hist);
            oracle_descriptor_single(&p, g, w, h, &kp, desc);
            (void)oracle_gauss_pixel_bilinear(g, w, h, xy[i][0] - 0.5f, xy[i][1] + 0.5f);
        }
        const float a[9] = {2, 1, 0, 1, 3, 1, 0, 1, 4}, bb[3] = {1, 2, 3};
        float x[3];
        oracle_solve_linear_system(a, bb, x);
        /* refinement on a one-octave DoG stack */
        const int nd = 5;
        float* dog = (float*)calloc((size_t)nd * w * h, sizeof(float));
        for (int l = 0; l < nd; l++)
            for (int k = 0; k < w * h; k++)
                dog[(size_t)l * w * h + k] = 0.01f * sinf(0.3f * (float)k + (float)l);
        oracle_keypoint_t cand = {20.0f, 18.0f, 1.6f, 0.0f, 0.05f, 0, 2}, out;
        oracle_refine_single(&p, dog, w, h, nd, &cand, &out);
        cand.x = 5.0f;  /* at the border */
        cand.y = 5.0f;
        oracle_refine_single(&p, dog, w, h, nd, &cand, &out);
        free(dog);
        free(g);
    }
    printf("SAN_CHECK_OK\n");
    return 0;
}
