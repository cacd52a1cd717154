/*
 * sift_oracle.c -- TEST INFRASTRUCTURE. A scalar C restatement of the reference
 * CPU SIFT path (stevel705/sift-wgpu, src/sift.rs:1200-1240), used as the parity
 * reference of the tests and, timed, as the CPU baseline. Nothing in the
 * product links it.
 *
 * Build: oracle/Makefile, gcc -O2 -ffp-contract=off (no FMA contraction, SSE
 * float arithmetic, FLT_EVAL_METHOD == 0) so that every f32 operation rounds
 * exactly as Rust's f32 arithmetic on x86-64 does. Transcendentals call the same
 * glibc libm routines Rust std calls on linux-gnu (expf, sinf, cosf, atan2f,
 * powf; sqrtf is the correctly rounded instruction on both).
 *
 * What pins each part (the tests under tests/ carry the checks):
 *  - solve_linear_system: the reference's own known answers
 *    (src/sift.rs:1424-1454).
 *  - get_gauss_pixel_bilinear: the reference's own known answers
 *    (src/sift.rs:1456-1525).
 *  - strict 26-neighbour extremum rule: uniform image -> 0 keypoints
 *    (tests/integration_test.rs:205-226) and brute force on tiny DoG stacks.
 *  - output invariants and determinism: tests/integration_test.rs:13-164.
 *  - gaussian kernel taps, level sigmas, Lanczos3 weights: closed forms.
 *  - imageproc gaussian_blur_f32 (crate 0.25.0, source not available here):
 *    closed forms (impulse response = truncated taps, flat field = truncated
 *    tap sum, separability) and an independent numpy restatement; PARITY OF THE
 *    CRATE'S KERNEL NORMALISATION IS UNPINNED (no reference test or golden
 *    vector fixes it; the restatement uses the crate's unnormalised taps).
 *  - image resize_exact Lanczos3 (crate 0.25.6, source not available here):
 *    flat fields, and PIL's Lanczos resample (same window/normalisation rule,
 *    fixed-point arithmetic) within 1 grey level. Exact rounding is UNPINNED.
 *  - to_luma8 (image 0.25.6 color.rs rgb_to_luma, source not available here):
 *    grey inputs map to themselves (the weights sum to the divisor), the
 *    primaries give the Rec. 709 luma of 255 truncated (54, 182, 18), and every
 *    one of the 2^24 RGB triples lies within [exact - 1, exact] of the real
 *    Rec. 709 value. The integer weights themselves are the crate's published
 *    constants; that the crate truncates (not rounds) is UNPINNED by any
 *    reference test.
 *  - draw_keypoints_to_image (src/lib.rs:70-116; imageproc 0.25.0 drawing,
 *    source not available here): the radius-1 circle is the 5-pixel cross, the
 *    radius-6 circle is symmetric under both reflections and the transposition
 *    and lies between the discs of radius 5.5 and 6.5, lines stay within 1.5 px
 *    of the segment and reach its far end, the 45-degree line from an integer
 *    point is the diagonal, and points outside the image are skipped. The exact pixel choice of the crate
 *    is UNPINNED by any reference test.
 *  - the wasm canvas conversion (src/wasm.rs:88-94, in the reference itself):
 *    all 2^24 triples against exact rational arithmetic on the f32 constants,
 *    which fixes every result except where the f32 sum rounds across an integer.
 */
#define _GNU_SOURCE
#include "sift_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_F 3.14159274101257324219f /* std::f32::consts::PI */

/* --------------------------------------------------------------------------
 * Rust scalar semantics
 * ------------------------------------------------------------------------ */

/* `x as i32` for f32: truncation toward zero, saturating, NaN -> 0. */
static int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

/* `x as u8` / imageproc Clamp<f32> for u8: truncation, saturating at 0 and 255. */
static uint8_t clamp_u8_trunc(float x) {
    if (x < 255.0f) {
        if (x > 0.0f) return (uint8_t)x;
        return 0;
    }
    return 255;
}

/* f32::rem_euclid (core): r = x % m; if r < 0 { r + |m| } else { r }. */
static float rem_euclid_f(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}

static int32_t rem_euclid_i(int32_t x, int32_t m) {
    int32_t r = x % m;
    return r < 0 ? r + m : r;
}

/* f32::powi with a run-time exponent lowers to compiler-rt __powisf2
 * (square-and-multiply). Used for k.powi(s) (src/sift.rs:135). */
static float powi_f(float a, int b) {
    int recip = b < 0;
    float r = 1.0f;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0f / r : r;
}

/* 2.0_f32.powi(o): an exact power of two for the octave range used. */
static float pow2i(int o) { return powi_f(2.0f, o); }

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* --------------------------------------------------------------------------
 * Parameters (src/sift.rs:11-80)
 * ------------------------------------------------------------------------ */

void oracle_params_new(oracle_params_t* p, float sigma, uint32_t num_octaves,
                       uint32_t num_intervals, float assumed_blur,
                       float contrast_threshold, float edge_threshold) {
    p->sigma = sigma;
    p->num_octaves = num_octaves;
    p->num_intervals = num_intervals;
    p->assumed_blur = assumed_blur;
    p->contrast_threshold = contrast_threshold / (float)num_intervals; /* :76 */
    p->edge_threshold = edge_threshold;
    p->image_border_width = 5; /* DEFAULT_IMAGE_BORDER_WIDTH, :17,78 */
}

/* --------------------------------------------------------------------------
 * imageproc 0.25.0 filter::gaussian_blur_f32 (called at src/sift.rs:159,1211)
 *   gaussian(x, r) = (sqrt(2*PI) * r).recip() * (-x.powi(2) / (2 * r.powi(2))).exp()
 *   kernel radius = ceil(2 sigma); taps[r +- i] = gaussian(i, sigma); NOT normalised.
 *   separable_filter: horizontal_filter then vertical_filter, each pass an
 *   Image<Luma<u8>> (Clamp<f32> for u8 after each pass), clamp-to-edge taps,
 *   acc = 0; acc = acc + p * w in tap order.
 * ------------------------------------------------------------------------ */

int oracle_gaussian_kernel(float sigma, float* taps, int cap) {
    int r = (int)ceilf(2.0f * sigma);
    int n = 2 * r + 1;
    if (taps && cap >= n) {
        float norm = 1.0f / (sqrtf(2.0f * PI_F) * sigma);
        float two_s2 = 2.0f * (sigma * sigma);
        for (int i = 0; i <= r; i++) {
            float x = (float)i;
            float v = norm * expf(-(x * x) / two_s2);
            taps[r + i] = v;
            taps[r - i] = v;
        }
    }
    return n;
}

static void hfilter_u8(const uint8_t* src, int w, int h, const float* k, int n, uint8_t* dst) {
    int half = n / 2;
    for (int y = 0; y < h; y++) {
        const uint8_t* row = src + (size_t)y * w;
        for (int x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int i = 0; i < n; i++) {
                int xx = x + i - half;
                xx = xx < 0 ? 0 : (xx > w - 1 ? w - 1 : xx);
                acc = acc + (float)row[xx] * k[i];
            }
            dst[(size_t)y * w + x] = clamp_u8_trunc(acc);
        }
    }
}

static void vfilter_u8(const uint8_t* src, int w, int h, const float* k, int n, uint8_t* dst) {
    int half = n / 2;
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int i = 0; i < n; i++) {
                int yy = y + i - half;
                yy = yy < 0 ? 0 : (yy > h - 1 ? h - 1 : yy);
                acc = acc + (float)src[(size_t)yy * w + x] * k[i];
            }
            dst[(size_t)y * w + x] = clamp_u8_trunc(acc);
        }
    }
}

void oracle_gaussian_blur(const uint8_t* src, int w, int h, float sigma, uint8_t* dst) {
    if (w <= 0 || h <= 0) return;
    int n = oracle_gaussian_kernel(sigma, NULL, 0);
    float* k = (float*)malloc(sizeof(float) * n);
    oracle_gaussian_kernel(sigma, k, n);
    uint8_t* tmp = (uint8_t*)malloc((size_t)w * h);
    hfilter_u8(src, w, h, k, n, tmp);
    vfilter_u8(tmp, w, h, k, n, dst);
    free(tmp);
    free(k);
}

/* --------------------------------------------------------------------------
 * image 0.25.6 imageops::sample::resize with FilterType::Lanczos3
 * (DynamicImage::resize_exact, src/sift.rs:83-88):
 *   vertical_sample (u8 -> f32 intermediate, no rounding), then
 *   horizontal_sample (f32 -> u8 with round(clamp(t, 0, 255))).
 *   Per output index: ratio = in/out; sratio = max(ratio, 1); support = 3*sratio;
 *   c = (j + 0.5) * ratio; left = clamp(floor(c - support), 0, in-1);
 *   right = clamp(ceil(c + support), left+1, in); weights
 *   w_i = lanczos3((i - (c - 0.5)) / sratio), each divided by their f32 sum.
 *   lanczos3(x) = |x| < 3 ? sinc(x) * sinc(x/3) : 0; sinc(t) = t == 0 ? 1 :
 *   sin(t*PI)/(t*PI).
 * ------------------------------------------------------------------------ */

static float sinc_f(float t) {
    float a = t * PI_F;
    if (t == 0.0f) return 1.0f;
    return sinf(a) / a;
}

static float lanczos3_f(float x) {
    if (fabsf(x) < 3.0f) return sinc_f(x) * sinc_f(x / 3.0f);
    return 0.0f;
}

int oracle_lanczos3_table(int in, int out, int* left_o, int* count_o, float* weights, int maxtaps) {
    float ratio = (float)in / (float)out;
    float sratio = ratio < 1.0f ? 1.0f : ratio;
    float support = 3.0f * sratio;
    int need = 0;
    for (int j = 0; j < out; j++) {
        float c = ((float)j + 0.5f) * ratio;
        long long left = (long long)floorf(c - support);
        if (left < 0) left = 0;
        if (left > (long long)in - 1) left = (long long)in - 1;
        long long right = (long long)ceilf(c + support);
        if (right < left + 1) right = left + 1;
        if (right > in) right = in;
        int cnt = (int)(right - left);
        if (cnt > need) need = cnt;
        if (weights && cnt <= maxtaps) {
            float cc = c - 0.5f;
            float sum = 0.0f;
            float* wrow = weights + (size_t)j * maxtaps;
            for (int i = 0; i < cnt; i++) {
                float wv = lanczos3_f(((float)(left + i) - cc) / sratio);
                wrow[i] = wv;
                sum += wv;
            }
            for (int i = 0; i < cnt; i++) wrow[i] /= sum;
            for (int i = cnt; i < maxtaps; i++) wrow[i] = 0.0f;
        }
        if (left_o) left_o[j] = (int)left;
        if (count_o) count_o[j] = cnt;
    }
    return need;
}

void oracle_resize_lanczos3(const uint8_t* src, int w, int h, uint8_t* dst, int nw, int nh) {
    if (nw <= 0 || nh <= 0) return;
    if (w <= 0 || h <= 0) { memset(dst, 0, (size_t)nw * nh); return; }
    if (nw == w && nh == h) { memcpy(dst, src, (size_t)w * h); return; }
    /* vertical pass: w x nh f32 */
    int mv = oracle_lanczos3_table(h, nh, NULL, NULL, NULL, 0);
    int* lv = (int*)malloc(sizeof(int) * nh);
    int* cv = (int*)malloc(sizeof(int) * nh);
    float* wv = (float*)malloc(sizeof(float) * (size_t)nh * mv);
    oracle_lanczos3_table(h, nh, lv, cv, wv, mv);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)w * nh);
    for (int oy = 0; oy < nh; oy++) {
        for (int x = 0; x < w; x++) {
            float t = 0.0f;
            for (int i = 0; i < cv[oy]; i++)
                t += (float)src[(size_t)(lv[oy] + i) * w + x] * wv[(size_t)oy * mv + i];
            tmp[(size_t)oy * w + x] = t;
        }
    }
    /* horizontal pass: nw x nh u8 */
    int mh = oracle_lanczos3_table(w, nw, NULL, NULL, NULL, 0);
    int* lh = (int*)malloc(sizeof(int) * nw);
    int* ch = (int*)malloc(sizeof(int) * nw);
    float* wh = (float*)malloc(sizeof(float) * (size_t)nw * mh);
    oracle_lanczos3_table(w, nw, lh, ch, wh, mh);
    for (int ox = 0; ox < nw; ox++) {
        for (int y = 0; y < nh; y++) {
            float t = 0.0f;
            for (int i = 0; i < ch[ox]; i++)
                t += tmp[(size_t)y * w + lh[ox] + i] * wh[(size_t)ox * mh + i];
            float c = t < 0.0f ? 0.0f : (t > 255.0f ? 255.0f : t);
            dst[(size_t)y * nw + ox] = (uint8_t)roundf(c);
        }
    }
    free(lv); free(cv); free(wv); free(tmp); free(lh); free(ch); free(wh);
}

/* --------------------------------------------------------------------------
 * Pyramid (src/sift.rs:129-177) and DoG (src/sift.rs:92-124,182-204)
 * ------------------------------------------------------------------------ */

void oracle_level_sigmas(const oracle_params_t* p, float* blur, float* initial_sigma) {
    int S = (int)p->num_intervals;
    int L = S + 3;
    /* :1205-1209 */
    if (initial_sigma) {
        float s0 = 0.0f;
        if (p->sigma > p->assumed_blur)
            s0 = sqrtf(p->sigma * p->sigma - p->assumed_blur * p->assumed_blur);
        *initial_sigma = s0;
    }
    if (!blur) return;
    float k = powf(2.0f, 1.0f / (float)S); /* :132 */
    float tgt[64];
    for (int s = 0; s < L && s < 64; s++) tgt[s] = p->sigma * powi_f(k, s); /* :133-136 */
    for (uint32_t o = 0; o < p->num_octaves; o++) {
        float prev = p->sigma * pow2i((int)o); /* :141 */
        blur[o * L + 0] = 0.0f;
        for (int s = 1; s < L; s++) {
            float t = tgt[s] * pow2i((int)o); /* :144-145 */
            blur[o * L + s] = sqrtf(t * t - prev * prev); /* :146-147 */
            prev = t;
        }
    }
}

void oracle_octave_shapes(const oracle_params_t* p, int w, int h, int* ws, int* hs) {
    for (uint32_t o = 0; o < p->num_octaves; o++) {
        ws[o] = w;
        hs[o] = h;
        w /= 2;
        h /= 2;
    }
}

size_t oracle_pyramid_size(const oracle_params_t* p, int w, int h) {
    size_t t = 0;
    for (uint32_t o = 0; o < p->num_octaves; o++) {
        t += (size_t)(p->num_intervals + 3) * w * h;
        w /= 2; h /= 2;
    }
    return t;
}

size_t oracle_dog_size(const oracle_params_t* p, int w, int h) {
    size_t t = 0;
    for (uint32_t o = 0; o < p->num_octaves; o++) {
        t += (size_t)(p->num_intervals + 2) * w * h;
        w /= 2; h /= 2;
    }
    return t;
}

typedef struct {
    int n_oct, L, S;
    int* w;
    int* h;
    uint8_t** g;  /* g[o*L + s] */
    float** d;    /* d[o*(L-1) + s] */
} pyr_t;

static void pyr_views(const oracle_params_t* p, int w, int h, uint8_t* gauss, float* dog, pyr_t* P) {
    P->n_oct = (int)p->num_octaves;
    P->S = (int)p->num_intervals;
    P->L = P->S + 3;
    P->w = (int*)malloc(sizeof(int) * P->n_oct);
    P->h = (int*)malloc(sizeof(int) * P->n_oct);
    P->g = (uint8_t**)malloc(sizeof(uint8_t*) * P->n_oct * P->L);
    P->d = (float**)malloc(sizeof(float*) * P->n_oct * (P->L - 1));
    oracle_octave_shapes(p, w, h, P->w, P->h);
    size_t go = 0, dgo = 0;
    for (int o = 0; o < P->n_oct; o++) {
        size_t a = (size_t)P->w[o] * P->h[o];
        for (int s = 0; s < P->L; s++) { P->g[o * P->L + s] = gauss + go; go += a; }
        for (int s = 0; s < P->L - 1; s++) { P->d[o * (P->L - 1) + s] = dog ? dog + dgo : NULL; dgo += a; }
    }
}

static void pyr_free(pyr_t* P) { free(P->w); free(P->h); free(P->g); free(P->d); }

static void build_gauss(const oracle_params_t* p, const uint8_t* img, int w, int h, pyr_t* P) {
    int L = P->L, S = P->S;
    float* blur = (float*)malloc(sizeof(float) * P->n_oct * L);
    float s0;
    oracle_level_sigmas(p, blur, &s0);
    /* initial blur (:1205-1214) writes octave 0 level 0 directly */
    if (s0 > 1e-4f) oracle_gaussian_blur(img, w, h, s0, P->g[0]);
    else memcpy(P->g[0], img, (size_t)w * h);
    for (int o = 0; o < P->n_oct; o++) {
        int ww = P->w[o], hh = P->h[o];
        size_t a = (size_t)ww * hh;
        for (int s = 1; s < L; s++) {
            float b = blur[o * L + s];
            if (b < 1e-4f) memcpy(P->g[o * L + s], P->g[o * L + s - 1], a); /* :149-150 */
            else oracle_gaussian_blur(P->g[o * L + s - 1], ww, hh, b, P->g[o * L + s]);
        }
        if (o < P->n_oct - 1) /* :166-173: next base = resize(level S, w/2, h/2) */
            oracle_resize_lanczos3(P->g[o * L + S], ww, hh, P->g[(o + 1) * L], ww / 2, hh / 2);
    }
    free(blur);
}

static void build_dog(pyr_t* P) {
    int L = P->L;
    for (int o = 0; o < P->n_oct; o++) {
        size_t a = (size_t)P->w[o] * P->h[o];
        for (int s = 0; s < L - 1; s++) {
            const uint8_t* g1 = P->g[o * L + s + 1];
            const uint8_t* g0 = P->g[o * L + s];
            float* d = P->d[o * (L - 1) + s];
            #pragma omp parallel for schedule(static) if (a > 65536)
            for (long long i = 0; i < (long long)a; i++)
                d[i] = (float)g1[i] / 255.0f - (float)g0[i] / 255.0f;
        }
    }
}

void oracle_build_pyramid(const oracle_params_t* p, const uint8_t* img, int w, int h,
                          uint8_t* gauss, float* dog) {
    pyr_t P;
    pyr_views(p, w, h, gauss, dog, &P);
    build_gauss(p, img, w, h, &P);
    if (dog) build_dog(&P);
    pyr_free(&P);
}

/* --------------------------------------------------------------------------
 * Pixel samplers (src/sift.rs:207-266)
 * ------------------------------------------------------------------------ */

static inline float dog_px(const float* d, int w, int h, int x, int y) {
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    return d[(size_t)y * w + x];
}

static inline float gauss_px(const uint8_t* g, int w, int h, int x, int y) {
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    return (float)g[(size_t)y * w + x] / 255.0f;
}

/* get_gauss_pixel_bilinear (src/sift.rs:238-266) */
static float gauss_bilinear(const uint8_t* g, int w, int h, float x, float y) {
    float xf = floorf(x), yf = floorf(y);
    float xc = xf + 1.0f, yc = yf + 1.0f;
    float dx = x - xf, dy = y - yf;
    int x0 = f2i_sat(xf), y0 = f2i_sat(yf), x1 = f2i_sat(xc), y1 = f2i_sat(yc);
    float q11 = gauss_px(g, w, h, x0, y0);
    float q21 = gauss_px(g, w, h, x1, y0);
    float q12 = gauss_px(g, w, h, x0, y1);
    float q22 = gauss_px(g, w, h, x1, y1);
    return q11 * (1.0f - dx) * (1.0f - dy) + q21 * dx * (1.0f - dy) + q12 * (1.0f - dx) * dy +
           q22 * dx * dy;
}

float oracle_gauss_pixel_bilinear(const uint8_t* img, int w, int h, float x, float y) {
    return gauss_bilinear(img, w, h, x, y);
}

/* --------------------------------------------------------------------------
 * solve_linear_system (src/sift.rs:478-503): Cramer's rule, |det| < 1e-10 -> None
 * ------------------------------------------------------------------------ */

static int solve3(const float a[3][3], const float b[3], float x[3]) {
    float det_a = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                  a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                  a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (fabsf(det_a) < 1e-10f) return 0;
    float det_x = b[0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                  a[0][1] * (b[1] * a[2][2] - a[1][2] * b[2]) +
                  a[0][2] * (b[1] * a[2][1] - a[1][1] * b[2]);
    float det_y = a[0][0] * (b[1] * a[2][2] - a[1][2] * b[2]) -
                  b[0] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                  a[0][2] * (a[1][0] * b[2] - b[1] * a[2][0]);
    float det_z = a[0][0] * (a[1][1] * b[2] - b[1] * a[2][1]) -
                  a[0][1] * (a[1][0] * b[2] - b[1] * a[2][0]) +
                  b[0] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    x[0] = det_x / det_a;
    x[1] = det_y / det_a;
    x[2] = det_z / det_a;
    return 1;
}

int oracle_solve_linear_system(const float a[9], const float b[3], float x[3]) {
    float m[3][3] = {{a[0], a[1], a[2]}, {a[3], a[4], a[5]}, {a[6], a[7], a[8]}};
    return solve3(m, b, x);
}

/* --------------------------------------------------------------------------
 * find_scale_space_extrema (src/sift.rs:510-603)
 * Order: octave -> layer -> row -> column (rayon collect keeps it).
 * ------------------------------------------------------------------------ */

typedef struct {
    oracle_keypoint_t* v;
    size_t n, cap;
} kpvec_t;

static void kv_push(kpvec_t* k, oracle_keypoint_t kp) {
    if (k->n == k->cap) {
        k->cap = k->cap ? k->cap * 2 : 64;
        k->v = (oracle_keypoint_t*)realloc(k->v, sizeof(oracle_keypoint_t) * k->cap);
    }
    k->v[k->n++] = kp;
}

static void kv_append(kpvec_t* dst, const kpvec_t* src) {
    for (size_t i = 0; i < src->n; i++) kv_push(dst, src->v[i]);
}

static void find_extrema(const oracle_params_t* p, const pyr_t* P, kpvec_t* out) {
    int S = P->S, nd = P->L - 1;
    int border = (int)p->image_border_width;
    for (int o = 0; o < P->n_oct; o++) {
        int w = P->w[o], h = P->h[o];
        float scale = pow2i(o);
        for (int s = 1; s <= S; s++) {
            if (!(s < nd - 1)) continue; /* :533 */
            const float* dp = P->d[o * nd + s - 1];
            const float* dc = P->d[o * nd + s];
            const float* dn = P->d[o * nd + s + 1];
            int y0 = border, y1 = h - border;
            int nrows = y1 > y0 ? y1 - y0 : 0;
            kpvec_t* rows = (kpvec_t*)calloc(nrows ? nrows : 1, sizeof(kpvec_t));
            #pragma omp parallel for schedule(dynamic, 4)
            for (int yi = 0; yi < nrows; yi++) {
                int y = y0 + yi;
                for (int x = border; x < w - border; x++) {
                    float val = dog_px(dc, w, h, x, y);
                    int is_max = 1, is_min = 1;
                    for (int dz = -1; dz <= 1 && (is_max || is_min); dz++) {
                        const float* img = dz < 0 ? dp : (dz == 0 ? dc : dn);
                        for (int dy = -1; dy <= 1 && (is_max || is_min); dy++) {
                            for (int dx = -1; dx <= 1; dx++) {
                                if (dz == 0 && dy == 0 && dx == 0) continue;
                                float nv = dog_px(img, w, h, x + dx, y + dy);
                                if (val <= nv) is_max = 0;
                                if (val >= nv) is_min = 0;
                                if (!is_max && !is_min) break;
                            }
                        }
                    }
                    if (is_max || is_min) {
                        oracle_keypoint_t kp;
                        kp.x = ((float)x + 0.5f) * scale;
                        kp.y = ((float)y + 0.5f) * scale;
                        kp.size = 0.0f;
                        kp.angle = 0.0f;
                        kp.response = val;
                        kp.octave = o;
                        kp.layer = s;
                        kv_push(&rows[yi], kp);
                    }
                }
            }
            for (int yi = 0; yi < nrows; yi++) { kv_append(out, &rows[yi]); free(rows[yi].v); }
            free(rows);
        }
    }
}

/* --------------------------------------------------------------------------
 * refine_and_filter_extrema (src/sift.rs:299-474); sequential over candidates
 * ------------------------------------------------------------------------ */

/* One candidate: dl[0..nd) are the octave's DoG layers (w x h each). Returns 1
 * and writes the refined keypoint when it passes, 0 when it is dropped. */
static int refine_one(const oracle_params_t* p, const float* const* dl, int nd, int w, int h,
                      const oracle_keypoint_t* kp, oracle_keypoint_t* outp) {
    float xi = kp->x / pow2i(kp->octave); /* :309-310 */
    float yi = kp->y / pow2i(kp->octave);
    int cx = f2i_sat(xi), cy = f2i_sat(yi), cl = kp->layer;
    for (int step = 0; step < 5; step++) { /* MAX_INTERPOLATION_STEPS */
        if (cl < 1 || cl >= nd - 1) break; /* :323 */
        const float* ip = dl[cl - 1];
        const float* ic = dl[cl];
        const float* in = dl[cl + 1];
        if (cx < 1 || cx >= w - 1 || cy < 1 || cy >= h - 1) break; /* :330-336 */
        float dx = (dog_px(ic, w, h, cx + 1, cy) - dog_px(ic, w, h, cx - 1, cy)) / 2.0f;
        float dy = (dog_px(ic, w, h, cx, cy + 1) - dog_px(ic, w, h, cx, cy - 1)) / 2.0f;
        float ds = (dog_px(in, w, h, cx, cy) - dog_px(ip, w, h, cx, cy)) / 2.0f;
        float cv = dog_px(ic, w, h, cx, cy);
        float dxx = dog_px(ic, w, h, cx + 1, cy) + dog_px(ic, w, h, cx - 1, cy) - 2.0f * cv;
        float dyy = dog_px(ic, w, h, cx, cy + 1) + dog_px(ic, w, h, cx, cy - 1) - 2.0f * cv;
        float dss = dog_px(in, w, h, cx, cy) + dog_px(ip, w, h, cx, cy) - 2.0f * cv;
        float dxy = (dog_px(ic, w, h, cx + 1, cy + 1) - dog_px(ic, w, h, cx - 1, cy + 1) -
                     dog_px(ic, w, h, cx + 1, cy - 1) + dog_px(ic, w, h, cx - 1, cy - 1)) /
                    4.0f;
        float dxs = (dog_px(in, w, h, cx + 1, cy) - dog_px(in, w, h, cx - 1, cy) -
                     (dog_px(ip, w, h, cx + 1, cy) - dog_px(ip, w, h, cx - 1, cy))) /
                    4.0f;
        float dys = (dog_px(in, w, h, cx, cy + 1) - dog_px(in, w, h, cx, cy - 1) -
                     (dog_px(ip, w, h, cx, cy + 1) - dog_px(ip, w, h, cx, cy - 1))) /
                    4.0f;
        float H[3][3] = {{dxx, dxy, dxs}, {dxy, dyy, dys}, {dxs, dys, dss}};
        float b[3] = {-dx, -dy, -ds};
        float off[3];
        if (!solve3(H, b, off)) break; /* :460-463 */
        if (fabsf(off[0]) < 0.5f && fabsf(off[1]) < 0.5f && fabsf(off[2]) < 0.5f) {
            float val = cv + 0.5f * (dx * off[0] + dy * off[1] + ds * off[2]); /* :394-398 */
            if (fabsf(val) < p->contrast_threshold) break;                     /* :401 */
            float tr = dxx + dyy;
            float trace_sq = tr * tr;
            float det = dxx * dyy - dxy * dxy;
            if (det <= 0.0f) break; /* :411 */
            float ratio = trace_sq / det;
            float e1 = p->edge_threshold + 1.0f;
            float thr = e1 * e1 / p->edge_threshold;
            if (ratio >= thr) break; /* :420 */
            float scale = pow2i(kp->octave);
            float fl = (float)cl + off[2];
            float psig = p->sigma * powf(2.0f, (float)kp->octave + fl / (float)p->num_intervals);
            oracle_keypoint_t r;
            r.x = ((float)cx + off[0]) * scale;
            r.y = ((float)cy + off[1]) * scale;
            r.size = psig;
            r.angle = 0.0f;
            r.response = val;
            r.octave = kp->octave;
            r.layer = f2i_sat(roundf(fl));
            *outp = r;
            return 1;
        } else {
            cx = f2i_sat(roundf((float)cx + off[0])); /* :451-453 */
            cy = f2i_sat(roundf((float)cy + off[1]));
            cl = f2i_sat(roundf((float)cl + off[2]));
            if (cl < 0 || cl >= nd) break; /* :456 */
        }
    }
    return 0;
}

static void refine(const oracle_params_t* p, const pyr_t* P, const kpvec_t* cand, kpvec_t* out) {
    int nd = P->L - 1;
    for (size_t ci = 0; ci < cand->n; ci++) {
        const oracle_keypoint_t* kp = &cand->v[ci];
        int o = kp->octave;
        oracle_keypoint_t r;
        if (refine_one(p, (const float* const*)&P->d[o * nd], nd, P->w[o], P->h[o], kp, &r))
            kv_push(out, r);
    }
}

int oracle_refine_single(const oracle_params_t* p, const float* dog, int w, int h, int nd,
                         const oracle_keypoint_t* cand, oracle_keypoint_t* out) {
    const float* dl[64];
    if (nd < 3 || nd > 64) return 0;
    for (int l = 0; l < nd; l++) dl[l] = dog + (size_t)l * w * h;
    return refine_one(p, dl, nd, w, h, cand, out);
}

/* --------------------------------------------------------------------------
 * assign_orientations (src/sift.rs:605-726)
 * ------------------------------------------------------------------------ */

/* The raw 36-bin histogram of one keypoint (src/sift.rs:614-671): 0 when the
 * keypoint has no valid level or size. */
static int orient_histogram(const oracle_params_t* p, const pyr_t* P, const oracle_keypoint_t* kp,
                            float* hist /* 36 */) {
    (void)p;
    int S = P->S, L = P->L;
    int o = kp->octave;
    int gl = kp->layer < 0 ? 0 : (kp->layer > S + 2 ? S + 2 : kp->layer); /* :618 */
    memset(hist, 0, 36 * sizeof(float));
    if (o < 0 || o >= P->n_oct || gl >= L) return 0;
    const uint8_t* g = P->g[o * L + gl];
    int w = P->w[o], h = P->h[o];
    float scale = pow2i(kp->octave);
    float xo = kp->x / scale, yo = kp->y / scale, so = kp->size / scale;
    if (so <= 0.0f) return 0;
    int R = f2i_sat(roundf(4.5f * so)); /* (3.0 * 1.5 * sigma_octave).round() */
    float ws = 1.5f * so;
    float denom = 2.0f * ws * ws;
    for (int dy = -R; dy <= R; dy++) {
        for (int dx = -R; dx <= R; dx++) {
            int xi = f2i_sat(roundf(xo + (float)dx));
            int yi = f2i_sat(roundf(yo + (float)dy));
            if (xi < 1 || xi >= w - 1 || yi < 1 || yi >= h - 1) continue;
            float gx = gauss_px(g, w, h, xi + 1, yi) - gauss_px(g, w, h, xi - 1, yi);
            float gy = gauss_px(g, w, h, xi, yi + 1) - gauss_px(g, w, h, xi, yi - 1);
            float mag = sqrtf(gx * gx + gy * gy);
            float ang = atan2f(gy, gx);
            float fdx = (float)dx, fdy = (float)dy;
            float wt = expf(-(fdx * fdx + fdy * fdy) / denom);
            float an = ang < 0.0f ? ang + 2.0f * PI_F : ang;
            float bf = an * 36.0f / (2.0f * PI_F);
            float fb = floorf(bf);
            /* `as usize`: saturating, NaN -> 0 */
            unsigned long long bu = (fb != fb || fb <= 0.0f) ? 0ULL : (unsigned long long)fb;
            int bin = (int)(bu % 36ULL);
            hist[bin] += mag * wt;
        }
    }
    return 1;
}

/* Two passes of the circular [1, 1, 1] / 3 smoothing (src/sift.rs:677-685). */
void oracle_smooth_orientation_histogram(float* hist /* 36 */) {
    for (int it = 0; it < 2; it++) { /* ORIENTATION_SMOOTHING_ITERATIONS */
        float prev[36];
        memcpy(prev, hist, sizeof prev);
        for (int i = 0; i < 36; i++)
            hist[i] = (prev[(i + 35) % 36] + prev[i] + prev[(i + 1) % 36]) / 3.0f;
    }
}

/* Peaks of a smoothed histogram (src/sift.rs:688-720): bins >= 0.8 max that are
 * strictly above both circular neighbours, with a parabolic fit; angles in
 * (-pi, pi], in ascending bin order. Returns their count. */
int oracle_orientation_peaks(const float* hist /* 36 */, float* angles /* 36 */) {
    float mx = 0.0f;
    for (int i = 0; i < 36; i++) mx = fmaxf(mx, hist[i]);
    float thr = mx * 0.8f;
    int n = 0;
    const float step = 2.0f * PI_F / 36.0f;
    for (int i = 0; i < 36; i++) {
        float c = hist[i];
        if (c >= thr) {
            float pv = hist[(i + 35) % 36], nv = hist[(i + 1) % 36];
            if (c > pv && c > nv) {
                float den = pv - 2.0f * c + nv;
                float off = fabsf(den) > 1e-5f ? 0.5f * (pv - nv) / den : 0.0f;
                float center = ((float)i + 0.5f) * step;
                float a = center + off * step;
                float fa = rem_euclid_f(a, 2.0f * PI_F);
                if (fa > PI_F) fa = fa - 2.0f * PI_F;
                angles[n++] = fa;
            }
        }
    }
    return n;
}

static int orient_one(const oracle_params_t* p, const pyr_t* P, const oracle_keypoint_t* kp,
                      oracle_keypoint_t* outv /* capacity 36 */) {
    float hist[36], ang[36];
    if (!orient_histogram(p, P, kp, hist)) return 0;
    oracle_smooth_orientation_histogram(hist);
    int n = oracle_orientation_peaks(hist, ang);
    for (int i = 0; i < n; i++) {
        outv[i] = *kp;
        outv[i].angle = ang[i];
    }
    return n;
}

/* --------------------------------------------------------------------------
 * normalize_and_clip_descriptor (src/sift.rs:889-913) and compute (917-1060)
 * ------------------------------------------------------------------------ */

void oracle_normalize_and_clip_descriptor(float* d, int n) {
    float ss = 0.0f;
    for (int i = 0; i < n; i++) ss += d[i] * d[i];
    float norm = sqrtf(ss);
    if (norm < 1e-8f) return;
    float inv = 1.0f / norm;
    float ss2 = 0.0f;
    for (int i = 0; i < n; i++) {
        d[i] *= inv;
        d[i] = fminf(d[i], 0.2f);
        ss2 += d[i] * d[i];
    }
    float n2 = sqrtf(ss2);
    if (n2 < 1e-8f) return;
    float inv2 = 1.0f / n2;
    for (int i = 0; i < n; i++) d[i] *= inv2;
}

/* The 4 x 4 x 8 histogram of compute (src/sift.rs:927-1052), before
 * normalize_and_clip_descriptor. */
static void descriptor_raw(const oracle_params_t* p, const pyr_t* P, const oracle_keypoint_t* kp,
                           float* hist /* 128, zeroed */) {
    int S = P->S, L = P->L;
    int o = kp->octave;
    int gl = kp->layer < 0 ? 0 : (kp->layer > S + 2 ? S + 2 : kp->layer); /* :931 */
    if (o < 0 || o >= P->n_oct || gl >= L) return;
    const uint8_t* g = P->g[o * L + gl];
    int w = P->w[o], h = P->h[o];
    float scale = pow2i(kp->octave);
    float xo = kp->x / scale, yo = kp->y / scale, so = kp->size / scale;
    if (so <= 0.0f) return;
    float angle = kp->angle;
    float ca = cosf(angle), sa = sinf(angle);
    float bw = 3.0f * so;          /* DESC_PATCH_SCALE_FACTOR */
    float win = bw * 4.0f;         /* DESC_WINDOW_WIDTH */
    float wsig = 0.5f * win;
    float denom = 2.0f * wsig * wsig;
    int R = f2i_sat(ceilf(win * sqrtf(2.0f) * 0.5f)); /* :960 */
    float wm1 = (float)(w - 1), hm1 = (float)(h - 1);
    for (int dyi = -R; dyi <= R; dyi++) {
        for (int dxi = -R; dxi <= R; dxi++) {
            float px = (float)dxi, py = (float)dyi;
            float rx = ca * px + sa * py;
            float ry = -sa * px + ca * py;
            float xb = rx / bw + 4.0f / 2.0f - 0.5f;
            float yb = ry / bw + 4.0f / 2.0f - 0.5f;
            if (!(xb > -1.0f && xb < 4.0f && yb > -1.0f && yb < 4.0f)) continue;
            float xs = xo + px, ys = yo + py;
            if (xs < 0.0f || xs >= wm1 || ys < 0.0f || ys >= hm1) continue;
            float gx = gauss_bilinear(g, w, h, xs + 1.0f, ys) - gauss_bilinear(g, w, h, xs - 1.0f, ys);
            float gy = gauss_bilinear(g, w, h, xs, ys + 1.0f) - gauss_bilinear(g, w, h, xs, ys - 1.0f);
            float mag = sqrtf(gx * gx + gy * gy);
            float pa = atan2f(gy, gx);
            float ar = rem_euclid_f(pa - angle, 2.0f * PI_F);
            float wt = expf(-(px * px + py * py) / denom);
            float wm = mag * wt;
            float ab = ar * 8.0f / (2.0f * PI_F);
            int xbi = f2i_sat(floorf(xb)), ybi = f2i_sat(floorf(yb)), abi = f2i_sat(floorf(ab));
            float dxi_ = xb - (float)xbi, dyi_ = yb - (float)ybi, dai = ab - (float)abi;
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    for (int k = 0; k < 2; k++) {
                        int ix = xbi + i, iy = ybi + j;
                        int ia = rem_euclid_i(abi + k, 8);
                        if (ix >= 0 && ix < 4 && iy >= 0 && iy < 4) {
                            float wx = i == 0 ? 1.0f - dxi_ : dxi_;
                            float wy = j == 0 ? 1.0f - dyi_ : dyi_;
                            float wa = k == 0 ? 1.0f - dai : dai;
                            float c = wm * wx * wy * wa;
                            hist[(iy * 4 + ix) * 8 + ia] += c;
                        }
                    }
        }
    }
}

static void descriptor_one(const oracle_params_t* p, const pyr_t* P, const oracle_keypoint_t* kp,
                           float* hist /* 128, zeroed */) {
    descriptor_raw(p, P, kp, hist);
    oracle_normalize_and_clip_descriptor(hist, 128); /* :1054 */
}

/* --------------------------------------------------------------------------
 * detect_and_compute_cpu (src/sift.rs:1200-1240)
 * ------------------------------------------------------------------------ */

int oracle_detect_and_compute(const oracle_params_t* p, const uint8_t* img, int w, int h,
                              int nthreads,
                              oracle_keypoint_t** cand_o, size_t* n_cand,
                              oracle_keypoint_t** refined_o, size_t* n_refined,
                              oracle_keypoint_t** kps_o, float** desc_o, size_t* n_kps,
                              double* stage_ms) {
#ifdef _OPENMP
    int prev_threads = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double t0 = now_ms();
    size_t gsz = oracle_pyramid_size(p, w, h), dsz = oracle_dog_size(p, w, h);
    uint8_t* gauss = (uint8_t*)malloc(gsz ? gsz : 1);
    float* dog = (float*)malloc(sizeof(float) * (dsz ? dsz : 1));
    pyr_t P;
    pyr_views(p, w, h, gauss, dog, &P);
    build_gauss(p, img, w, h, &P);
    double t1 = now_ms();
    build_dog(&P);
    double t2 = now_ms();
    kpvec_t cand = {0}, ref = {0};
    find_extrema(p, &P, &cand);
    double t3 = now_ms();
    refine(p, &P, &cand, &ref);
    double t4 = now_ms();

    /* orientation: per keypoint in parallel, flat_map keeps order */
    size_t nr = ref.n;
    oracle_keypoint_t* per = (oracle_keypoint_t*)malloc(sizeof(oracle_keypoint_t) * 36 * (nr ? nr : 1));
    int* cnt = (int*)malloc(sizeof(int) * (nr ? nr : 1));
    #pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < (long long)nr; i++) cnt[i] = orient_one(p, &P, &ref.v[i], per + 36 * i);
    size_t nk = 0;
    for (size_t i = 0; i < nr; i++) nk += (size_t)cnt[i];
    oracle_keypoint_t* kps = (oracle_keypoint_t*)malloc(sizeof(oracle_keypoint_t) * (nk ? nk : 1));
    size_t pos = 0;
    for (size_t i = 0; i < nr; i++)
        for (int j = 0; j < cnt[i]; j++) kps[pos++] = per[36 * i + j];
    free(per);
    free(cnt);
    double t5 = now_ms();

    float* desc = (float*)calloc(nk ? nk * 128 : 1, sizeof(float));
    #pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < (long long)nk; i++) descriptor_one(p, &P, &kps[i], desc + 128 * i);
    double t6 = now_ms();

    if (stage_ms) {
        stage_ms[0] = t1 - t0; stage_ms[1] = t2 - t1; stage_ms[2] = t3 - t2;
        stage_ms[3] = t4 - t3; stage_ms[4] = t5 - t4; stage_ms[5] = t6 - t5;
    }
    if (cand_o) *cand_o = cand.v; else free(cand.v);
    if (n_cand) *n_cand = cand.n;
    if (refined_o) *refined_o = ref.v; else free(ref.v);
    if (n_refined) *n_refined = ref.n;
    if (kps_o) *kps_o = kps; else free(kps);
    if (desc_o) *desc_o = desc; else free(desc);
    if (n_kps) *n_kps = nk;
    pyr_free(&P);
    free(gauss);
    free(dog);
#ifdef _OPENMP
    omp_set_num_threads(prev_threads);
#endif
    return 0;
}

/* Single-keypoint access to assign_orientations / compute for closed-form
 * tests: every level of a one-octave pyramid is the given u8 image. */
static void single_level_pyr(const oracle_params_t* p, const uint8_t* g, int w, int h, pyr_t* P,
                             uint8_t** gp, int* ws, int* hs) {
    P->n_oct = 1;
    P->S = (int)p->num_intervals;
    P->L = P->S + 3;
    ws[0] = w;
    hs[0] = h;
    P->w = ws;
    P->h = hs;
    for (int s = 0; s < P->L; s++) gp[s] = (uint8_t*)g;
    P->g = gp;
    P->d = NULL;
}

int oracle_orientation_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                              const oracle_keypoint_t* kp, float* angles /* 36 */) {
    pyr_t P;
    uint8_t* gp[64];
    int ws[1], hs[1];
    single_level_pyr(p, g, w, h, &P, gp, ws, hs);
    oracle_keypoint_t out[36];
    oracle_keypoint_t k = *kp;
    k.octave = 0;
    int n = orient_one(p, &P, &k, out);
    for (int i = 0; i < n; i++) angles[i] = out[i].angle;
    return n;
}

int oracle_orientation_histogram_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                                        const oracle_keypoint_t* kp, float* hist /* 36 */) {
    pyr_t P;
    uint8_t* gp[64];
    int ws[1], hs[1];
    single_level_pyr(p, g, w, h, &P, gp, ws, hs);
    oracle_keypoint_t k = *kp;
    k.octave = 0;
    return orient_histogram(p, &P, &k, hist);
}

void oracle_descriptor_raw_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                                  const oracle_keypoint_t* kp, float* hist /* 128 */) {
    pyr_t P;
    uint8_t* gp[64];
    int ws[1], hs[1];
    single_level_pyr(p, g, w, h, &P, gp, ws, hs);
    oracle_keypoint_t k = *kp;
    k.octave = 0;
    memset(hist, 0, 128 * sizeof(float));
    descriptor_raw(p, &P, &k, hist);
}

void oracle_descriptor_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                              const oracle_keypoint_t* kp, float* desc /* 128 */) {
    pyr_t P;
    uint8_t* gp[64];
    int ws[1], hs[1];
    single_level_pyr(p, g, w, h, &P, gp, ws, hs);
    oracle_keypoint_t k = *kp;
    k.octave = 0;
    memset(desc, 0, 128 * sizeof(float));
    descriptor_one(p, &P, &k, desc);
}

void oracle_free(void* ptr) { free(ptr); }

/* --------------------------------------------------------------------------
 * Colour ingest and the WebGPU output convention
 * ------------------------------------------------------------------------ */

/* image 0.25.6 color.rs rgb_to_luma for u8: the u32 sum
 * 2126 R + 7152 G + 722 B divided by 10000 (integer division, truncating);
 * Rgba and LumaA drop alpha. Canvas RGBA: src/wasm.rs:88-94,
 * ((0.299 r + 0.587 g) + 0.114 b) in f32 (Rust evaluates left to right, no
 * FMA), then `luma as u8`. */
void oracle_to_luma8(const uint8_t* src, int w, int h, size_t stride, int format, uint8_t* dst) {
    const int bpp = format == 1 ? 3 : format == 3 ? 2 : format == 0 ? 1 : 4;
    for (int y = 0; y < h; y++) {
        const uint8_t* row = src + (size_t)y * stride;
        for (int x = 0; x < w; x++) {
            const uint8_t* p = row + (size_t)x * bpp;
            uint8_t l;
            if (format == 0 || format == 3) {
                l = p[0];
            } else if (format == 4) {
                const float r = (float)p[0], g = (float)p[1], b = (float)p[2];
                const float v = 0.299f * r + 0.587f * g + 0.114f * b;
                l = clamp_u8_trunc(v);
            } else {
                l = (uint8_t)((2126u * p[0] + 7152u * p[1] + 722u * p[2]) / 10000u);
            }
            dst[(size_t)y * w + x] = l;
        }
    }
}

/* descriptor.wgsl:258-261: u32(clamp(normalized * 512.0, 0.0, 255.0)); the
 * conversion to u32 truncates. */
void oracle_descriptor_u8(const float* desc, size_t n, uint8_t* out) {
    for (size_t i = 0; i < n; i++) {
        float v = desc[i] * 512.0f;
        v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
        out[i] = (uint8_t)v;
    }
}

/* imageproc 0.25.0 drawing: BresenhamLineIter over f32 end points (steep
 * lines swap axes, the walk runs from the smaller x; error starts at dx/2 and
 * steps y when it goes negative), points outside the canvas skipped;
 * draw_filled_circle_mut: midpoint circle (p = 1 - r), four horizontal
 * segments per step. */
static void draw_px(uint8_t* rgb, int w, int h, const uint8_t c[3], int x, int y) {
    if (x < 0 || y < 0 || x >= w || y >= h) return;
    uint8_t* p = rgb + ((size_t)y * w + x) * 3;
    p[0] = c[0];
    p[1] = c[1];
    p[2] = c[2];
}

static void draw_line(uint8_t* rgb, int w, int h, const uint8_t c[3], float x0, float y0, float x1,
                      float y1) {
    int steep = fabsf(y1 - y0) > fabsf(x1 - x0);
    float t;
    if (steep) {
        t = x0; x0 = y0; y0 = t;
        t = x1; x1 = y1; y1 = t;
    }
    if (x0 > x1) {
        t = x0; x0 = x1; x1 = t;
        t = y0; y0 = y1; y1 = t;
    }
    float dx = x1 - x0, dy = fabsf(y1 - y0), err = dx / 2.0f;
    int x = f2i_sat(x0), y = f2i_sat(y0), end_x = f2i_sat(x1), ystep = y0 < y1 ? 1 : -1;
    while (x <= end_x) {
        if (steep) draw_px(rgb, w, h, c, y, x);
        else draw_px(rgb, w, h, c, x, y);
        x++;
        err -= dy;
        if (err < 0.0f) {
            y += ystep;
            err += dx;
        }
    }
}

static void draw_disc(uint8_t* rgb, int w, int h, const uint8_t c[3], int x0, int y0, int r) {
    int x = 0, y = r, p = 1 - r;
    while (x <= y) {
        draw_line(rgb, w, h, c, (float)(x0 - x), (float)(y0 + y), (float)(x0 + x), (float)(y0 + y));
        draw_line(rgb, w, h, c, (float)(x0 - y), (float)(y0 + x), (float)(x0 + y), (float)(y0 + x));
        draw_line(rgb, w, h, c, (float)(x0 - x), (float)(y0 - y), (float)(x0 + x), (float)(y0 - y));
        draw_line(rgb, w, h, c, (float)(x0 - y), (float)(y0 - x), (float)(x0 + y), (float)(y0 - x));
        x++;
        if (p < 0) {
            p += 2 * x + 1;
        } else {
            y--;
            p += 2 * (x - y) + 1;
        }
    }
}

void oracle_draw_keypoints(const uint8_t* src, int w, int h, size_t stride, int format,
                           const oracle_keypoint_t* kps, size_t n, const uint8_t c[3],
                           uint8_t* rgb) {
    const int bpp = format == 1 ? 3 : format == 3 ? 2 : format == 0 ? 1 : 4;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t* p = src + (size_t)y * stride + (size_t)x * bpp;
            uint8_t* d = rgb + ((size_t)y * w + x) * 3;
            if (bpp <= 2) {
                d[0] = d[1] = d[2] = p[0];
            } else {
                d[0] = p[0];
                d[1] = p[1];
                d[2] = p[2];
            }
        }
    for (size_t i = 0; i < n; i++) {
        const float x = kps[i].x, y = kps[i].y;
        const float size = 2.0f; /* src/lib.rs:87 */
        int radius = f2i_sat(roundf(size * 3.0f));
        int disp = radius > 2 ? radius : 2;
        draw_disc(rgb, w, h, c, f2i_sat(x), f2i_sat(y), disp);
        float xe = x + (float)disp * cosf(kps[i].angle);
        float ye = y + (float)disp * sinf(kps[i].angle);
        draw_line(rgb, w, h, c, x, y, xe, ye);
        draw_disc(rgb, w, h, c, f2i_sat(x), f2i_sat(y), 1);
    }
}
