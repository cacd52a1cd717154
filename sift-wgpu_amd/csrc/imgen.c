/*
 * imgen.c -- seeded synthetic grey-scale test images (host C, no GPU).
 *
 * Generator used by the tests and the benchmark for every workload:
 *   background 0.5; add B Gaussian blobs (centre U[0,W)xU[0,H), sigma U[2,20] px,
 *   amplitude U[-0.4,0.4], evaluated within +-ceil(4 sigma)); then paint R
 *   axis-aligned rectangles (top-left U[0,W)xU[0,H), width U[4,W/8], height
 *   U[4,H/8], intensity U[0,1], overwrite); then add i.i.d. N(0, noise_sigma)
 *   (Box-Muller); clamp to [0,1]; u8 = floor(255 x + 0.5).
 * RNG: SplitMix64; uniform doubles use the top 53 bits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct { uint64_t s; } sm64_t;

static uint64_t sm64_next(sm64_t* r) {
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double sm64_unit(sm64_t* r) { return (double)(sm64_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static double uni(sm64_t* r, double a, double b) { return a + (b - a) * sm64_unit(r); }

int sift_synth_image(int w, int h, int n_blobs, int n_rects, double noise_sigma, uint64_t seed,
                     uint8_t* out) {
    if (w <= 0 || h <= 0 || !out) return -1;
    sm64_t r = {seed * 0x2545F4914F6CDD1DULL + 0x1234567ULL};
    double* f = (double*)malloc(sizeof(double) * (size_t)w * h);
    if (!f) return -2;
    for (size_t i = 0; i < (size_t)w * h; i++) f[i] = 0.5;
    for (int b = 0; b < n_blobs; b++) {
        double cx = uni(&r, 0.0, (double)w), cy = uni(&r, 0.0, (double)h);
        double s = uni(&r, 2.0, 20.0), a = uni(&r, -0.4, 0.4);
        int rad = (int)ceil(4.0 * s);
        int x0 = (int)floor(cx) - rad, x1 = (int)floor(cx) + rad;
        int y0 = (int)floor(cy) - rad, y1 = (int)floor(cy) + rad;
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > w - 1) x1 = w - 1;
        if (y1 > h - 1) y1 = h - 1;
        double inv = 1.0 / (2.0 * s * s);
        for (int y = y0; y <= y1; y++) {
            double dy = (double)y - cy;
            for (int x = x0; x <= x1; x++) {
                double dx = (double)x - cx;
                f[(size_t)y * w + x] += a * exp(-(dx * dx + dy * dy) * inv);
            }
        }
    }
    for (int k = 0; k < n_rects; k++) {
        int x0 = (int)uni(&r, 0.0, (double)w), y0 = (int)uni(&r, 0.0, (double)h);
        double wmax = (double)w / 8.0, hmax = (double)h / 8.0;
        int rw = (int)uni(&r, 4.0, wmax > 4.0 ? wmax : 4.0);
        int rh = (int)uni(&r, 4.0, hmax > 4.0 ? hmax : 4.0);
        double v = uni(&r, 0.0, 1.0);
        for (int y = y0; y < y0 + rh && y < h; y++)
            for (int x = x0; x < x0 + rw && x < w; x++) f[(size_t)y * w + x] = v;
    }
    if (noise_sigma > 0.0) {
        size_t n = (size_t)w * h;
        for (size_t i = 0; i < n; i += 2) {
            double u1 = sm64_unit(&r), u2 = sm64_unit(&r);
            if (u1 < 1e-300) u1 = 1e-300;
            double rr = sqrt(-2.0 * log(u1));
            f[i] += noise_sigma * rr * cos(6.283185307179586 * u2);
            if (i + 1 < n) f[i + 1] += noise_sigma * rr * sin(6.283185307179586 * u2);
        }
    }
    for (size_t i = 0; i < (size_t)w * h; i++) {
        double v = f[i] < 0.0 ? 0.0 : (f[i] > 1.0 ? 1.0 : f[i]);
        out[i] = (uint8_t)floor(255.0 * v + 0.5);
    }
    free(f);
    return 0;
}
