// draw.cpp -- sift_draw_keypoints: draw_keypoints_to_image (src/lib.rs:70-116)
// on the host, the step after the path that turns its keypoints into a picture.
//
// The image becomes RGB8 as DynamicImage::to_rgb8 does (grey and LumaA repeat
// the luma byte, RGBA drops alpha). Per keypoint, in the reference's order and
// f32 arithmetic: a filled circle of radius max(round(2 * 3), 2) = 6 at
// (x as i32, y as i32), a line from (x, y) to (x + 6 cos(angle), y + 6 sin(angle)),
// and a filled circle of radius 1 at the centre. Circles and lines rasterise as
// imageproc 0.25.0 does: draw_filled_circle_mut walks the midpoint circle and
// fills horizontal spans; draw_line_segment_mut walks Bresenham's line over f32
// end points and skips points outside the image.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "sift_hip.h"

namespace {

struct Canvas {
    uint8_t* px;
    int w, h;
    uint8_t c[3];
    void put(int x, int y) const {
        if (x < 0 || y < 0 || x >= w || y >= h) return;
        uint8_t* p = px + ((size_t)y * w + x) * 3;
        p[0] = c[0];
        p[1] = c[1];
        p[2] = c[2];
    }
};

// `v as i32` for f32 (truncation toward zero, saturating, NaN -> 0).
int32_t as_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

void line(const Canvas& cv, float x0, float y0, float x1, float y1) {
    const bool steep = fabsf(y1 - y0) > fabsf(x1 - x0);
    if (steep) {
        float t = x0; x0 = y0; y0 = t;
        t = x1; x1 = y1; y1 = t;
    }
    if (x0 > x1) {
        float t = x0; x0 = x1; x1 = t;
        t = y0; y0 = y1; y1 = t;
    }
    const float dx = x1 - x0, dy = fabsf(y1 - y0);
    float err = dx / 2.0f;
    int x = as_i32(x0), y = as_i32(y0);
    const int end_x = as_i32(x1), ystep = y0 < y1 ? 1 : -1;
    for (; x <= end_x; x++) {
        if (steep) cv.put(y, x);
        else cv.put(x, y);
        err -= dy;
        if (err < 0.0f) {
            y += ystep;
            err += dx;
        }
    }
}

void filled_circle(const Canvas& cv, int cx, int cy, int r) {
    // each horizontal span has integer end points, which Bresenham's walk
    // covers pixel for pixel
    auto span = [&](int xa, int xb, int y) {
        for (int x = xa; x <= xb; x++) cv.put(x, y);
    };
    int x = 0, y = r, p = 1 - r;
    while (x <= y) {
        span(cx - x, cx + x, cy + y);
        span(cx - y, cx + y, cy + x);
        span(cx - x, cx + x, cy - y);
        span(cx - y, cx + y, cy - x);
        x++;
        if (p < 0) {
            p += 2 * x + 1;
        } else {
            y--;
            p += 2 * (x - y) + 1;
        }
    }
}

}  // namespace

extern "C" int sift_draw_keypoints(const sift_color_image_t* image, const sift_keypoint_t* keypoints,
                                   size_t count, const uint8_t color[3], uint8_t* out_rgb) {
    if (!image || !image->data || !color || !out_rgb || (count && !keypoints)) return SIFT_ERR_INVALID;
    const uint32_t f = image->format;
    const int bpp = f == SIFT_PIXEL_GRAY8 ? 1 : f == SIFT_PIXEL_RGB8 ? 3 : f == SIFT_PIXEL_LUMAA8 ? 2
                  : (f == SIFT_PIXEL_RGBA8 || f == SIFT_PIXEL_RGBA8_CANVAS) ? 4 : 0;
    if (bpp == 0 || image->width == 0 || image->height == 0) return SIFT_ERR_INVALID;
    const int w = (int)image->width, h = (int)image->height;
    const size_t stride = image->stride ? image->stride : (size_t)w * bpp;
    if (stride < (size_t)w * bpp) return SIFT_ERR_INVALID;
    for (int y = 0; y < h; y++) {
        const uint8_t* s = image->data + (size_t)y * stride;
        uint8_t* d = out_rgb + (size_t)y * w * 3;
        for (int x = 0; x < w; x++, s += bpp, d += 3) {
            if (bpp <= 2) {
                d[0] = d[1] = d[2] = s[0];
            } else {
                d[0] = s[0];
                d[1] = s[1];
                d[2] = s[2];
            }
        }
    }
    Canvas cv{out_rgb, w, h, {color[0], color[1], color[2]}};
    const float size = 2.0f;  // src/lib.rs:87 draws every point at size 2
    const int radius = as_i32(roundf(size * 3.0f));
    const int rd = radius > 2 ? radius : 2;
    for (size_t i = 0; i < count; i++) {
        const float x = keypoints[i].x, y = keypoints[i].y, a = keypoints[i].angle;
        filled_circle(cv, as_i32(x), as_i32(y), rd);
        const float xe = x + (float)rd * cosf(a);
        const float ye = y + (float)rd * sinf(a);
        line(cv, x, y, xe, ye);
        filled_circle(cv, as_i32(x), as_i32(y), 1);
    }
    return SIFT_OK;
}
