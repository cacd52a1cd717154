/*
 * sift_oracle.h -- CPU restatement of the reference SIFT path, used ONLY as
 * test infrastructure (the reference the parity tests compare against, and
 * the CPU baseline timing).
 *
 * The product (libsift_hip.so) never links, loads or calls anything here.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Reference followed: src/sift.rs of stevel705/sift-wgpu (cited per function in
 * sift_oracle.c), plus the two third-party routines that path calls:
 *   imageproc 0.25.0  filter::gaussian_blur_f32   (src/sift.rs:159,1211)
 *   image     0.25.6  DynamicImage::resize_exact(.., Lanczos3) (src/sift.rs:83-88)
 * Those crate sources are not available here; their restatement follows the
 * crates' published algorithm (see the header of sift_oracle.c for what pins it).
 */
#ifndef SIFT_ORACLE_H
#define SIFT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 28-byte layout as sift_keypoint_t in include/sift_hip.h and the
 * reference KeyPoint (src/keypoints.rs:1-10). */
typedef struct {
    float x, y, size, angle, response;
    int32_t octave, layer;
} oracle_keypoint_t;

/* Fields of the reference `Sift` (src/sift.rs:36-45). contrast_threshold is the
 * STORED value, i.e. already divided by num_intervals (src/sift.rs:76). */
typedef struct {
    float sigma;
    uint32_t num_octaves;
    uint32_t num_intervals;
    float assumed_blur;
    float contrast_threshold;
    float edge_threshold;
    uint32_t image_border_width;
} oracle_params_t;

/* Sift::new (src/sift.rs:63-80): stores contrast_threshold / num_intervals. */
void oracle_params_new(oracle_params_t* p, float sigma, uint32_t num_octaves,
                       uint32_t num_intervals, float assumed_blur,
                       float contrast_threshold, float edge_threshold);

/* imageproc gaussian_kernel_f32: writes 2r+1 taps (r = ceil(2 sigma)) if cap
 * allows; returns the tap count. */
int oracle_gaussian_kernel(float sigma, float* taps, int cap);

/* imageproc gaussian_blur_f32 on a u8 gray image (H pass, u8, V pass, u8). */
void oracle_gaussian_blur(const uint8_t* src, int w, int h, float sigma, uint8_t* dst);

/* image resize_exact(.., Lanczos3) for Luma<u8>. */
void oracle_resize_lanczos3(const uint8_t* src, int w, int h, uint8_t* dst, int nw, int nh);

/* Lanczos3 weight table for one axis in->out: left[out], count[out],
 * weights[out*maxtaps]. Returns the max taps needed (call with NULL first). */
int oracle_lanczos3_table(int in, int out, int* left, int* count, float* weights, int maxtaps);

/* Level sigmas (src/sift.rs:132-147): blur[o*(S+3)+s] for s>=1 (s=0 unused, 0)
 * and the initial blur sigma (src/sift.rs:1205-1209). */
void oracle_level_sigmas(const oracle_params_t* p, float* blur, float* initial_sigma);

/* Octave shapes (W_o, H_o) following the floor(w/2) rule (src/sift.rs:169-173). */
void oracle_octave_shapes(const oracle_params_t* p, int w, int h, int* ws, int* hs);

/* Pyramid bytes: sum over octaves of (S+3)*W_o*H_o; DoG floats: sum (S+2)*W_o*H_o. */
size_t oracle_pyramid_size(const oracle_params_t* p, int w, int h);
size_t oracle_dog_size(const oracle_params_t* p, int w, int h);

/* Initial blur + generate_gaussian_pyramid + generate_dog_pyramid
 * (src/sift.rs:1205-1220, 129-204). Levels are packed octave-major, level-major,
 * row-major with no padding. dog may be NULL. */
void oracle_build_pyramid(const oracle_params_t* p, const uint8_t* img, int w, int h,
                          uint8_t* gauss, float* dog);

/* Full detect_and_compute_cpu (src/sift.rs:1200-1240). Stage outputs are
 * returned through malloc'ed arrays freed with oracle_free. Any of the
 * candidate/refined outputs may be NULL. nthreads: 0 = all, 1 = serial.
 * stage_ms[6] (may be NULL): pyramid(incl. initial blur), dog, extrema, refine,
 * orientation, descriptor wall times. Returns 0. */
int oracle_detect_and_compute(const oracle_params_t* p, const uint8_t* img, int w, int h,
                              int nthreads,
                              oracle_keypoint_t** cand, size_t* n_cand,
                              oracle_keypoint_t** refined, size_t* n_refined,
                              oracle_keypoint_t** kps, float** desc, size_t* n_kps,
                              double* stage_ms);

void oracle_free(void* ptr);

/* Grey image of a colour input (w*h bytes out, no padding). format: 0 grey,
 * 1 RGB8, 2 RGBA8, 3 LumaA8 -- DynamicImage::to_luma8 of image 0.25.6, called
 * at src/sift.rs:1202 -- and 4, the wasm entry's RGBA canvas conversion
 * (src/wasm.rs:79-94). stride = bytes between input rows. */
void oracle_to_luma8(const uint8_t* src, int w, int h, size_t stride, int format, uint8_t* dst);

/* draw_keypoints_to_image (src/lib.rs:70-116): rgb (w*h*3) = to_rgb8 of the
 * input (format as for oracle_to_luma8), then per keypoint imageproc 0.25.0's
 * draw_filled_circle_mut (radius 6), draw_line_segment_mut (length 6 along
 * the angle) and draw_filled_circle_mut (radius 1) in colour c. */
void oracle_draw_keypoints(const uint8_t* src, int w, int h, size_t stride, int format,
                           const oracle_keypoint_t* kps, size_t n, const uint8_t c[3],
                           uint8_t* rgb);

/* u8 descriptors of the WebGPU backends (src/shaders/descriptor.wgsl:256-262):
 * u32(clamp(d * 512, 0, 255)) per value, n values. */
void oracle_descriptor_u8(const float* desc, size_t n, uint8_t* out);

/* Unit-level entry points (for the reference's own known-answer tests). */
int oracle_solve_linear_system(const float a[9], const float b[3], float x[3]);
float oracle_gauss_pixel_bilinear(const uint8_t* img, int w, int h, float x, float y);
void oracle_normalize_and_clip_descriptor(float* desc, int n);

/* assign_orientations / compute for one keypoint on a one-octave pyramid whose
 * levels all equal `g` (kp->octave is treated as 0). Returns the peak count. */
int oracle_orientation_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                              const oracle_keypoint_t* kp, float* angles);
void oracle_descriptor_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                              const oracle_keypoint_t* kp, float* desc);

/* The pieces of those two, for closed-form checks: the raw orientation
 * histogram (before smoothing), the two smoothing passes, the peak rule, and
 * the descriptor histogram before normalize_and_clip_descriptor. */
int oracle_orientation_histogram_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                                        const oracle_keypoint_t* kp, float* hist);
void oracle_smooth_orientation_histogram(float* hist);
int oracle_orientation_peaks(const float* hist, float* angles);
void oracle_descriptor_raw_single(const oracle_params_t* p, const uint8_t* g, int w, int h,
                                  const oracle_keypoint_t* kp, float* hist);

/* refine_and_filter_extrema (src/sift.rs:299-474) of one candidate on one
 * octave's DoG layers `dog` (nd layers of w x h f32, layer-major). Returns 1 and
 * writes *out when the candidate passes, 0 when it is dropped. */
int oracle_refine_single(const oracle_params_t* p, const float* dog, int w, int h, int nd,
                         const oracle_keypoint_t* cand, oracle_keypoint_t* out);

#ifdef __cplusplus
}
#endif
#endif
