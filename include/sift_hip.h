/*
 * sift_hip.h -- C ABI of libsift_hip.so, the MI355X (gfx950) SIFT backend.
 *
 * This is the drop-in boundary for the reference's SIFT path. A host binding
 * (the Rust `SiftBackend::Hip` arm, a ctypes wrapper, C++) calls these entry
 * points; no C++ exception or Rust panic crosses them. Return codes: 0 = OK,
 * negative = error, with a thread-local message in sift_hip_last_error().
 *
 * Reference interfaces replaced (stevel705/sift-wgpu):
 *   sift_keypoint_t      <- KeyPoint                      src/keypoints.rs:1-10
 *   sift_params_t        <- Sift {..}, Sift::new          src/sift.rs:36-80
 *   sift_hip_detect      <- Sift::detect_and_compute_cpu  src/sift.rs:1200-1240
 *                           via detect_and_compute_with_backend src/sift.rs:1244-1355
 *   sift_hip_pyramid     <- generate_gaussian_pyramid + generate_dog_pyramid
 *                           (with the initial blur)        src/sift.rs:129-204,1205-1214
 *   sift_hip_detect_stages <- find_scale_space_extrema / refine_and_filter_extrema
 *                           stage outputs                  src/sift.rs:510-603,299-474
 *   sift_hip_to_luma8,    <- DynamicImage::to_luma8 before the CPU path src/sift.rs:1202,
 *   sift_hip_detect*_color   img_from_raw of the wasm entry          src/wasm.rs:79-103
 *   sift_draw_keypoints  <- draw_keypoints_to_image                  src/lib.rs:70-116
 *   sift_hip_detect_batch_u8 <- the WebGPU arms' u8 descriptors and V2 keypoints
 *                           src/shaders/descriptor.wgsl:256-262, src/sift.rs:1288-1291,
 *                           src/gpu_sift_v2.rs:1699-1707
 */
#ifndef SIFT_HIP_H
#define SIFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define SIFT_OK 0
#define SIFT_ERR_INVALID -1   /* bad argument (null pointer, zero size, bad params) */
#define SIFT_ERR_DEVICE -2    /* HIP runtime error (no device, launch failure, ...) */
#define SIFT_ERR_NOMEM -3     /* host or device allocation failed */
#define SIFT_ERR_BUSY -4      /* serving queue full / admission refused */
#define SIFT_ERR_TIMEOUT -5   /* sift_hip_poll timed out with nothing completed */

/* KeyPoint (src/keypoints.rs:1-10), #[repr(C)] mirror: 5 x f32 + 2 x i32 = 28 B.
 * x, y in input-image pixels; size = sigma; angle in radians, (-pi, pi]. */
typedef struct sift_keypoint_t {
    float x, y, size, angle, response;
    int32_t octave, layer;
} sift_keypoint_t;

/* Fields of `Sift` (src/sift.rs:36-45). contrast_threshold is the STORED value,
 * i.e. already divided by num_intervals as Sift::new does (src/sift.rs:76);
 * pass `sift.contrast_threshold` through unchanged. image_border_width is 5 in
 * the reference (src/sift.rs:17,78). */
typedef struct sift_params_t {
    float sigma;
    uint32_t num_octaves;
    uint32_t num_intervals;
    float assumed_blur;
    float contrast_threshold;
    float edge_threshold;
    uint32_t image_border_width;
} sift_params_t;

/* 8-bit grey image, row-major. stride = bytes between rows (>= width; 0 means
 * width). For the *_device entry points `data` is a device pointer. */
typedef struct sift_image_t {
    const uint8_t* data;
    uint32_t width, height, stride;
} sift_image_t;

/* Host result, allocated by the library, released with sift_result_free.
 * descriptors: count x 128 f32, row-major, same convention as the CPU path
 * (normalised, clipped at 0.2, renormalised; src/sift.rs:889-913). Order: as the
 * CPU path emits them (octave, layer, row, column of the extremum, then
 * orientation peaks in ascending bin order). */
typedef struct sift_result_t {
    sift_keypoint_t* keypoints;
    float* descriptors;
    size_t count;
} sift_result_t;

typedef struct sift_hip_ctx sift_hip_ctx;

/* Sift::default() values (src/sift.rs:47-60): sigma 1.6, 4 octaves, 3 intervals,
 * assumed_blur 0.5, contrast 0.04/3 (stored), edge 10, border 5. */
void sift_params_default(sift_params_t* out);

/* Sift::new(sigma, octaves, intervals, assumed_blur, contrast, edge)
 * (src/sift.rs:63-80): stores contrast / intervals. */
void sift_params_new(sift_params_t* out, float sigma, uint32_t num_octaves,
                     uint32_t num_intervals, float assumed_blur, float contrast_threshold,
                     float edge_threshold);

/* Creates a context on HIP device `device` (ordinal). Fails with
 * SIFT_ERR_DEVICE when no device is visible. One stream per context; calls on
 * one context are serialised internally. */
int sift_hip_create(const sift_params_t* params, int device, sift_hip_ctx** out);
void sift_hip_destroy(sift_hip_ctx* ctx);

/* Thread-local message for the last failing call on this thread ("" if none). */
const char* sift_hip_last_error(void);

/* One image, host in / host out (the `SiftBackend::Hip` arm's call). */
int sift_hip_detect(sift_hip_ctx* ctx, const sift_image_t* image, sift_result_t* out);

/* n images of any sizes in one pass (all stages batched across images).
 * params: NULL = the context's params for every image, else n entries.
 * out: n results. */
int sift_hip_detect_batch(sift_hip_ctx* ctx, const sift_image_t* images,
                          const sift_params_t* params, size_t n, sift_result_t* out);

/* Device-resident variant: `images[i].data` are device pointers; results stay
 * in device memory owned by the context and valid until the next call on it.
 * out_counts (host, n entries, may be NULL) receive per-image keypoint counts
 * after the call returns (it synchronises once). d_keypoints/d_descriptors
 * (may be NULL) receive device pointers to the concatenated results in image
 * order; d_offsets (may be NULL) a device pointer to n+1 uint32 offsets. */
int sift_hip_detect_batch_device(sift_hip_ctx* ctx, const sift_image_t* images,
                                 const sift_params_t* params, size_t n, size_t* out_counts,
                                 const sift_keypoint_t** d_keypoints,
                                 const float** d_descriptors, const uint32_t** d_offsets);

/* No counterpart in the reference: sift_hip_detect_batch_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It calls no device or pinned
 * allocator: its working set (pyramids, temporaries, candidate lists, weight
 * tables, job records) lives in one static device array of
 * SIFT_HIP_SCRATCH_BYTES per GPU, held from the library's first use on that GPU
 * until the process ends, and it writes the results to the caller's buffers:
 * d_keypoints / d_descriptors (device, keypoint_capacity entries, 128 f32 each),
 * d_offsets (device, n + 1 uint32 per-image keypoint offsets) and out_counts
 * (host, n entries, may be NULL). It synchronises once, like the variant above.
 * Returns SIFT_ERR_NOMEM, with a message, when the batch's working set exceeds
 * SIFT_HIP_SCRATCH_BYTES, and when the keypoints exceed keypoint_capacity (then
 * out_counts and d_offsets hold the counts and offsets found, and nothing past
 * keypoint_capacity is written). All calls on one GPU share its array, from any
 * context or thread: calls that overlap on one GPU race and must be serialised
 * by the caller. */
#define SIFT_HIP_SCRATCH_BYTES ((size_t)256 << 20)
int sift_hip_detect_batch_device_noalloc(sift_hip_ctx* ctx, const sift_image_t* images,
                                         const sift_params_t* params, size_t n,
                                         sift_keypoint_t* d_keypoints, float* d_descriptors,
                                         size_t keypoint_capacity, uint32_t* d_offsets,
                                         size_t* out_counts);

void sift_result_free(sift_result_t* r);

/* Pyramid introspection: initial blur + Gaussian pyramid + DoG pyramid
 * (src/sift.rs:1205-1214,129-204) for one image, copied to host.
 * gauss: sum_o (S+3)*W_o*H_o bytes; dog: sum_o (S+2)*W_o*H_o floats (may be NULL);
 * octave-major, level-major, row-major, no padding. W_{o+1} = floor(W_o/2). */
int sift_hip_pyramid(sift_hip_ctx* ctx, const sift_image_t* image, uint8_t* gauss, float* dog);

/* Stage outputs for parity checks: extremum candidates (pre-refinement,
 * src/sift.rs:510-603) and refined keypoints (src/sift.rs:299-474, before
 * orientation), plus the final result. Arrays are library-allocated and freed
 * by sift_stages_free. */
typedef struct sift_stages_t {
    sift_keypoint_t* candidates;
    size_t n_candidates;
    sift_keypoint_t* refined;
    size_t n_refined;
    sift_result_t result;
} sift_stages_t;

int sift_hip_detect_stages(sift_hip_ctx* ctx, const sift_image_t* image, sift_stages_t* out);
void sift_stages_free(sift_stages_t* s);

/* ---- Kernel timing --------------------------------------------------------
 * When enabled, every kernel launch of the context is bracketed by HIP events
 * on the context's stream; durations, the launch's algorithmic HBM bytes (u8
 * levels in and out for the blur and resize passes, levels read plus bitmask
 * written for the extremum scan; 0 for the gather-bound keypoint stages) and
 * its prescribed f32 operations (a multiply and an add per Gaussian tap or
 * Lanczos weight and output pixel, the reference's arithmetic, which the
 * kernels keep without FMA contraction; 0 for other kernels) are accumulated
 * per kernel name. Enabling resets the counters. */
typedef struct sift_kernel_stat_t {
    char name[32];
    uint64_t launches;
    double total_ms;
    double bytes;
    double ops;
} sift_kernel_stat_t;

int sift_hip_profile(sift_hip_ctx* ctx, int enable);

/* ---- Blur tuning ----------------------------------------------------------
 * No counterpart in the reference: the measurement behind the tuning table that
 * chooses each pyramid level's blur form (its tile shape and outputs per lane)
 * by level width, height and tap count. Times one separable blur of sigma on
 * `jobs` levels of width x height in one launch per pass, with the given
 * horizontal form ("h8", "h2", "hsmall", "hgeneric", "h16") and vertical form ("v",
 * "vnarrow", "vsmall", "vgeneric", "vsmall8", "v8"; NULL or "" = the built-in choice), and
 * writes the median over `reps` of the horizontal and the vertical pass, in ms,
 * to out_ms[0] and out_ms[1], and the forms used (indices in the two lists
 * above) to out_forms[0] and [1] (may be NULL). SIFT_ERR_INVALID for a form not
 * valid for the level. The table the library loads at context creation is blur_tune.json
 * next to it (or the file SIFT_HIP_TUNE_TABLE names; "0" = none); its entries
 * are [width, height, taps, many, "<h form>", "<v form>"], many = 1 for launches
 * carrying a batch of 8 or more frames. */
int sift_hip_tune_blur(sift_hip_ctx* ctx, uint32_t width, uint32_t height, float sigma,
                       uint32_t jobs, const char* h_form, const char* v_form, uint32_t reps,
                       double* out_ms, int* out_forms);
/* Restricts the timing to launches of one kernel (its name as in
 * sift_kernel_stat_t, e.g. "k_blur_h<8>"), so that the events add little to
 * the timed work; NULL or "" times every kernel. Unknown name: SIFT_ERR_INVALID. */
int sift_hip_profile_filter(sift_hip_ctx* ctx, const char* kernel);
int sift_hip_profile_read(sift_hip_ctx* ctx, sift_kernel_stat_t* out, size_t cap, size_t* n);

/* ---- Serving: continuous batching on one device -------------------------
 * Frames are admitted while device memory for their pyramids fits within
 * `hbm_budget_bytes`; frames arriving while others run are batched into the
 * next launch, so their work overlaps on the device. Completed frames are
 * returned by sift_hip_poll with the submit tag. */
typedef struct sift_server_config_t {
    size_t hbm_budget_bytes;  /* device bytes the in-flight pyramid pool may use */
    uint32_t max_batch;       /* max frames per device launch (0 = 64) */
    uint32_t max_queue;       /* max queued (not yet admitted) frames (0 = 4096) */
} sift_server_config_t;

typedef struct sift_completion_t {
    uint64_t tag;
    int status;               /* SIFT_OK or an error code */
    sift_result_t result;     /* owned by the caller after poll; free with sift_result_free */
    double submit_ms, start_ms, done_ms;  /* steady-clock ms, for latency accounting */
} sift_completion_t;

/* ---- Colour ingest and the WebGPU backends' output conventions ----------
 * The CPU path starts with DynamicImage::to_luma8 (src/sift.rs:1202); these
 * entry points take the colour pixels and convert them on the device. */
#define SIFT_PIXEL_GRAY8 0        /* one byte per pixel, used as is */
#define SIFT_PIXEL_RGB8 1         /* R,G,B: to_luma8 of image 0.25.6, (2126 R + 7152 G + 722 B) / 10000 */
#define SIFT_PIXEL_RGBA8 2        /* R,G,B,A: the same, alpha ignored */
#define SIFT_PIXEL_LUMAA8 3       /* L,A: L */
#define SIFT_PIXEL_RGBA8_CANVAS 4 /* R,G,B,A: the wasm entry's conversion (src/wasm.rs:79-94),
                                     (0.299 r + 0.587 g + 0.114 b) in f32, truncated to u8 */

/* Row-major pixels of `format`; stride = bytes between rows (0 means
 * width * bytes per pixel). */
typedef struct sift_color_image_t {
    const uint8_t* data;
    uint32_t width, height, stride;
    uint32_t format;
} sift_color_image_t;

/* Grey image of one colour image (width*height bytes, no padding), converted
 * on the device: to_luma8 / img_from_raw (src/sift.rs:1202, src/wasm.rs:79-94). */
int sift_hip_to_luma8(sift_hip_ctx* ctx, const sift_color_image_t* image, uint8_t* out);

/* sift_hip_detect / _batch on colour images: the result equals that of the grey
 * image sift_hip_to_luma8 gives. */
int sift_hip_detect_color(sift_hip_ctx* ctx, const sift_color_image_t* image, sift_result_t* out);
int sift_hip_detect_batch_color(sift_hip_ctx* ctx, const sift_color_image_t* images,
                                const sift_params_t* params, size_t n, sift_result_t* out);

/* Descriptors in the u8 form of the WebGPU backends (src/shaders/descriptor.wgsl:256-262):
 * u8 = trunc(clamp(d * 512, 0, 255)) of the CPU-convention descriptor d; the
 * reference hands them to its callers as u8 / 255 (src/sift.rs:1288-1291).
 * 128 B per keypoint instead of 512 B, also on the device-to-host copy. */
#define SIFT_KEYPOINTS_CPU 0      /* keypoints as the CPU path gives them */
#define SIFT_KEYPOINTS_WGPU_V2 1  /* as the V2 backend reports them (src/gpu_sift_v2.rs:1699-1707):
                                     size = 2 * sigma; response, octave and layer zero */
typedef struct sift_result_u8_t {
    sift_keypoint_t* keypoints;
    uint8_t* descriptors;  /* count x 128 */
    size_t count;
} sift_result_u8_t;

int sift_hip_detect_batch_u8(sift_hip_ctx* ctx, const sift_color_image_t* images,
                             const sift_params_t* params, size_t n, uint32_t keypoint_convention,
                             sift_result_u8_t* out);
void sift_result_u8_free(sift_result_u8_t* r);

/* draw_keypoints_to_image (src/lib.rs:70-116), on the host (no device work):
 * the image as RGB8 (DynamicImage::to_rgb8), then per keypoint a filled circle
 * of radius 6 at (x as i32, y as i32), a line of length 6 along its angle and a
 * filled circle of radius 1, in `color`, rasterised as imageproc 0.25.0's
 * draw_filled_circle_mut / draw_line_segment_mut. out_rgb: width*height*3 bytes. */
int sift_draw_keypoints(const sift_color_image_t* image, const sift_keypoint_t* keypoints,
                        size_t count, const uint8_t color[3], uint8_t* out_rgb);

int sift_hip_server_start(sift_hip_ctx* ctx, const sift_server_config_t* cfg);
/* Copies the pixels before returning. params: NULL = the context's params,
 * else this frame's (e.g. an octave count chosen by frame size).
 * SIFT_ERR_BUSY when the queue is full, SIFT_ERR_NOMEM when the frame alone
 * exceeds the HBM budget. */
int sift_hip_submit(sift_hip_ctx* ctx, const sift_image_t* image, const sift_params_t* params,
                    uint64_t tag);
/* sift_hip_submit for a colour frame (converted to grey on the device). */
int sift_hip_submit_color(sift_hip_ctx* ctx, const sift_color_image_t* image,
                          const sift_params_t* params, uint64_t tag);
int sift_hip_poll(sift_hip_ctx* ctx, sift_completion_t* out, size_t cap, size_t* n,
                  int timeout_ms);
int sift_hip_server_stop(sift_hip_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
