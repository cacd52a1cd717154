// sift_internal.h -- device-visible job descriptors and host launch wrappers
// shared by the gfx950 kernels and the host orchestration of libsift_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sh {

// Gaussian taps are zero-padded to a multiple of this (register tap block).
constexpr int kTapBlock = 8;
// Max orientation peaks per keypoint: strict local maxima of a circular 36-bin
// histogram (src/sift.rs:691-719) are at most 18.
constexpr int kMaxPeaks = 18;

// The dispatcher places work-group b on XCD b % 8 (each XCD has its own L2).
// xcd_remap(b, n) is a bijection of [0, n) that gives the work-groups of one
// XCD consecutive logical indices, so neighbouring tiles (shared halo lines)
// are served by one L2. (The per-keypoint kernels keep the plain order: it
// balances keypoint sizes across the XCDs better than locality pays.)
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n / kNumXcd, r = n % kNumXcd, x = b % kNumXcd;
    return x * q + min(x, r) + b / kNumXcd;
}

// One separable-blur pass over one image level (horizontal or vertical).
struct BlurJob {
    const uint8_t* src;
    uint8_t* dst;
    const uint8_t* prev;  // V pass DoG epilogue: level s-1 (may be null)
    float* dog;           // V pass DoG epilogue: dog = q[dst] - q[prev] (may be null)
    const float* taps;    // ntaps_pad floats, zero-padded
    int ntaps;            // 2r+1
    int ntaps_pad;        // multiple of kTapBlock
    int w, h;
    int src_pitch, dst_pitch, dog_pitch;
    int tiles_x;          // tiles along x for this job
    int tile_begin;       // prefix over jobs in the launch
    int tiles;            // tiles of this job
};

// One Lanczos3 resize pass (vertical: u8 -> f32 rows; horizontal: f32 -> u8).
struct ResizeJob {
    const uint8_t* src_u8;  // vertical input
    float* tmp;             // vertical output / horizontal input (pitch tmp_pitch floats)
    uint8_t* dst;           // horizontal output
    const int* left;        // per output index
    const int* count;
    const float* weights;   // [out][maxtaps]
    int maxtaps;
    // the vertical table, for the fused kernel (its left/count/weights are horizontal)
    const int* vleft;
    const int* vcount;
    const float* vweights;
    int vmaxtaps;
    int in_w, in_h, out_w, out_h;
    int src_pitch, tmp_pitch, dst_pitch;
    int tiles_x;
    int tile_begin;
    int tiles;
};

// Colour -> grey conversion of one input image (to_luma8 / the wasm canvas rule).
// Pixel formats are the SIFT_PIXEL_* values of the C ABI.
struct LumaJob {
    const uint8_t* src;
    uint8_t* dst;
    int w, h;
    int src_pitch, dst_pitch;
    int format;
    int tiles_x;
    int tile_begin;
    int tiles;
};
constexpr int kLumaTileW = 1024;  // 256 lanes x 4 px

// Per-image parameters read by the detection kernels.
struct ImageParams {
    float sigma;
    float contrast_threshold;  // stored value (already / S)
    float edge_threshold;
    int num_intervals;
    int border;
    int num_octaves;
};

// One (image, octave) of the pyramid.
struct OctaveDesc {
    const uint8_t* levels;  // level s at levels + s*level_stride
    long long level_stride;
    int w, h, pitch;
    int octave;
    int img;
    int nlevels;            // S+3
    int slot_begin;         // first extremum row-slot of this octave: (s-1)*h + y
    long long bits_begin;   // first u64 word of the extremum bitmask
    int nwords;             // u64 words per row-slot
    int row_begin;          // prefix of rows over octaves (extremum launch grid)
};

// Candidate extremum (before refinement): octave job index, position, layer.
struct Cand {
    int job;      // index into the OctaveDesc table
    int x, y;
    int layer;
};

struct Keypoint {
    float x, y, size, angle, response;
    int octave, layer;
};

// Launch wrappers (kernels_pyramid.hip / kernels_detect.hip).
// px: outputs per lane (8 for wide rows, 2 for narrow rows).
void launch_blur_h(const BlurJob* d_jobs, int njobs, int total_tiles, int max_pad, int px,
                   hipStream_t s);
// narrow_rows > 0 selects the 64x16-tile LDS variant with that many staged rows.
void launch_blur_v(const BlurJob* d_jobs, int njobs, int total_tiles, int narrow_rows,
                   hipStream_t s);
// 8 rows per lane (256 x 32 tiles); only for 7, 9, 11, 13 or 15 taps.
void launch_blur_v8(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
void launch_resize_v(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
// Both Lanczos3 passes in one kernel, the f32 intermediate kept in LDS: for jobs
// whose every tile of kResizeRows x 256 outputs needs at most kResizeFusedSpan
// input columns (from left[ox0] rounded down to a multiple of 4).
void launch_resize_fused(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
void launch_resize_h(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
void launch_copy_level(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
// Small levels: one output per lane (tiles of 64 x 4). max_taps4 = the largest
// tap count of the launch rounded up to 4; max_rows = the largest number of
// distinct input rows a vertical tile stages (min(h, 4 + 2r)).
void launch_blur_h_small(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                         hipStream_t s);
void launch_blur_v_small(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                         int max_rows, hipStream_t s);
// Byte-staged form of launch_blur_v_small (max_rows * 64 B of rows instead of f32).
void launch_blur_v_small8(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                          int max_rows, hipStream_t s);
void launch_blur_h_generic(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
// One tile = kLumaTileW pixels of one row.
void launch_luma(const LumaJob* d_jobs, int njobs, int total_tiles, hipStream_t s);
void launch_blur_v_generic(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s);

void launch_extrema(const OctaveDesc* d_oct, int noct, int total_rows, const ImageParams* d_params,
                    unsigned long long* d_bits, unsigned* d_counts, int nslots, hipStream_t s);
// Tiled extremum scan (64x16 tiles) for S in [1, 6]: zeroes the row-slot
// counts, then one workgroup per tile; OctaveDesc.row_begin holds the tile prefix.
bool extrema_tile_supported(int S);
void launch_extrema_tile(int S, int strip_rows, const OctaveDesc* d_oct, int noct, int total_tiles,
                         const ImageParams* d_params, unsigned long long* d_bits, unsigned* d_counts,
                         int nslots, hipStream_t s);
void launch_scan_dev(const unsigned* d_in, int n_max, const unsigned* n_dev, unsigned* d_out,
                     unsigned* d_tmp, hipStream_t s);
void launch_compact(const OctaveDesc* d_oct, int noct, int nslots, const unsigned long long* d_bits,
                    const unsigned* d_counts, const unsigned* d_offsets, Cand* d_cand, int cap,
                    hipStream_t s);
void launch_refine(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, Keypoint* d_ref, int* d_valid,
                   float* d_cand_resp, hipStream_t s);
// d_queue / d_queues: zeroed work-queue counters (kNumXcd for k_orient, kNumXcd
// for each of the descriptor's two instantiations), among the kCountExtra after
// the row slots.
void launch_orient(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, const Keypoint* d_ref, const int* d_valid,
                   unsigned* d_npeaks, float* d_angles, unsigned* d_queue, hipStream_t s);
void launch_descriptor(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                       const unsigned* d_total, int cap, const Keypoint* d_ref,
                       const unsigned* d_npeaks, const float* d_angles, const unsigned* d_koff,
                       int kcap, Keypoint* d_kps, float* d_desc, unsigned* d_wide_count,
                       int* d_wide_list, unsigned* d_mid_count, int* d_mid_list, unsigned* d_queues,
                       hipStream_t s);
// Writes summary[0] = candidates, summary[1] = keypoints, summary[2..2+nimg] =
// per-image keypoint offsets (nimg+1 entries), the offsets also to
// d_out_offsets when it is not null.
void launch_finalize(const unsigned* d_slot_offsets, const int* d_img_slot_begin, int nimg,
                     const unsigned* d_koff, const unsigned* d_total, int cap,
                     unsigned* d_summary, unsigned* d_out_offsets, hipStream_t s);
// u8 descriptors of the WebGPU backends: out = trunc(clamp(d * 512, 0, 255)) for
// the first min(*d_count, kcap) descriptors.
void launch_desc_u8(const float* d_desc, const unsigned* d_count, int kcap, uint8_t* d_out,
                    hipStream_t s);
// Max padded tap count handled by the LDS row-tile horizontal pass.
constexpr int kMaxHPad = 3576;
// Small-level kernels: taps (rounded up to 4) and staged rows they accept.
constexpr int kSmallMaxTaps4 = 2800;
constexpr int kSmallMaxRows = 128;
// Staged rows of the byte-staged small vertical form: with kSmallMaxTaps4 taps,
// 11.2 KB + 50 KB of LDS, within 64 KB whatever jobs share a launch.
constexpr int kSmallMaxRows8 = 800;
constexpr int kScanTileElems = 4096;
// Counters after the extremum row-slot counts, zeroed with them: the
// descriptor's two lists, then the work-queue counters (kNumXcd each) of
// k_orient and of the two descriptor instantiations.
constexpr int kCountExtra = 2 + 3 * kNumXcd;
// Lanczos3 resize tiles: output rows per work-group of the horizontal and fused
// kernels, and the input columns the fused kernel stages per tile.
constexpr int kResizeRows = 8;
constexpr int kResizeFusedSpan = 576;
// Rows of one extremum-test strip (64 columns wide, tested as bands of 16 rows):
// kExtStripRows when a launch has at least kExtFillBands bands of 16 rows (enough
// strips to fill the GPU several times over), else one band per strip.
constexpr int kExtStripRows = 64;
constexpr int kExtFillBands = 18432;

}  // namespace sh
