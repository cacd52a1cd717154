// kernels_pyramid.hip -- gfx950 kernels for the Gaussian pyramid:
//   k_blur_h / k_blur_v : one separable pass of imageproc gaussian_blur_f32 on u8
//                         (called at src/sift.rs:159,1211), optional DoG epilogue
//                         (src/sift.rs:92-124,182-204) in the vertical pass.
//   k_resize_v/_h       : image resize_exact(.., Lanczos3) (src/sift.rs:83-88).
//
// Exactness of the blur. The reference computes, per output pixel,
//   s = 0; for i in taps: s = s + p_i * w_i      (f32, no FMA, in tap order)
// then truncates s to u8. Every p_i is an integer in [0,255] and every w_i > 0,
// so all terms are non-negative. The kernel first computes t with one fused
// multiply-add per tap (single rounding each step). Standard recursive-summation
// bounds give |s - S| <= g(n+1) S and |t - S| <= g(n) S with S the exact sum and
// g(k) = k u / (1 - k u), u = 2^-24, so |s - t| <= 2 g(n+1) S <= eps_rel * t
// where eps_rel = 2.05 (n+2) u is a host-computed constant (with margin). If the
// u8 truncations of t - E and t + E agree (E = eps_rel t + 6.2e-5; the absolute
// term covers the rounding of t +- E), trunc(s) is that value; otherwise the
// kernel recomputes s exactly in the reference order (mul, then add) for that
// pixel. The result equals the reference's u8 for every pixel.
#include "sift_internal.h"

namespace sh {

namespace {

constexpr int kThreads = 256;
constexpr int kHTileW = 512;  // 64 lanes x 8 px
constexpr int kHTileH = 4;    // one row per wave
constexpr int kVTileW = 256;  // 64 lanes x 4 px
constexpr int kVTileH = 32;   // 4 waves x 8 rows
constexpr float kAbsMargin = 6.2e-5f;

template <typename Job>
__device__ __forceinline__ int find_job(const Job* __restrict__ jobs, int njobs, int tile) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// imageproc Clamp<f32> for u8 == Rust `as u8`: truncate, saturate at 0 and 255.
__device__ __forceinline__ unsigned u8_trunc(float x) {
    return x < 255.0f ? (x > 0.0f ? (unsigned)x : 0u) : 255u;
}

__device__ __forceinline__ float mul_add_exact(float acc, float a, float b) {
    return __fadd_rn(acc, __fmul_rn(a, b));
}

// Decides trunc(s) from the FMA sum t, or reports that it is ambiguous.
__device__ __forceinline__ bool fast_u8(float t, float eps_rel, unsigned& out) {
    float e = t * eps_rel + kAbsMargin;
    unsigned lo = u8_trunc(t - e), hi = u8_trunc(t + e);
    out = lo;
    return lo == hi;
}

__device__ __forceinline__ float q255(unsigned v) { return __fdiv_rn((float)v, 255.0f); }

// ---------------------------------------------------------------------------
// Horizontal pass. Tile = 512 columns x 4 rows; wave w computes row y0+w; lane l
// computes the 8 outputs x0+8l .. x0+8l+7 from a register window of 16 inputs
// that slides over the taps in blocks of 8 (64 FMAs per 2 LDS reads). Taps are
// wave-uniform (scalar registers). The row segment [x0-r, x0+512+pad) is staged
// in LDS as f32 with clamp-to-edge indexing, which is the reference's edge rule.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_blur_h(const BlurJob* __restrict__ jobs, int njobs) {
    extern __shared__ float lds[];
    const int tile = blockIdx.x;
    const int j = find_job(jobs, njobs, tile);
    const BlurJob& J = jobs[j];
    const int t = tile - J.tile_begin;
    const int tx = t % J.tiles_x, ty = t / J.tiles_x;
    const int x0 = tx * kHTileW, y0 = ty * kHTileH;
    const int w = J.w, h = J.h;
    const int r = J.ntaps >> 1;
    const int npad = J.ntaps_pad;
    const int rowlen = kHTileW + npad + 8;

    for (int rr = 0; rr < kHTileH; rr++) {
        const int y = min(y0 + rr, h - 1);
        const uint8_t* src = J.src + (size_t)y * J.src_pitch;
        float* dstrow = lds + rr * rowlen;
        for (int c = threadIdx.x; c < rowlen; c += kThreads) {
            int x = min(max(x0 - r + c, 0), w - 1);
            dstrow[c] = (float)src[x];
        }
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = y0 + wave;
    if (y >= h) return;
    const float* row = lds + wave * rowlen + lane * 8;
    const float* __restrict__ taps = J.taps;

    float acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0f;
    float4 lo0 = *(const float4*)(row + 0);
    float4 lo1 = *(const float4*)(row + 4);
    for (int b = 0; b < npad; b += 8) {
        float4 hi0 = *(const float4*)(row + b + 8);
        float4 hi1 = *(const float4*)(row + b + 12);
        float win[16] = {lo0.x, lo0.y, lo0.z, lo0.w, lo1.x, lo1.y, lo1.z, lo1.w,
                         hi0.x, hi0.y, hi0.z, hi0.w, hi1.x, hi1.y, hi1.z, hi1.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wk = taps[b + k];
#pragma unroll
            for (int q = 0; q < 8; q++) acc[q] = __builtin_fmaf(win[q + k], wk, acc[q]);
        }
        lo0 = hi0;
        lo1 = hi1;
    }

    const float eps = J.eps_rel;
    unsigned res[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (!fast_u8(acc[q], eps, res[q])) {
            float s = 0.0f;
            for (int i = 0; i < J.ntaps; i++) s = mul_add_exact(s, row[q + i], taps[i]);
            res[q] = u8_trunc(s);
        }
    }

    const int x = x0 + lane * 8;
    if (x >= w) return;
    uint8_t* dst = J.dst + (size_t)y * J.dst_pitch + x;
    if (x + 8 <= w) {
        uint2 v;
        v.x = res[0] | (res[1] << 8) | (res[2] << 16) | (res[3] << 24);
        v.y = res[4] | (res[5] << 8) | (res[6] << 16) | (res[7] << 24);
        *(uint2*)dst = v;
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (x + q < w) dst[q] = (uint8_t)res[q];
    }
}

// ---------------------------------------------------------------------------
// Vertical pass. Tile = 256 columns x 32 rows; lane l owns columns x0+4l..+3
// (one u32 load per input row), wave w owns rows y0+8w..+7. A register window
// of 16 input rows x 4 columns slides over the taps in blocks of 8 (256 FMAs per
// 8 loads). Optional DoG epilogue: dog = q[G_s] - q[G_{s-1}], q[v] = v / 255.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cvt4(unsigned v, float* o) {
    o[0] = (float)(v & 0xff);
    o[1] = (float)((v >> 8) & 0xff);
    o[2] = (float)((v >> 16) & 0xff);
    o[3] = (float)(v >> 24);
}

__global__ __launch_bounds__(kThreads) void k_blur_v(const BlurJob* __restrict__ jobs, int njobs) {
    const int tile = blockIdx.x;
    const int j = find_job(jobs, njobs, tile);
    const BlurJob& J = jobs[j];
    const int t = tile - J.tile_begin;
    const int tx = t % J.tiles_x, ty = t / J.tiles_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w = J.w, h = J.h;
    const int x = tx * kVTileW + lane * 4;
    const int yb = ty * kVTileH + wave * 8;
    if (yb >= h || x >= w) return;
    const int r = J.ntaps >> 1;
    const int npad = J.ntaps_pad;
    const int pitch = J.src_pitch;
    const uint8_t* col = J.src + x;
    const float* __restrict__ taps = J.taps;

    float acc[8][4];
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[q][c] = 0.0f;

    float lo[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int yy = min(max(yb - r + i, 0), h - 1);
        cvt4(*(const unsigned*)(col + (size_t)yy * pitch), lo[i]);
    }
    for (int b = 0; b < npad; b += 8) {
        float hi[8][4];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            int yy = min(max(yb - r + b + 8 + i, 0), h - 1);
            cvt4(*(const unsigned*)(col + (size_t)yy * pitch), hi[i]);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float wk = taps[b + k];
#pragma unroll
            for (int q = 0; q < 8; q++) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float p = (q + k < 8) ? lo[q + k][c] : hi[q + k - 8][c];
                    acc[q][c] = __builtin_fmaf(p, wk, acc[q][c]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int c = 0; c < 4; c++) lo[i][c] = hi[i][c];
    }

    const float eps = J.eps_rel;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = yb + q;
        if (y >= h) break;
        unsigned res[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (!fast_u8(acc[q][c], eps, res[c])) {
                float s = 0.0f;
                for (int i = 0; i < J.ntaps; i++) {
                    int yy = min(max(y - r + i, 0), h - 1);
                    s = mul_add_exact(s, (float)col[(size_t)yy * pitch + c], taps[i]);
                }
                res[c] = u8_trunc(s);
            }
        }
        uint8_t* dst = J.dst + (size_t)y * J.dst_pitch + x;
        if (x + 4 <= w) {
            *(unsigned*)dst = res[0] | (res[1] << 8) | (res[2] << 16) | (res[3] << 24);
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (x + c < w) dst[c] = (uint8_t)res[c];
        }
        if (J.dog) {
            const uint8_t* prev = J.prev + (size_t)y * J.dst_pitch + x;
            float* dg = J.dog + (size_t)y * J.dog_pitch + x;
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (x + c < w) dg[c] = q255(res[c]) - q255(prev[c]);
        }
    }
}

// Exact thread-per-pixel passes for kernels too wide for the LDS row tile.
__global__ __launch_bounds__(kThreads) void k_blur_h_generic(const BlurJob* __restrict__ jobs,
                                                             int njobs) {
    const int tile = blockIdx.x;
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x = (t % J.tiles_x) * kThreads + threadIdx.x, y = t / J.tiles_x;
    if (x >= J.w || y >= J.h) return;
    const int r = J.ntaps >> 1;
    const uint8_t* src = J.src + (size_t)y * J.src_pitch;
    float s = 0.0f;
    for (int i = 0; i < J.ntaps; i++) {
        int xx = min(max(x + i - r, 0), J.w - 1);
        s = mul_add_exact(s, (float)src[xx], J.taps[i]);
    }
    J.dst[(size_t)y * J.dst_pitch + x] = (uint8_t)u8_trunc(s);
}

__global__ __launch_bounds__(kThreads) void k_blur_v_generic(const BlurJob* __restrict__ jobs,
                                                             int njobs) {
    const int tile = blockIdx.x;
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x = (t % J.tiles_x) * kThreads + threadIdx.x, y = t / J.tiles_x;
    if (x >= J.w || y >= J.h) return;
    const int r = J.ntaps >> 1;
    float s = 0.0f;
    for (int i = 0; i < J.ntaps; i++) {
        int yy = min(max(y + i - r, 0), J.h - 1);
        s = mul_add_exact(s, (float)J.src[(size_t)yy * J.src_pitch + x], J.taps[i]);
    }
    unsigned v = u8_trunc(s);
    J.dst[(size_t)y * J.dst_pitch + x] = (uint8_t)v;
    if (J.dog)
        J.dog[(size_t)y * J.dog_pitch + x] = q255(v) - q255(J.prev[(size_t)y * J.dst_pitch + x]);
}

// Level copy (blur sigma < 1e-4, src/sift.rs:149-150) with its (zero) DoG.
__global__ __launch_bounds__(kThreads) void k_copy_level(const BlurJob* __restrict__ jobs,
                                                         int njobs) {
    const int tile = blockIdx.x;
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x = (t % J.tiles_x) * kThreads + threadIdx.x, y = t / J.tiles_x;
    if (x >= J.w || y >= J.h) return;
    uint8_t v = J.src[(size_t)y * J.src_pitch + x];
    J.dst[(size_t)y * J.dst_pitch + x] = v;
    if (J.dog) J.dog[(size_t)y * J.dog_pitch + x] = q255(v) - q255(v);
}

// ---------------------------------------------------------------------------
// Lanczos3 resize (image 0.25.6 sample::resize): vertical pass into f32 rows
// (t += p * w in window order, unrounded), then horizontal pass with
// round(clamp(t, 0, 255)). Exact reference order (no FMA).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_resize_v(const ResizeJob* __restrict__ jobs,
                                                       int njobs) {
    const int tile = blockIdx.x;
    const ResizeJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int oy = t / J.tiles_x;
    const int x = (t % J.tiles_x) * kThreads + threadIdx.x;
    if (x >= J.in_w) return;
    const int l = J.left[oy], n = J.count[oy];
    const float* wr = J.weights + (size_t)oy * J.maxtaps;
    const uint8_t* src = J.src_u8 + (size_t)l * J.src_pitch + x;
    float s = 0.0f;
    for (int i = 0; i < n; i++) s = mul_add_exact(s, (float)src[(size_t)i * J.src_pitch], wr[i]);
    J.tmp[(size_t)oy * J.tmp_pitch + x] = s;
}

__global__ __launch_bounds__(kThreads) void k_resize_h(const ResizeJob* __restrict__ jobs,
                                                       int njobs) {
    const int tile = blockIdx.x;
    const ResizeJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int y = t / J.tiles_x;
    const int ox = (t % J.tiles_x) * kThreads + threadIdx.x;
    if (ox >= J.out_w) return;
    const int l = J.left[ox], n = J.count[ox];
    const float* wr = J.weights + (size_t)ox * J.maxtaps;
    const float* src = J.tmp + (size_t)y * J.tmp_pitch + l;
    float s = 0.0f;
    for (int i = 0; i < n; i++) s = mul_add_exact(s, src[i], wr[i]);
    float c = s < 0.0f ? 0.0f : (s > 255.0f ? 255.0f : s);
    J.dst[(size_t)y * J.dst_pitch + ox] = (uint8_t)roundf(c);
}

}  // namespace

void launch_blur_h(const BlurJob* d_jobs, int njobs, int total_tiles, int max_pad, hipStream_t s) {
    if (total_tiles <= 0) return;
    size_t lds = (size_t)kHTileH * (kHTileW + max_pad + 8) * sizeof(float);
    hipLaunchKernelGGL(k_blur_h, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs);
}

void launch_blur_v(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_blur_v, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_blur_h_generic(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_blur_h_generic, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_blur_v_generic(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_blur_v_generic, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_copy_level(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_copy_level, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_resize_v(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_resize_v, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_resize_h(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_resize_h, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

}  // namespace sh
