// kernels_pyramid.hip -- gfx950 kernels for the Gaussian pyramid:
//   k_blur_h / k_blur_v : one separable pass of imageproc gaussian_blur_f32 on u8
//                         (called at src/sift.rs:159,1211), optional DoG epilogue
//                         (src/sift.rs:92-124,182-204) in the vertical pass.
//   k_resize_fused      : image resize_exact(.., Lanczos3) (src/sift.rs:83-88), both
//                         passes in one kernel; k_resize_v/_h: the two-pass form.
//   k_luma              : colour -> grey (image 0.25.6 to_luma8, the wasm canvas rule).
//
// Exactness: every pass computes, per output pixel, acc = 0; acc = acc + p_i * w_i
// in tap order with separate multiply and add (no FMA; the file is compiled with
// -ffp-contract=off and uses __fmul_rn/__fadd_rn), then truncates to u8 -- the
// reference's arithmetic, so the levels are bit-identical to it.
#include "sift_hip.h"
#include "sift_internal.h"

namespace sh {

namespace {

constexpr int kThreads = 256;
constexpr int kHTileH = 4;    // one row per wave
constexpr int kVTileW = 256;  // 64 lanes x 4 px
constexpr int kVRows = 4;     // rows per lane in k_blur_v (a tile: 4 waves x kVRows rows)

// The job of a tile (jobs[j].tile_begin <= tile < jobs[j+1].tile_begin); every
// launch's grid is its total tile count. With more than 4 jobs the proportional
// guess tile * njobs / total is checked first: it is exact when the jobs have
// equal tile counts (a batch of equal frames), and then the lookup is two
// independent loads instead of a chain of dependent ones; otherwise, and for
// other layouts, a binary search.
template <typename Job>
__device__ __forceinline__ int find_job(const Job* __restrict__ jobs, int njobs, int tile) {
    if (njobs > 4) {
        const int g = (int)(((long long)tile * njobs) / (long long)gridDim.x);
        if (g < njobs && jobs[g].tile_begin <= tile && (g + 1 == njobs || jobs[g + 1].tile_begin > tile))
            return g;
    }
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// imageproc Clamp<f32> for u8 == Rust `as u8`: truncate, saturate at 0 and 255.
// Every caller converts a sum of products of non-negative values (pixel values
// and Gaussian taps, or the luma weights), which is never negative or NaN and
// stays far below 2^32: the conversion truncates and only the upper bound needs
// a test (two instructions instead of five).
__device__ __forceinline__ unsigned u8_trunc(float x) {
    return min((unsigned)x, 255u);
}

__device__ __forceinline__ float mul_add_exact(float acc, float a, float b) {
    return __fadd_rn(acc, __fmul_rn(a, b));
}

__device__ __forceinline__ float q255(unsigned v) { return __fdiv_rn((float)v, 255.0f); }

// The tap tables are written by host copies before any launch and only read
// by kernels: reading them through the constant address space lets the
// wave-uniform tap loads go through scalar loads instead of vector loads.
typedef __attribute__((address_space(4))) const float ctap_t;
// Global memory (source rows): loads through it are global, not flat, loads.
typedef __attribute__((address_space(1))) const uint8_t gbyte_t;
__device__ __forceinline__ const ctap_t* as_const_taps(const float* p) {
    return (const ctap_t*)(p);
}

// ---------------------------------------------------------------------------
// Horizontal pass. Tile = (64*PX) columns x 4 rows; wave w computes row y0+w;
// lane l computes PX consecutive outputs. The row segment from the aligned word
// at or before x0-r (d = (x0-r) mod 4 elements early) is staged in LDS as f32
// with clamp-to-edge indexing (the reference's edge rule), one word of 4 pixels
// per lane and step, with pad words so that the lanes' reads, PX elements apart,
// fall in distinct banks: PX = 16 puts one pad word after every 16 elements
// (element e at e + e/16, lane stride 17 words), PX = 8 one after every 8
// (e + e/8, lane stride 9 words), PX = 2 one after every 32 (e + e/32). Either
// way a word's 4 pixels land in 4 consecutive slots.
// For PX = 8 and 16 and the tap counts of the default scale schedule (7 to 15
// taps) the taps are fully unrolled: the PX + n - 1 window inputs, elements
// base + d + k, sit at offsets hpad(d + k) from the lane's first slot (the same
// for every lane), so they are read once with immediate offsets, one code path
// per d, and every tap is a scalar register. PX = 16 gives each wave twice the
// arithmetic per staged row, against which the row's load latency is spread. Otherwise a register window of PX+4 inputs
// slides over the taps in blocks of 4; taps are zero-padded to a multiple of 4
// (p * 0 added to a non-negative sum leaves it unchanged), except that a single
// leftover tap is applied on its own, and are wave-uniform (scalar registers).
// ---------------------------------------------------------------------------
template <int PX>
__host__ __device__ __forceinline__ int hpad(int c) {
    return PX == 16 ? c + (c >> 4) : PX == 8 ? c + (c >> 3) : c + (c >> 5);
}

// NT taps over the window elements base + D + k, k < PX + NT - 1, at
// p[hpad(D + k)] (p = the slot of element base, base a multiple of PX, which is
// the padding period for PX = 8 and 16), in tap order per output; the first
// product is the first partial sum (0 + p * w0 is p * w0: p * w0 is never -0).
template <int PX, int NT, int D>
__device__ __forceinline__ void hconv_fixed(const float* __restrict__ p, const ctap_t* taps,
                                            float (&acc)[PX]) {
    constexpr int NW = PX + NT - 1;
    float win[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) win[k] = p[hpad<PX>(D + k)];
    float wk[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) wk[k] = taps[k];
#pragma unroll
    for (int q = 0; q < PX; q++) acc[q] = __fmul_rn(win[q], wk[0]);
#pragma unroll
    for (int k = 1; k < NT; k++)
#pragma unroll
        for (int q = 0; q < PX; q++) acc[q] = mul_add_exact(acc[q], win[q + k], wk[k]);
}

template <int PX, int NT>
__device__ __forceinline__ void hconv_fixed_d(int d, const float* __restrict__ p,
                                              const ctap_t* taps, float (&acc)[PX]) {
    switch (d) {
        case 0: hconv_fixed<PX, NT, 0>(p, taps, acc); break;
        case 1: hconv_fixed<PX, NT, 1>(p, taps, acc); break;
        case 2: hconv_fixed<PX, NT, 2>(p, taps, acc); break;
        default: hconv_fixed<PX, NT, 3>(p, taps, acc); break;
    }
}

// Any tap count: taps in blocks of 4 (taps beyond ntaps are zero, <= ntaps_pad).
// When one tap is left over (ntaps = 4k + 1), it is applied after the blocks from
// the window registers, which then hold row[base + 4k ..], instead of a block of
// three zero taps.
template <int PX>
__device__ __forceinline__ void hconv_blocks(const float* __restrict__ row, int base,
                                             const ctap_t* taps, int ntaps, float (&acc)[PX]) {
    const bool one_left = (ntaps & 3) == 1;
    const int nb = one_left ? (ntaps & ~3) : ((ntaps + 3) & ~3);
#pragma unroll
    for (int q = 0; q < PX; q++) acc[q] = 0.0f;
    float win[PX + 4];
#pragma unroll
    for (int k = 0; k < PX; k++) win[k] = row[hpad<PX>(base + k)];
    for (int b = 0; b < nb; b += 4) {
#pragma unroll
        for (int k = 0; k < 4; k++) win[PX + k] = row[hpad<PX>(base + PX + b + k)];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float wk = taps[b + k];
#pragma unroll
            for (int q = 0; q < PX; q++) acc[q] = mul_add_exact(acc[q], win[q + k], wk);
        }
#pragma unroll
        for (int k = 0; k < PX; k++) win[k] = win[k + 4];
    }
    if (one_left) {
        const float wk = taps[nb];
#pragma unroll
        for (int q = 0; q < PX; q++) acc[q] = mul_add_exact(acc[q], win[q], wk);
    }
}

template <int PX>
__global__ __launch_bounds__(kThreads) void k_blur_h(const BlurJob* __restrict__ jobs, int njobs) {
    extern __shared__ float lds[];
    constexpr int TW = 64 * PX;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int j = find_job(jobs, njobs, tile);
    const BlurJob& J = jobs[j];
    const int t = tile - J.tile_begin;
    const int tx = t % J.tiles_x, ty = t / J.tiles_x;
    const int x0 = tx * TW, y0 = ty * kHTileH;
    const int w = J.w, h = J.h;
    const int r = J.ntaps >> 1;
    const int npad = J.ntaps_pad;
    const int rowlen = TW + npad + 8;
    const int rowstride = hpad<PX>(rowlen + 8) + 1;

    // Wave w stages row y0+w, the words [ws, ws + 4 * nwords) with ws the aligned
    // word start at or before x0-r (clamped) as f32, and reads only that row, so
    // the waves need no work-group barrier. Aligned 4-byte words fully inside
    // [0, w) are read with one load; words at the image edges (and unaligned
    // sources) fall back to clamped bytes. The first KU words of a lane are all
    // loaded before any is written, so their loads overlap.
    const int xs = x0 - r;
    const int ws = xs & ~3;                      // first word start (may be < 0)
    const int d = xs - ws;                       // element base + k is staged element base + d + k
    const int nwords = (d + rowlen + 3) >> 2;
    const bool aligned = (((uintptr_t)J.src | (uintptr_t)J.src_pitch) & 3) == 0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = y0 + wave;
    if (y >= h) return;
    gbyte_t* src = (gbyte_t*)J.src + (size_t)y * J.src_pitch;
    float* dstrow = lds + wave * rowstride;
    auto fetch = [&](int k) -> unsigned {
        const int xw = ws + 4 * k;
        if (aligned && xw >= 0 && xw + 4 <= w)
            return *(__attribute__((address_space(1))) const unsigned*)(src + xw);
        unsigned u = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) u |= (unsigned)src[min(max(xw + b, 0), w - 1)] << (8 * b);
        return u;
    };
    auto put = [&](int k, unsigned u) {  // pixels 4k..4k+3: 4 consecutive slots
        float* dp = dstrow + hpad<PX>(4 * k);
        dp[0] = (float)(u & 0xff);
        dp[1] = (float)((u >> 8) & 0xff);
        dp[2] = (float)((u >> 16) & 0xff);
        dp[3] = (float)(u >> 24);
    };
    constexpr int KU = PX == 16 ? 5 : PX == 8 ? 3 : 1;  // words per lane of a row of the default kernels
    unsigned uw[KU];
#pragma unroll
    for (int i = 0; i < KU; i++)
        if (lane + 64 * i < nwords) uw[i] = fetch(lane + 64 * i);
#pragma unroll
    for (int i = 0; i < KU; i++)
        if (lane + 64 * i < nwords) put(lane + 64 * i, uw[i]);
    for (int k = lane + 64 * KU; k < nwords; k += 64) put(k, fetch(k));
    // the row's LDS writes are complete and visible to the wave before its reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float* row = lds + wave * rowstride;
    const int base = lane * PX;
    const ctap_t* taps = as_const_taps(J.taps);
    float acc[PX];
    bool fixed = false;
    if constexpr (PX >= 8) {
        const float* p = row + hpad<PX>(base);
        fixed = true;
        switch (J.ntaps) {
            case 7: hconv_fixed_d<PX, 7>(d, p, taps, acc); break;
            case 9: hconv_fixed_d<PX, 9>(d, p, taps, acc); break;
            case 11: hconv_fixed_d<PX, 11>(d, p, taps, acc); break;
            case 13: hconv_fixed_d<PX, 13>(d, p, taps, acc); break;
            case 15: hconv_fixed_d<PX, 15>(d, p, taps, acc); break;
            default: fixed = false;
        }
    }
    if (!fixed) hconv_blocks<PX>(row, base + d, taps, J.ntaps, acc);

    const int x = x0 + base;
    if (x >= w) return;
    uint8_t* dst = J.dst + (size_t)y * J.dst_pitch + x;
    unsigned res[PX];
#pragma unroll
    for (int q = 0; q < PX; q++) res[q] = u8_trunc(acc[q]);
    if constexpr (PX == 8) {
        if (x + 8 <= w) {
            uint2 v;
            v.x = res[0] | (res[1] << 8) | (res[2] << 16) | (res[3] << 24);
            v.y = res[4] | (res[5] << 8) | (res[6] << 16) | (res[7] << 24);
            *(uint2*)dst = v;
            return;
        }
    }
    if constexpr (PX == 16) {
        if (x + 16 <= w) {
            uint4 v;
            v.x = res[0] | (res[1] << 8) | (res[2] << 16) | (res[3] << 24);
            v.y = res[4] | (res[5] << 8) | (res[6] << 16) | (res[7] << 24);
            v.z = res[8] | (res[9] << 8) | (res[10] << 16) | (res[11] << 24);
            v.w = res[12] | (res[13] << 8) | (res[14] << 16) | (res[15] << 24);
            *(uint4*)dst = v;
            return;
        }
    }
#pragma unroll
    for (int q = 0; q < PX; q++)
        if (x + q < w) dst[q] = (uint8_t)res[q];
}

// ---------------------------------------------------------------------------
// Vertical pass, wide octaves. Tile = 256 columns x 16 rows; lane l owns columns
// x0+4l..+3 (one u32 load per input row), wave w owns rows y0+4w..+3. For the
// tap counts of the default scale schedule (7 to 15) the taps are unrolled and
// each input row is loaded and converted once; otherwise a register window of
// 8 input rows x 4 columns slides over the taps in blocks of 4.
// Optional DoG epilogue: dog = q[G_s] - q[G_{s-1}], q[v] = v / 255.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cvt4(unsigned v, float* o) {
    o[0] = (float)(v & 0xff);
    o[1] = (float)((v >> 8) & 0xff);
    o[2] = (float)((v >> 16) & 0xff);
    o[3] = (float)(v >> 24);
}

__device__ __forceinline__ void store_row4(const BlurJob& J, int y, int x, const float* acc) {
    unsigned res[4];
#pragma unroll
    for (int c = 0; c < 4; c++) res[c] = u8_trunc(acc[c]);
    uint8_t* dst = J.dst + (size_t)y * J.dst_pitch + x;
    const int w = J.w;
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

// Any tap count: a register window of RV input rows x 4 columns slides over the
// taps in blocks of RV; the remaining taps (< RV) are applied one at a time.
template <int RV>
__device__ __forceinline__ void vconv_blocks(const uint8_t* col, int pitch, int y0, int h, int n,
                                             const ctap_t* taps, float (&acc)[RV][4]) {
#pragma unroll
    for (int q = 0; q < RV; q++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[q][c] = 0.0f;

    float lo[RV][4];
#pragma unroll
    for (int i = 0; i < RV; i++) {
        int yy = min(max(y0 + i, 0), h - 1);
        cvt4(*(const unsigned*)(col + (size_t)yy * pitch), lo[i]);
    }
    int b = 0;
    for (; b + RV <= n; b += RV) {
        float hi[RV][4];
#pragma unroll
        for (int i = 0; i < RV; i++) {
            int yy = min(max(y0 + b + RV + i, 0), h - 1);
            cvt4(*(const unsigned*)(col + (size_t)yy * pitch), hi[i]);
        }
#pragma unroll
        for (int k = 0; k < RV; k++) {
            const float wk = taps[b + k];
#pragma unroll
            for (int q = 0; q < RV; q++) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float p = (q + k < RV) ? lo[q + k][c] : hi[q + k - RV][c];
                    acc[q][c] = mul_add_exact(acc[q][c], p, wk);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < RV; i++)
#pragma unroll
            for (int c = 0; c < 4; c++) lo[i][c] = hi[i][c];
    }
    // remaining taps (< RV): lo holds rows y0 + b .. + RV - 1
    for (; b < n; b++) {
        const float wk = taps[b];
#pragma unroll
        for (int q = 0; q < RV; q++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[q][c] = mul_add_exact(acc[q][c], lo[q][c], wk);
#pragma unroll
        for (int i = 0; i < RV - 1; i++)
#pragma unroll
            for (int c = 0; c < 4; c++) lo[i][c] = lo[i + 1][c];
        int yy = min(max(y0 + b + RV, 0), h - 1);
        cvt4(*(const unsigned*)(col + (size_t)yy * pitch), lo[RV - 1]);
    }
}

// NT taps, fully unrolled: the RV + NT - 1 input rows y0.. (clamped; y0 is
// wave-uniform, so each row is one scalar base plus the lane's column offset)
// are streamed once in increasing order, and input row i adds tap i - q to
// output row q -- per output, the taps in tap order. The first product is the
// first partial sum (p * w0 is never -0).
template <int RV, int NT>
__device__ __forceinline__ void vconv_fixed(const uint8_t* src, unsigned x, int pitch, int y0,
                                            int h, const ctap_t* taps, float (&acc)[RV][4]) {
    float wk[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) wk[k] = taps[k];
#pragma unroll
    for (int i = 0; i < RV + NT - 1; i++) {
        const int yy = min(max(y0 + i, 0), h - 1);
        gbyte_t* rowp = (gbyte_t*)src + (size_t)yy * pitch;
        float p[4];
        cvt4(*(__attribute__((address_space(1))) const unsigned*)(rowp + x), p);
#pragma unroll
        for (int q = 0; q < RV; q++) {
            const int k = i - q;
            if (k < 0 || k >= NT) continue;
#pragma unroll
            for (int c = 0; c < 4; c++)
                acc[q][c] = k == 0 ? __fmul_rn(p[c], wk[0]) : mul_add_exact(acc[q][c], p[c], wk[k]);
        }
    }
}

// RV = kVRows (4) rows per lane: tiles of 256 x 16. RV = 8 (the "v8" form, for
// the unrolled tap counts only): 256 x 32 tiles, twice the arithmetic per wave
// against the same load latency, and 8 + n - 1 input rows per 8 outputs rows.
template <int RV>
__global__ __launch_bounds__(kThreads) void k_blur_v(const BlurJob* __restrict__ jobs, int njobs) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int j = find_job(jobs, njobs, tile);
    const BlurJob& J = jobs[j];
    const int t = tile - J.tile_begin;
    const int tx = t % J.tiles_x, ty = t / J.tiles_x;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int w = J.w, h = J.h;
    const int x = tx * kVTileW + lane * 4;
    const int yb = ty * (4 * RV) + wave * RV;
    if (yb >= h || x >= w) return;
    const int r = J.ntaps >> 1;
    const int n = J.ntaps;
    const int pitch = J.src_pitch;
    const ctap_t* taps = as_const_taps(J.taps);

    float acc[RV][4];
    bool fixed = true;
    switch (n) {
        case 7: vconv_fixed<RV, 7>(J.src, x, pitch, yb - r, h, taps, acc); break;
        case 9: vconv_fixed<RV, 9>(J.src, x, pitch, yb - r, h, taps, acc); break;
        case 11: vconv_fixed<RV, 11>(J.src, x, pitch, yb - r, h, taps, acc); break;
        case 13: vconv_fixed<RV, 13>(J.src, x, pitch, yb - r, h, taps, acc); break;
        case 15: vconv_fixed<RV, 15>(J.src, x, pitch, yb - r, h, taps, acc); break;
        default: fixed = false;
    }
    if constexpr (RV == kVRows) {
        if (!fixed) vconv_blocks<RV>(J.src + x, pitch, yb - r, h, n, taps, acc);
    }
#pragma unroll
    for (int q = 0; q < RV; q++) {
        const int y = yb + q;
        if (y >= h) break;
        store_row4(J, y, x, acc[q]);
    }
}

// ---------------------------------------------------------------------------
// Vertical pass, narrow octaves (and wide kernels). Tile = 64 columns x 16 rows;
// lane = column, wave w owns rows y0+4w..+3. The tile's input rows
// [y0-r, y0+16+r), clamped, are staged once in LDS as u8 (at most min(h, 16+2r)
// distinct rows, so the LDS footprint stays small even when r >> h).
// ---------------------------------------------------------------------------
constexpr int kVnTileW = 64, kVnTileH = 16;

__global__ __launch_bounds__(kThreads) void k_blur_v_narrow(const BlurJob* __restrict__ jobs,
                                                            int njobs) {
    extern __shared__ uint8_t vl[];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int j = find_job(jobs, njobs, tile);
    const BlurJob& J = jobs[j];
    const int t = tile - J.tile_begin;
    const int tx = t % J.tiles_x, ty = t / J.tiles_x;
    const int w = J.w, h = J.h;
    const int r = J.ntaps >> 1;
    const int n = J.ntaps;
    const int x0 = tx * kVnTileW, y0 = ty * kVnTileH;
    const int rlo = max(y0 - r, 0), rhi = min(y0 + kVnTileH + r, h);  // distinct rows [rlo, rhi)
    const int nrows = rhi - rlo;
    for (int i = threadIdx.x; i < nrows * kVnTileW; i += kThreads) {
        const int rr = i / kVnTileW, c = i - rr * kVnTileW;
        const int x = min(x0 + c, w - 1);
        vl[i] = J.src[(size_t)(rlo + rr) * J.src_pitch + x];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x = x0 + lane;
    const int yb = y0 + wave * 4;
    if (x >= w || yb >= h) return;
    const ctap_t* taps = as_const_taps(J.taps);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float cur[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int yy = min(max(yb + q - r, 0), h - 1);
        cur[q] = (float)vl[(yy - rlo) * kVnTileW + lane];
    }
    for (int i = 0; i < n; i++) {
        const float wk = taps[i];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = mul_add_exact(acc[q], cur[q], wk);
#pragma unroll
        for (int q = 0; q < 3; q++) cur[q] = cur[q + 1];
        if (i + 1 < n) {  // next input row; stays inside the staged rows
            const int yy = min(max(yb + 4 + i - r, 0), h - 1);
            cur[3] = (float)vl[(yy - rlo) * kVnTileW + lane];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int y = yb + q;
        if (y >= h) break;
        const unsigned v = u8_trunc(acc[q]);
        J.dst[(size_t)y * J.dst_pitch + x] = (uint8_t)v;
        if (J.dog)
            J.dog[(size_t)y * J.dog_pitch + x] = q255(v) - q255(J.prev[(size_t)y * J.dst_pitch + x]);
    }
}

// ---------------------------------------------------------------------------
// Small levels (the deep octaves: a few thousand pixels, blur radii up to
// several hundred): one output per lane, so each wave's serial tap chain is as
// short as it can be and the few tiles still spread over many waves.
// Tile = 64 columns x 4 rows, wave w computes row y0+w, lane l column x0+l.
// The taps (zero-padded to a multiple of 4) sit in LDS and are read four at a
// time by all lanes at once. Every output keeps the reference's tap order and
// separate f32 multiply and add.
// ---------------------------------------------------------------------------
constexpr int kSTileW = 64, kSTileH = 4;

// Horizontal: the clamped row segments [x0-r, x0+64+n4-r) of the tile's 4 rows
// are staged as f32 (segment stride seg_stride >= 64 + n4).
__global__ __launch_bounds__(kThreads) void k_blur_h_small(const BlurJob* __restrict__ jobs,
                                                           int njobs, int seg_stride) {
    extern __shared__ float sm[];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x0 = (t % J.tiles_x) * kSTileW, y0 = (t / J.tiles_x) * kSTileH;
    const int w = J.w, h = J.h, n = J.ntaps, r = n >> 1;
    const int n4 = (n + 3) & ~3;
    float* tp = sm;
    float* seg = sm + n4;
    for (int i = threadIdx.x; i < n4; i += kThreads) tp[i] = i < n ? J.taps[i] : 0.0f;
    const int segw = kSTileW + n4;
    const float inv_segw = 1.0f / (float)segw;
    for (int i = threadIdx.x; i < kSTileH * segw; i += kThreads) {
        const int rr = (int)(((float)i + 0.5f) * inv_segw), c = i - rr * segw;
        const int y = min(y0 + rr, h - 1);
        const int x = min(max(x0 - r + c, 0), w - 1);
        seg[rr * seg_stride + c] = (float)J.src[(size_t)y * J.src_pitch + x];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = y0 + wave, x = x0 + lane;
    if (y >= h || x >= w) return;
    const float* sg = seg + wave * seg_stride + lane;
    float acc = 0.0f;
#pragma unroll 2
    for (int k = 0; k < n4; k += 4) {
        const float4 t4 = *reinterpret_cast<const float4*>(tp + k);
        acc = mul_add_exact(acc, sg[k], t4.x);
        acc = mul_add_exact(acc, sg[k + 1], t4.y);
        acc = mul_add_exact(acc, sg[k + 2], t4.z);
        acc = mul_add_exact(acc, sg[k + 3], t4.w);
    }
    J.dst[(size_t)y * J.dst_pitch + x] = (uint8_t)u8_trunc(acc);
}

// acc + p * tp[k] for k in [k0, k1) in order (one input value, a run of taps):
// the taps are read 4 at a time once k is a multiple of 4 (tp is 16-byte aligned).
__device__ __forceinline__ float run_const(float acc, float p, const float* tp, int k0, int k1) {
    int k = k0;
    for (; k < k1 && (k & 3); k++) acc = mul_add_exact(acc, p, tp[k]);
#pragma unroll 2
    for (; k + 4 <= k1; k += 4) {
        const float4 t4 = *reinterpret_cast<const float4*>(tp + k);
        acc = mul_add_exact(acc, p, t4.x);
        acc = mul_add_exact(acc, p, t4.y);
        acc = mul_add_exact(acc, p, t4.z);
        acc = mul_add_exact(acc, p, t4.w);
    }
    for (; k < k1; k++) acc = mul_add_exact(acc, p, tp[k]);
    return acc;
}

// Vertical: the tile's distinct input rows [max(0, y0-r), min(h, y0+4+r)) are
// staged (at most h rows): as f32 (k_blur_v_small, up to kSmallMaxRows rows), or
// as bytes (k_blur_v_small8: 4x the rows in the same LDS, so levels of a few
// hundred rows with radii of dozens of pixels, the middle octaves, also get one
// output per lane; each tap then converts its byte). For output row y, taps
// k < r-y read row 0 and taps k >= h+r-y read row h-1 (clamp to edge): those
// runs need no loads and their bounds are the same for the whole wave.
template <typename T>
__device__ __forceinline__ void blur_v_small_tile(const BlurJob* __restrict__ jobs, int njobs,
                                                  int max_taps4, float* sm) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x0 = (t % J.tiles_x) * kSTileW, y0 = (t / J.tiles_x) * kSTileH;
    const int w = J.w, h = J.h, n = J.ntaps, r = n >> 1;
    const int n4 = (n + 3) & ~3;
    float* tp = sm;
    T* col = reinterpret_cast<T*>(sm + max_taps4);
    for (int i = threadIdx.x; i < n4; i += kThreads) tp[i] = i < n ? J.taps[i] : 0.0f;
    const int rlo = max(y0 - r, 0), rhi = min(y0 + kSTileH + r, h);
    for (int i = threadIdx.x; i < (rhi - rlo) * kSTileW; i += kThreads) {
        const int rr = i >> 6, c = i & 63;
        col[i] = (T)J.src[(size_t)(rlo + rr) * J.src_pitch + min(x0 + c, w - 1)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int y = y0 + wave, x = x0 + lane;
    if (y >= h || x >= w) return;
    const int ka = min(max(r - y, 0), n);      // taps reading row 0
    const int kc = max(min(h + r - y, n), ka);  // taps from kc on read row h-1
    float acc = 0.0f;
    if (ka > 0) acc = run_const(acc, (float)col[(0 - rlo) * kSTileW + lane], tp, 0, ka);
    const T* cp = col + (y - r - rlo) * kSTileW + lane;
#pragma unroll 4
    for (int k = ka; k < kc; k++) acc = mul_add_exact(acc, (float)cp[k * kSTileW], tp[k]);
    if (kc < n) acc = run_const(acc, (float)col[(h - 1 - rlo) * kSTileW + lane], tp, kc, n);
    const unsigned v = u8_trunc(acc);
    J.dst[(size_t)y * J.dst_pitch + x] = (uint8_t)v;
    if (J.dog)
        J.dog[(size_t)y * J.dog_pitch + x] = q255(v) - q255(J.prev[(size_t)y * J.dst_pitch + x]);
}

__global__ __launch_bounds__(kThreads) void k_blur_v_small(const BlurJob* __restrict__ jobs,
                                                           int njobs, int max_taps4) {
    extern __shared__ __align__(16) float sm[];
    blur_v_small_tile<float>(jobs, njobs, max_taps4, sm);
}

__global__ __launch_bounds__(kThreads) void k_blur_v_small8(const BlurJob* __restrict__ jobs,
                                                            int njobs, int max_taps4) {
    extern __shared__ __align__(16) float sm8[];
    blur_v_small_tile<uint8_t>(jobs, njobs, max_taps4, sm8);
}

// Exact thread-per-pixel passes for kernels too wide for the LDS row tile.
__global__ __launch_bounds__(kThreads) void k_blur_h_generic(const BlurJob* __restrict__ jobs,
                                                             int njobs) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
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
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
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
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const BlurJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int x = (t % J.tiles_x) * kThreads + threadIdx.x, y = t / J.tiles_x;
    if (x >= J.w || y >= J.h) return;
    uint8_t v = J.src[(size_t)y * J.src_pitch + x];
    J.dst[(size_t)y * J.dst_pitch + x] = v;
    if (J.dog) J.dog[(size_t)y * J.dog_pitch + x] = q255(v) - q255(v);
}

// ---------------------------------------------------------------------------
// Colour -> grey before the pyramid. RGB8 / RGBA8: DynamicImage::to_luma8 of
// image 0.25.6 (color.rs rgb_to_luma: u32 sum 2126 R + 7152 G + 722 B, divided
// by 10000 with truncation; alpha dropped), the conversion at src/sift.rs:1202.
// LUMAA8: the luma byte. RGBA8_CANVAS: src/wasm.rs:88-94, f32
// (0.299 r + 0.587 g) + 0.114 b without FMA, then `as u8` (truncate, saturate).
// A lane converts 4 consecutive pixels of one row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned luma_srgb(unsigned r, unsigned g, unsigned b) {
    return (2126u * r + 7152u * g + 722u * b) / 10000u;
}

__device__ __forceinline__ unsigned luma_canvas(unsigned r, unsigned g, unsigned b) {
    const float l = __fadd_rn(__fadd_rn(__fmul_rn(0.299f, (float)r), __fmul_rn(0.587f, (float)g)),
                              __fmul_rn(0.114f, (float)b));
    return u8_trunc(l);
}

__global__ __launch_bounds__(kThreads) void k_luma(const LumaJob* __restrict__ jobs, int njobs) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const LumaJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int y = t / J.tiles_x;
    const int x = (t % J.tiles_x) * kLumaTileW + threadIdx.x * 4;
    if (x >= J.w || y >= J.h) return;
    const int n = min(4, J.w - x);
    const uint8_t* src = J.src + (size_t)y * J.src_pitch;
    uint8_t* dst = J.dst + (size_t)y * J.dst_pitch + x;
    unsigned l[4] = {0u, 0u, 0u, 0u};
    const int f = J.format;
    const int bpp = f == SIFT_PIXEL_RGB8 ? 3 : f == SIFT_PIXEL_LUMAA8 ? 2 : f == SIFT_PIXEL_GRAY8 ? 1 : 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k >= n) break;
        const uint8_t* p = src + (size_t)(x + k) * bpp;
        if (bpp >= 3) {
            const unsigned r = p[0], g = p[1], b = p[2];
            l[k] = f == SIFT_PIXEL_RGBA8_CANVAS ? luma_canvas(r, g, b) : luma_srgb(r, g, b);
        } else {
            l[k] = p[0];
        }
    }
    if (n == 4 && (((uintptr_t)dst) & 3) == 0) {
        *(unsigned*)dst = l[0] | (l[1] << 8) | (l[2] << 16) | (l[3] << 24);
    } else {
        for (int k = 0; k < n; k++) dst[k] = (uint8_t)l[k];
    }
}

// ---------------------------------------------------------------------------
// Lanczos3 resize (image 0.25.6 sample::resize): vertical pass into f32 rows
// (t += p * w in window order, unrounded), then horizontal pass with
// round(clamp(t, 0, 255)). Exact reference order (separate mul and add).
// Vertical: a thread computes 4 adjacent columns of one output row (one u32
// load per tap). Horizontal: a workgroup computes 256 outputs of one row from
// the input span they need, staged in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_resize_v(const ResizeJob* __restrict__ jobs,
                                                       int njobs) {
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const ResizeJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int oy = t / J.tiles_x;
    const int x = ((t % J.tiles_x) * kThreads + threadIdx.x) * 4;
    if (x >= J.in_w) return;
    const int l = J.left[oy], n = J.count[oy];
    const ctap_t* wr = as_const_taps(J.weights) + (size_t)oy * J.maxtaps;  // one row for the wave
    const uint8_t* src = J.src_u8 + (size_t)l * J.src_pitch + x;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int i = 0;
    for (; i + 4 <= n; i += 4) {
        unsigned u[4];
#pragma unroll
        for (int k = 0; k < 4; k++) u[k] = *(const unsigned*)(src + (size_t)(i + k) * J.src_pitch);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float wk = wr[i + k];
            float p[4];
            cvt4(u[k], p);
#pragma unroll
            for (int c = 0; c < 4; c++) s[c] = mul_add_exact(s[c], p[c], wk);
        }
    }
    for (; i < n; i++) {
        const float wk = wr[i];
        float p[4];
        cvt4(*(const unsigned*)(src + (size_t)i * J.src_pitch), p);
#pragma unroll
        for (int c = 0; c < 4; c++) s[c] = mul_add_exact(s[c], p[c], wk);
    }
    float* out = J.tmp + (size_t)oy * J.tmp_pitch + x;
    if (x + 4 <= J.in_w) {
        *(float4*)out = make_float4(s[0], s[1], s[2], s[3]);
    } else {
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (x + c < J.in_w) out[c] = s[c];
    }
}

constexpr int kResizeSpan = 2 * kThreads + 64;  // input floats for 256 outputs (ratio ~2)

__global__ __launch_bounds__(kThreads) void k_resize_h(const ResizeJob* __restrict__ jobs,
                                                       int njobs) {
    __shared__ float span[kResizeRows][kResizeSpan];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const ResizeJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int y0 = (t / J.tiles_x) * kResizeRows;
    const int ox0 = (t % J.tiles_x) * kThreads;
    const int ox = ox0 + threadIdx.x;
    const int oxl = min(ox0 + kThreads, J.out_w) - 1;
    const int nrows = min(kResizeRows, J.out_h - y0);
    const int lo = J.left[ox0];
    const int hi = J.left[oxl] + J.count[oxl];  // exclusive
    const bool staged = hi - lo <= kResizeSpan;
    if (staged) {
        for (int r = 0; r < nrows; r++) {
            const float* src = J.tmp + (size_t)(y0 + r) * J.tmp_pitch;
            for (int i = threadIdx.x; i < hi - lo; i += kThreads) span[r][i] = src[lo + i];
        }
    }
    __syncthreads();
    if (ox >= J.out_w) return;
    const int l = J.left[ox], n = J.count[ox];
    const float* wr = J.weights + (size_t)ox * J.maxtaps;
    float wv[16];  // this column's weights, loaded once for all rows
    if (n <= 16) {
#pragma unroll
        for (int i = 0; i < 16; i++) wv[i] = i < n ? wr[i] : 0.0f;
    }
    for (int r = 0; r < nrows; r++) {
        const float* in = staged ? &span[r][l - lo] : J.tmp + (size_t)(y0 + r) * J.tmp_pitch + l;
        float s = 0.0f;
        if (n <= 16) {
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i < n) s = mul_add_exact(s, in[i], wv[i]);
        } else {
            for (int i = 0; i < n; i++) s = mul_add_exact(s, in[i], wr[i]);
        }
        float c = s < 0.0f ? 0.0f : (s > 255.0f ? 255.0f : s);
        J.dst[(size_t)(y0 + r) * J.dst_pitch + ox] = (uint8_t)roundf(c);
    }
}

// ---------------------------------------------------------------------------
// Fused Lanczos3 resize: one work-group computes kResizeRows output rows x 256
// output columns. The vertical pass for those rows over the input columns the
// tile's horizontal windows need, [lo4, hi) with lo4 = left[ox0] rounded down to
// a multiple of 4, goes to LDS as f32 (the same t += p * w in window order as
// k_resize_v: wave w takes output rows w and w + 4, so its weights are
// wave-uniform, and lane l the column quads l, l + 64, ...); the horizontal pass
// then reads it from LDS as k_resize_h does. The f32 intermediate never goes to
// HBM. The host takes this kernel only when every tile's span fits
// kResizeFusedSpan (octave steps, ratio ~2, need about 524 columns).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_resize_fused(const ResizeJob* __restrict__ jobs,
                                                           int njobs) {
    __shared__ __align__(16) float vt[kResizeRows][kResizeFusedSpan];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const ResizeJob& J = jobs[find_job(jobs, njobs, tile)];
    const int t = tile - J.tile_begin;
    const int y0 = (t / J.tiles_x) * kResizeRows;
    const int ox0 = (t % J.tiles_x) * kThreads;
    const int oxl = min(ox0 + kThreads, J.out_w) - 1;
    const int nrows = min(kResizeRows, J.out_h - y0);
    const int lo4 = J.left[ox0] & ~3;
    const int hi = J.left[oxl] + J.count[oxl];  // exclusive, <= in_w
    const int nq = (hi - lo4 + 3) >> 2;         // column quads: 4 * nq <= kResizeFusedSpan
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int in_w = J.in_w, pitch = J.src_pitch;
    for (int r = wave; r < nrows; r += kThreads / 64) {
        const int oy = y0 + r;
        const int l = J.vleft[oy], n = J.vcount[oy];
        const ctap_t* wr = as_const_taps(J.vweights) + (size_t)oy * J.vmaxtaps;
        for (int q = lane; q < nq; q += 64) {
            const int x = lo4 + 4 * q;
            const uint8_t* src = J.src_u8 + (size_t)l * pitch + x;
            float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (x + 4 <= in_w) {
                int i = 0;
                for (; i + 4 <= n; i += 4) {
                    unsigned u[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) u[k] = *(const unsigned*)(src + (size_t)(i + k) * pitch);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float wk = wr[i + k];
                        float p[4];
                        cvt4(u[k], p);
#pragma unroll
                        for (int c = 0; c < 4; c++) s[c] = mul_add_exact(s[c], p[c], wk);
                    }
                }
                for (; i < n; i++) {
                    const float wk = wr[i];
                    float p[4];
                    cvt4(*(const unsigned*)(src + (size_t)i * pitch), p);
#pragma unroll
                    for (int c = 0; c < 4; c++) s[c] = mul_add_exact(s[c], p[c], wk);
                }
            } else {  // the quad at the right edge: columns >= in_w are never read back
                for (int i = 0; i < n; i++) {
                    const float wk = wr[i];
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        if (x + c < in_w) s[c] = mul_add_exact(s[c], (float)src[(size_t)i * pitch + c], wk);
                }
            }
            *(float4*)&vt[r][4 * q] = make_float4(s[0], s[1], s[2], s[3]);
        }
    }
    __syncthreads();
    const int ox = ox0 + threadIdx.x;
    if (ox >= J.out_w) return;
    const int l = J.left[ox], n = J.count[ox];
    const float* wr = J.weights + (size_t)ox * J.maxtaps;
    float wv[16];  // this column's weights, loaded once for all rows
    if (n <= 16) {
#pragma unroll
        for (int i = 0; i < 16; i++) wv[i] = i < n ? wr[i] : 0.0f;
    }
    for (int r = 0; r < nrows; r++) {
        const float* in = &vt[r][l - lo4];
        float s = 0.0f;
        if (n <= 16) {
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i < n) s = mul_add_exact(s, in[i], wv[i]);
        } else {
            for (int i = 0; i < n; i++) s = mul_add_exact(s, in[i], wr[i]);
        }
        float c = s < 0.0f ? 0.0f : (s > 255.0f ? 255.0f : s);
        J.dst[(size_t)(y0 + r) * J.dst_pitch + ox] = (uint8_t)roundf(c);
    }
}

}  // namespace

void launch_blur_h(const BlurJob* d_jobs, int njobs, int total_tiles, int max_pad, int px,
                   hipStream_t s) {
    if (total_tiles <= 0) return;
    const int rowlen = 64 * px + max_pad + 8;
    const int padded = px == 16 ? hpad<16>(rowlen + 8) : px == 8 ? hpad<8>(rowlen + 8) : hpad<2>(rowlen + 8);
    const size_t lds = (size_t)kHTileH * (padded + 1) * sizeof(float);
    if (px == 16)
        hipLaunchKernelGGL(k_blur_h<16>, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs);
    else if (px == 8)
        hipLaunchKernelGGL(k_blur_h<8>, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs);
    else
        hipLaunchKernelGGL(k_blur_h<2>, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs);
}

void launch_blur_v(const BlurJob* d_jobs, int njobs, int total_tiles, int narrow_rows,
                   hipStream_t s) {
    if (total_tiles <= 0) return;
    if (narrow_rows > 0) {
        size_t lds = (size_t)narrow_rows * kVnTileW;
        hipLaunchKernelGGL(k_blur_v_narrow, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs);
    } else {
        hipLaunchKernelGGL(k_blur_v<kVRows>, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
    }
}

void launch_blur_v8(const BlurJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_blur_v<8>, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_blur_h_small(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                         hipStream_t s) {
    if (total_tiles <= 0) return;
    const int seg_stride = kSTileW + max_taps4 + 1;
    const size_t lds = ((size_t)max_taps4 + (size_t)kSTileH * seg_stride) * sizeof(float);
    hipLaunchKernelGGL(k_blur_h_small, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs,
                       seg_stride);
}

void launch_blur_v_small(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                         int max_rows, hipStream_t s) {
    if (total_tiles <= 0) return;
    const size_t lds = ((size_t)max_taps4 + (size_t)max_rows * kSTileW) * sizeof(float);
    hipLaunchKernelGGL(k_blur_v_small, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs,
                       max_taps4);
}

void launch_blur_v_small8(const BlurJob* d_jobs, int njobs, int total_tiles, int max_taps4,
                          int max_rows, hipStream_t s) {
    if (total_tiles <= 0) return;
    const size_t lds = (size_t)max_taps4 * sizeof(float) + (size_t)max_rows * kSTileW;
    hipLaunchKernelGGL(k_blur_v_small8, dim3(total_tiles), dim3(kThreads), lds, s, d_jobs, njobs,
                       max_taps4);
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

void launch_luma(const LumaJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_luma, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_resize_v(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_resize_v, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_resize_fused(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_resize_fused, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

void launch_resize_h(const ResizeJob* d_jobs, int njobs, int total_tiles, hipStream_t s) {
    if (total_tiles <= 0) return;
    hipLaunchKernelGGL(k_resize_h, dim3(total_tiles), dim3(kThreads), 0, s, d_jobs, njobs);
}

}  // namespace sh
