// kernels_detect.hip -- gfx950 kernels for everything after the pyramid:
//   k_extrema_tile / k_extrema  find_scale_space_extrema  src/sift.rs:510-603
//   k_scan_*      ordered compaction offsets (exclusive scan)
//   k_compact     candidate list in the reference order (octave, layer, row, column)
//   k_refine      refine_and_filter_extrema     src/sift.rs:299-474 (+ solve 478-503)
//   k_orient      assign_orientations           src/sift.rs:605-726
//   k_descriptor  compute + normalize           src/sift.rs:917-1060, 889-913
//   k_descriptor_wide  the same for windows wider than the LDS patch
//
// DoG values are never stored: every DoG sample is recomputed from the u8
// Gaussian levels as q[G_{s+1}] - q[G_s] with q[v] = v / 255 (IEEE division,
// formed without it by q255f, or read from an LDS table of it in the row-wise
// extremum and wide-window kernels), which is bit-identical to the reference's
// (G_{s+1} as f32 / 255.0) - (G_s as f32 / 255.0) (src/sift.rs:97,120).
#include <algorithm>
#include <map>
#include <mutex>

#include "sift_internal.h"

namespace sh {

namespace {

constexpr float kPi = 3.14159274101257324219f;  // std::f32::consts::PI
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kDescFix = 1048576.0f;  // 2^20: u32 fixed point of the descriptor sums, times a per-keypoint power-of-two gain
constexpr int kThreads = 256;
constexpr int kExtChunk = 1024;  // pixels of one row per chunk in k_extrema
constexpr int kScanTile = 4096;  // 256 threads x 16 elements

// Rust `f as i32`: truncation toward zero, saturating, NaN -> 0.
__device__ __forceinline__ int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

// f32::rem_euclid: r = x % m; r < 0 ? r + |m| : r.
__device__ __forceinline__ float rem_euclid(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}

__device__ __forceinline__ float pow2f(int o) { return ldexpf(1.0f, o); }

// atan2 for the gradient angles: octant reduction a = min(|x|,|y|)/max(|x|,|y|),
// atan(a) = a * P(a^2) with P a degree-8 least-squares (Lawson-reweighted)
// fit to atan(a)/a on [0,1], evaluated with FMA (max error 1.0e-7 rad on [0,1];
// about 2e-7 rad after the quadrant corrections, comparable to a 1-2 ulp
// atan2f). Gradients here are differences of non-negative values, never -0.
__device__ __forceinline__ float atan2_grad(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    // selects rather than fmaxf / fminf: the operands are never NaN, and one
    // comparison serves both and the octant test below
    const bool steep = ay > ax;
    const float mx = steep ? ay : ax, mn = steep ? ax : ay;
    const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = a * a;
    float p = 0.002456706715747714f;
    p = __builtin_fmaf(p, s, -0.014401290565729141f);
    p = __builtin_fmaf(p, s, 0.0397811196744442f);
    p = __builtin_fmaf(p, s, -0.07234849780797958f);
    p = __builtin_fmaf(p, s, 0.1049894317984581f);
    p = __builtin_fmaf(p, s, -0.14161229133605957f);
    p = __builtin_fmaf(p, s, 0.19985906779766083f);
    p = __builtin_fmaf(p, s, -0.33332598209381104f);
    p = __builtin_fmaf(p, s, 0.9999998807907104f);
    float r = a * p;
    if (steep) r = 1.57079637050628662109f - r;
    if (x < 0.0f) r = 3.14159274101257324219f - r;
    return copysignf(r, y);
}

// Gradient angle in descriptor bin units (pi/4 per bin, result in (-4, 4]):
// the same octant reduction, with a 5-term fit of atan(a)/a scaled by 4/pi
// (max error 1.5e-5 bin, 1.2e-5 rad; far inside the descriptor tolerance).
__device__ __forceinline__ float atan2_bins(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mn * __builtin_amdgcn_rcpf(fmaxf(mx, 1e-30f));  // 0 when x = y = 0
    const float s = a * a;
    float p = 0.026541974395513535f;
    p = __builtin_fmaf(p, s, -0.10842673480510712f);
    p = __builtin_fmaf(p, s, 0.22938744723796844f);
    p = __builtin_fmaf(p, s, -0.4205574691295624f);
    p = __builtin_fmaf(p, s, 1.2730693817138672f);
    float r = a * p;
    if (ay > ax) r = 2.0f - r;
    if (x < 0.0f) r = 4.0f - r;
    return copysignf(r, y);
}

// x / 2pi (kTwoPi) for x in [0, 36 * 2pi], as y = x * RN(1/2pi) plus one FMA
// correction: three full-rate operations in place of the IEEE division (about
// eleven). The orientation binning only takes the floor, and that equals the
// floor of the IEEE quotient for every f32 x of the range (checked
// exhaustively on the CPU, tests/c/div_two_pi_check.c); the quotient itself can
// differ in the last place, and the product alone changes the floor of 7 x.
__device__ __forceinline__ float div_two_pi(float x) {
    constexpr float rc = 1.0f / kTwoPi;
    const float y = x * rc;
    return __builtin_fmaf(__builtin_fmaf(-y, kTwoPi, x), rc, y);
}

// A wave-uniform f32 value moved to a scalar register (the compiler cannot prove
// that values read from LDS or computed in VALU from uniform inputs are uniform).
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// v / 255 in f32 without a table, for v an integer in [0, 255] held as f32:
// 1/255 split into c = RN(1/255) and cl = RN(1/255 - c), then v * c + v * cl
// with one FMA. Equal to the IEEE division v / 255.0f for every byte value
// (checked exhaustively: v / 255 lies at least 2^-33 relative from any rounding
// midpoint, and the split's error is near 2^-48); three full-rate f32
// operations with the conversion.
__device__ __forceinline__ float q255f(float v) {
    constexpr float c = 0x1.010102p-8f;
    constexpr float cl = -0x1.fdfdfep-33f;
    return __builtin_fmaf(v, c, v * cl);
}

__device__ __forceinline__ void fill_q(float* q) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) q[i] = __fdiv_rn((float)i, 255.0f);
}

template <typename T>
__device__ __forceinline__ int find_by(const T* __restrict__ a, int n, int key,
                                       int T::*field) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a[mid].*field <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long ballot64(bool p) {
    return __ballot(p);
}

// ---------------------------------------------------------------------------
// k_extrema: one workgroup per (octave, row). For each layer s in 1..S it stages
// DoG rows y-1, y, y+1 (layers s-1, s, s+1) of a 1024-pixel chunk in LDS and
// tests each pixel against its 26 neighbours with the reference's strict rule
// (val <= n clears is_max, val >= n clears is_min). A wave ballot turns the 64
// results of 64 consecutive pixels into one bitmask word; popcounts give the
// per-(layer,row) candidate count. Rows outside the border get count 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_extrema(const OctaveDesc* __restrict__ oct, int noct,
                                                      const ImageParams* __restrict__ params,
                                                      unsigned long long* __restrict__ bits,
                                                      unsigned* __restrict__ counts) {
    __shared__ float q[256];
    __shared__ float D[3][3][kExtChunk + 2];
    __shared__ unsigned cnt;
    const int row = blockIdx.x;
    const int j = find_by(oct, noct, row, &OctaveDesc::row_begin);
    const OctaveDesc& O = oct[j];
    const ImageParams& P = params[O.img];
    const int y = row - O.row_begin;
    const int w = O.w, h = O.h, S = P.num_intervals, border = P.border;
    const bool row_ok = (y >= border) && (y < h - border);
    if (!row_ok) {
        for (int s = 1 + threadIdx.x; s <= S; s += kThreads) counts[O.slot_begin + (s - 1) * h + y] = 0;
        return;
    }
    fill_q(q);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int s = 1; s <= S; s++) {
        if (threadIdx.x == 0) cnt = 0;
        for (int cx0 = 0; cx0 < w; cx0 += kExtChunk) {
            __syncthreads();
            // stage DoG layers s-1..s+1, rows y-1..y+1, x in [cx0-1, cx0+kExtChunk+1)
            const int nstage = kExtChunk + 2;
            for (int i = threadIdx.x; i < 3 * 3 * nstage; i += kThreads) {
                const int l = i / (3 * nstage);
                const int rem = i - l * 3 * nstage;
                const int rr = rem / nstage;
                const int c = rem - rr * nstage;
                const int x = min(max(cx0 - 1 + c, 0), w - 1);
                const int yy = min(max(y - 1 + rr, 0), h - 1);
                const int lay = s - 1 + l;   // DoG layer: G_{lay+1} - G_{lay}
                const uint8_t* g0 = O.levels + (long long)lay * O.level_stride + (size_t)yy * O.pitch;
                const uint8_t* g1 = g0 + O.level_stride;
                D[l][rr][c] = q[g1[x]] - q[g0[x]];
            }
            __syncthreads();
            for (int k = 0; k < kExtChunk / kThreads; k++) {
                const int c = k * kThreads + threadIdx.x;  // wave-contiguous 64 pixels
                const int x = cx0 + c;
                bool ext = false;
                if (x >= border && x < w - border) {
                    const float v = D[1][1][c + 1];
                    bool is_max = true, is_min = true;
#pragma unroll
                    for (int l = 0; l < 3; l++) {
#pragma unroll
                        for (int rr = 0; rr < 3; rr++) {
#pragma unroll
                            for (int dx = 0; dx < 3; dx++) {
                                if (l == 1 && rr == 1 && dx == 1) continue;
                                const float nv = D[l][rr][c + dx];
                                is_max = is_max && (v > nv);
                                is_min = is_min && (v < nv);
                            }
                        }
                    }
                    ext = is_max || is_min;
                }
                const unsigned long long m = ballot64(ext);
                const int word_x = cx0 + k * kThreads + wave * 64;
                if (lane == 0 && word_x < w) {
                    bits[O.bits_begin + (long long)((s - 1) * h + y) * O.nwords + (word_x >> 6)] = m;
                    if (m) atomicAdd(&cnt, (unsigned)__popcll(m));
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) counts[O.slot_begin + (s - 1) * h + y] = cnt;
    }
}

// ---------------------------------------------------------------------------
// k_extrema_tile<S>: the same test on strips of 64 columns x strip_rows rows
// (16 or kExtStripRows), taken as bands of 16 rows. The S+2 DoG layers of a band plus a 1-pixel halo are
// staged in LDS as f32 (each thread holds one 8-byte word of every Gaussian level
// and writes q[G_{l+1}] - q[G_l] for its 8 pixels); the words of the next band
// are loaded into registers before the current band is tested, so their latency
// overlaps the test. In a band, lane l owns column x0+l and walks down 4 rows (one
// wave per 4 rows). Per row it reads the DoG values of every layer at x-1, x, x+1
// and forms their horizontal max / min; a rolling window of three rows gives the
// 3x3 max / min of each layer.
// The 26-neighbour maximum of layer s is then max(3x3 of s-1, 3x3 of s+1,
// row max above and below in s, left and right in s), so "v greater than every
// neighbour" is v > that maximum (likewise for the minimum) -- the reference's
// strict comparisons (src/sift.rs:571-576). The 64 results of a row form one
// bitmask word; the row-slot counts are integer atomics (order-independent).
// ---------------------------------------------------------------------------
constexpr int kTileW = 64, kTileH = 16;
static_assert(kExtStripRows % kTileH == 0, "strips are whole bands");
constexpr int kExtNW = (kTileW + 16) / 8;         // 8-byte words covering [x0-8, x0+72)
constexpr int kExtItems = (kTileH + 2) * kExtNW;  // (row, word) items of a band
constexpr int kExtRW = kTileW + 8;                // staged D row: columns [x0-4, x0+68)
static_assert(kExtItems <= kThreads, "one item per thread");
typedef __attribute__((address_space(1))) const uint8_t gbyte_t;
typedef __attribute__((address_space(1))) const unsigned long long gu64_t;

// The NG level words of this thread's item of the band starting at row y0 (rows
// y0-1 .. y0+16, clamped): 8-byte words inside [0, w) with one load (x0 is a
// multiple of 64 and rows are 64-byte aligned), the others from clamped bytes.
template <int NG>
__device__ __forceinline__ void ext_fetch(const OctaveDesc& O, int x0, int y0,
                                          unsigned long long (&u)[NG]) {
    const int i = threadIdx.x;
    if (i >= kExtItems) return;
    const int w = O.w, h = O.h;
    const int r = i / kExtNW, k = i - r * kExtNW;
    const int y = min(max(y0 - 1 + r, 0), h - 1);
    const int xw = x0 - 8 + 8 * k;
    gbyte_t* const lv = (gbyte_t*)O.levels;
    // fits 32 bits: images are at most 65536 x 65536 (larger ones are rejected)
    const unsigned rowoff = (unsigned)y * (unsigned)O.pitch;
    if (xw >= 0 && xw + 8 <= w) {
#pragma unroll
        for (int l = 0; l < NG; l++)
            u[l] = *(gu64_t*)(lv + (size_t)l * O.level_stride + rowoff + (unsigned)xw);
    } else {
#pragma unroll
        for (int l = 0; l < NG; l++) {
            unsigned long long v = 0;
#pragma unroll 1
            for (int b = 0; b < 8; b++)
                v |= (unsigned long long)lv[(size_t)l * O.level_stride + rowoff +
                                             (unsigned)min(max(xw + b, 0), w - 1)]
                     << (8 * b);
            u[l] = v;
        }
    }
}

// D = q[G_{l+1}] - q[G_l] for columns [x0-4, x0+68) of the band's 18 rows (the
// test reads [x0-1, x0+65)), with q formed by q255f, equal to v / 255.0f for
// every byte value; D is the same f32 subtraction the reference makes. Column
// x0-4+j sits at D index j, so the 4-column halves of a thread's 8-byte word
// (columns x0-8+8k ..) start at 16-byte boundaries and each goes to LDS as one
// 16-byte write (the first half of word 0 and the second half of the last word
// lie outside the staged columns).
template <int NL, int NG>
__device__ __forceinline__ void ext_stage(const unsigned long long (&u)[NG],
                                          float (*D)[kTileH + 2][kExtRW]) {
    const int i = threadIdx.x;
    if (i >= kExtItems) return;
    const int r = i / kExtNW, k = i - r * kExtNW;
#pragma unroll
    for (int hh = 0; hh < 2; hh++) {
        if (hh == 0 ? k == 0 : k == kExtNW - 1) continue;
        float* const dst = &D[0][r][8 * k + 4 * hh - 4];
        const unsigned w0 = (unsigned)(u[0] >> (32 * hh));
        float qlo[4];
#pragma unroll
        for (int b = 0; b < 4; b++) qlo[b] = q255f((float)((w0 >> (8 * b)) & 0xffu));
#pragma unroll
        for (int l = 0; l < NL; l++) {
            const unsigned w1 = (unsigned)(u[l + 1] >> (32 * hh));
            float qhi[4];
#pragma unroll
            for (int b = 0; b < 4; b++) qhi[b] = q255f((float)((w1 >> (8 * b)) & 0xffu));
            *reinterpret_cast<float4*>(dst + l * (kTileH + 2) * kExtRW) =
                make_float4(qhi[0] - qlo[0], qhi[1] - qlo[1], qhi[2] - qlo[2], qhi[3] - qlo[3]);
#pragma unroll
            for (int b = 0; b < 4; b++) qlo[b] = qhi[b];
        }
    }
}

// Occupancy target: the LDS bound (6 work-groups per CU for S <= 3).
template <int S>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(S <= 3 ? 6 : 3)))
void k_extrema_tile(const OctaveDesc* __restrict__ oct, int noct,
                    const ImageParams* __restrict__ params, unsigned long long* __restrict__ bits,
                    unsigned* __restrict__ counts, int strip_rows) {
    constexpr int NL = S + 2;  // DoG layers
    constexpr int NG = S + 3;  // Gaussian levels
    __shared__ __align__(16) float D[NL][kTileH + 2][kExtRW];
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int j = find_by(oct, noct, tile, &OctaveDesc::row_begin);
    const OctaveDesc& O = oct[j];
    const ImageParams& P = params[O.img];
    const int t = tile - O.row_begin;
    const int w = O.w, h = O.h, border = P.border;
    const int tiles_x = (w + kTileW - 1) / kTileW;
    const int x0 = (t % tiles_x) * kTileW, ys = (t / tiles_x) * strip_rows;
    const int nbands = min(strip_rows, h - ys + kTileH - 1) / kTileH;  // >= 1

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x = x0 + lane;
    const bool col_ok = x >= border && x < w - border;
    unsigned long long u[NG];
    ext_fetch<NG>(O, x0, ys, u);
#pragma unroll 1
    for (int band = 0; band < nbands; band++) {
        const int y0 = ys + band * kTileH;
        if (band > 0) __syncthreads();  // the previous band's tests are done with D
        ext_stage<NL, NG>(u, D);
        if (band + 1 < nbands) ext_fetch<NG>(O, x0, y0 + kTileH, u);
        __syncthreads();

        float hmx[3][NL], hmn[3][NL], dl[3][NL], dc[3][NL], dr[3][NL];
        auto load_row = [&](int slot, int r) {
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const float a = D[l][r][lane + 3];  // column x0 + lane - 1
                const float b = D[l][r][lane + 4];
                const float c = D[l][r][lane + 5];
                dl[slot][l] = a;
                dc[slot][l] = b;
                dr[slot][l] = c;
                hmx[slot][l] = fmaxf(a, fmaxf(b, c));
                hmn[slot][l] = fminf(a, fminf(b, c));
            }
        };
        const int rbase = wave * 4;  // local D row of output row y0 + wave*4 is rbase + 1
        load_row(0, rbase);
        load_row(1, rbase + 1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int prv = k % 3, cur = (k + 1) % 3, nxt = (k + 2) % 3;
            load_row(nxt, rbase + k + 2);
            const int y = y0 + wave * 4 + k;
            const bool row_ok = y >= border && y < h - border;
#pragma unroll
            for (int s = 1; s <= S; s++) {
                float mx = fmaxf(fmaxf(hmx[prv][s - 1], hmx[cur][s - 1]), hmx[nxt][s - 1]);
                mx = fmaxf(mx, fmaxf(fmaxf(hmx[prv][s + 1], hmx[cur][s + 1]), hmx[nxt][s + 1]));
                mx = fmaxf(mx, fmaxf(fmaxf(hmx[prv][s], hmx[nxt][s]), fmaxf(dl[cur][s], dr[cur][s])));
                float mn = fminf(fminf(hmn[prv][s - 1], hmn[cur][s - 1]), hmn[nxt][s - 1]);
                mn = fminf(mn, fminf(fminf(hmn[prv][s + 1], hmn[cur][s + 1]), hmn[nxt][s + 1]));
                mn = fminf(mn, fminf(fminf(hmn[prv][s], hmn[nxt][s]), fminf(dl[cur][s], dr[cur][s])));
                const float v = dc[cur][s];
                const bool ext = row_ok && col_ok && (v > mx || v < mn);
                const unsigned long long m = ballot64(ext);
                if (lane == 0 && y < h) {
                    const long long slot_local = (long long)(s - 1) * h + y;
                    bits[O.bits_begin + slot_local * O.nwords + (x0 >> 6)] = m;
                    if (m) atomicAdd(&counts[O.slot_begin + slot_local], (unsigned)__popcll(m));
                }
            }
        }
    }
}

__global__ void k_fill_u32(unsigned* __restrict__ p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Exclusive scan over n uint32 (n read from device memory when n_dev != null,
// capped by n_max): three passes over 4096-element tiles.
// out has n+1 entries; out[n] = total.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned* sh, unsigned& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    unsigned wpre = 0, tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
        unsigned t = sh[i];
        if (i < wave) wpre += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wpre + x - v;
}

__device__ __forceinline__ int scan_n(int n_static, const unsigned* n_dev) {
    if (!n_dev) return n_static;
    unsigned v = *n_dev;
    return (int)min(v, (unsigned)n_static);
}

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const unsigned* __restrict__ in, int n_max,
                                                          const unsigned* n_dev,
                                                          unsigned* __restrict__ tile_sums) {
    __shared__ unsigned sh[8];
    const int n = scan_n(n_max, n_dev);
    const int base = blockIdx.x * kScanTile;
    unsigned s = 0;
    for (int i = 0; i < kScanTile / kThreads; i++) {
        int k = base + i * kThreads + threadIdx.x;
        if (k < n) s += in[k];
    }
    unsigned tot;
    block_excl_scan(s, sh, tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void k_scan_tiles(unsigned* __restrict__ tile_sums, int ntiles) {
    __shared__ unsigned sh[8];
    unsigned carry = 0;
    for (int b = 0; b < ntiles; b += kThreads) {
        int k = b + threadIdx.x;
        unsigned v = k < ntiles ? tile_sums[k] : 0u;
        unsigned tot;
        unsigned ex = block_excl_scan(v, sh, tot);
        if (k < ntiles) tile_sums[k] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_down(const unsigned* __restrict__ in, int n_max,
                                                        const unsigned* n_dev,
                                                        const unsigned* __restrict__ tile_prefix,
                                                        unsigned* __restrict__ out) {
    __shared__ unsigned sh[8];
    const int n = scan_n(n_max, n_dev);
    const int base = blockIdx.x * kScanTile + threadIdx.x * 16;
    unsigned v[16];
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    unsigned tot;
    unsigned ex = block_excl_scan(s, sh, tot);
    unsigned run = tile_prefix[blockIdx.x] + ex;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
        if (base + i == n - 1) out[n] = run;
    }
    if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

// The same scan in one work-group of 1024 threads, for small n_max (single
// frames, small batches): 16384-element chunks with a running carry, one
// launch instead of three.
constexpr int kScanSingleThreads = 1024;
constexpr int kScanSingleMax = 1 << 16;
__global__ __launch_bounds__(kScanSingleThreads) void k_scan_single(const unsigned* __restrict__ in,
                                                                   int n_max, const unsigned* n_dev,
                                                                   unsigned* __restrict__ out) {
    __shared__ unsigned sh[kScanSingleThreads / 64];
    const int n = scan_n(n_max, n_dev);
    unsigned carry = 0;
    for (int b = 0; b < n; b += kScanSingleThreads * 16) {
        const int base = b + threadIdx.x * 16;
        unsigned v[16];
        unsigned s = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            v[i] = (base + i < n) ? in[base + i] : 0u;
            s += v[i];
        }
        unsigned tot;
        unsigned run = carry + block_excl_scan(s, sh, tot);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if (base + i < n) out[base + i] = run;
            run += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// ---------------------------------------------------------------------------
// k_compact: one wave per spw consecutive (octave, layer, row) slots (spw <=
// kCompactSlots: 8 for launches with many slots, 1 for single frames, whose
// few slots would otherwise leave most of the GPU idle).
// The slots' counts and offsets are read once (one lane each) and the octave is
// found once; the bitmask words of every slot with candidates (one lane per
// word, rows of at most 64 words) are loaded before any slot is written, so
// their loads overlap. For each slot the set bits are ranked with popcounts and
// a wave prefix sum and the candidates written at offsets[slot] + rank, i.e. in
// the reference's order. Rows of more than 64 words take a chunked loop.
// ---------------------------------------------------------------------------
constexpr int kCompactSlots = 8;

__device__ __forceinline__ void compact_word(unsigned long long m, int wi, int j, int y, int s,
                                             unsigned& base, int lane, Cand* __restrict__ cand,
                                             int cap) {
    const unsigned c = (unsigned)__popcll(m);
    unsigned x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    unsigned pre = base + x - c;
    const unsigned tot = __shfl(x, 63, 64);
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        if (pre < (unsigned)cap) {
            Cand cd;
            cd.job = j;
            cd.x = wi * 64 + b;
            cd.y = y;
            cd.layer = s;
            cand[pre] = cd;
        }
        pre++;
    }
    base += tot;
}

__global__ __launch_bounds__(kThreads) void k_compact(const OctaveDesc* __restrict__ oct, int noct,
                                                      int nslots,
                                                      const unsigned long long* __restrict__ bits,
                                                      const unsigned* __restrict__ counts,
                                                      const unsigned* __restrict__ offsets,
                                                      Cand* __restrict__ cand, int cap, int spw) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s0 = (blockIdx.x * (kThreads / 64) + wave) * spw;
    if (s0 >= nslots) return;
    const int ns = min(spw, nslots - s0);
    const unsigned cnt_l = lane < ns ? counts[s0 + lane] : 0u;
    const unsigned off_l = lane < ns ? offsets[s0 + lane] : 0u;
    const unsigned long long any = ballot64(cnt_l != 0u);
    if (any == 0ull) return;
    // each slot's octave (the slots are ordered by octave)
    int jk[kCompactSlots];
    int j = find_by(oct, noct, s0, &OctaveDesc::slot_begin);
#pragma unroll
    for (int k = 0; k < kCompactSlots; k++) {
        while (k < ns && j + 1 < noct && oct[j + 1].slot_begin <= s0 + k) j++;
        jk[k] = j;
    }
    // the bitmask words of the slots with candidates, all loads issued first
    unsigned long long m[kCompactSlots];
#pragma unroll
    for (int k = 0; k < kCompactSlots; k++) {
        m[k] = 0ull;
        if ((any >> k) & 1ull) {
            const OctaveDesc& O = oct[jk[k]];
            if (lane < O.nwords)
                m[k] = bits[O.bits_begin + (long long)(s0 + k - O.slot_begin) * O.nwords + lane];
        }
    }
#pragma unroll
    for (int k = 0; k < kCompactSlots; k++) {
        if (!((any >> k) & 1ull)) continue;
        const OctaveDesc& O = oct[jk[k]];
        const int local = s0 + k - O.slot_begin;
        const int s = local / O.h + 1, y = local - (s - 1) * O.h;
        unsigned base = __shfl(off_l, k, 64);
        compact_word(m[k], lane, jk[k], y, s, base, lane, cand, cap);
        if (O.nwords > 64) {  // rows wider than 4096 px: the remaining words by chunks
            const unsigned long long* wrow = bits + O.bits_begin + (long long)local * O.nwords;
            for (int w0 = 64; w0 < O.nwords; w0 += 64) {
                const int wi = w0 + lane;
                compact_word(wi < O.nwords ? wrow[wi] : 0ull, wi, jk[k], y, s, base, lane, cand, cap);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_refine: one thread per candidate; the reference's arithmetic in the same
// f32 operation order (compiled without FMA contraction).
// ---------------------------------------------------------------------------
struct DogView {
    const uint8_t* levels;
    long long stride;
    int pitch, w, h;
    __device__ __forceinline__ float at(int l, int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        const uint8_t* g0 = levels + (long long)l * stride + (size_t)y * pitch + x;
        return q255f((float)g0[stride]) - q255f((float)g0[0]);
    }
};

__global__ __launch_bounds__(kThreads) void k_refine(const OctaveDesc* __restrict__ oct,
                                                     const ImageParams* __restrict__ params,
                                                     const Cand* __restrict__ cand,
                                                     const unsigned* __restrict__ total, int cap,
                                                     Keypoint* __restrict__ out,
                                                     int* __restrict__ valid,
                                                     float* __restrict__ cand_resp) {
    const int n = (int)min(*total, (unsigned)cap);
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
        const Cand cd = cand[c];
        const OctaveDesc& O = oct[cd.job];
        const ImageParams& P = params[O.img];
        DogView V{O.levels, O.level_stride, O.pitch, O.w, O.h};
        const int nd = P.num_intervals + 2;  // DoG layers per octave
        const float scale0 = pow2f(O.octave);
        // kp.x / 2^o of the candidate ((x + 0.5) * 2^o), `as i32` (src/sift.rs:309-313)
        const float xi = (((float)cd.x + 0.5f) * scale0) / scale0;
        const float yi = (((float)cd.y + 0.5f) * scale0) / scale0;
        int cx = f2i_sat(xi), cy = f2i_sat(yi), cl = cd.layer;
        if (cand_resp) cand_resp[c] = V.at(cd.layer, cd.x, cd.y);
        int ok = 0;
        Keypoint r;
        for (int step = 0; step < 5; step++) {
            if (cl < 1 || cl >= nd - 1) break;
            if (cx < 1 || cx >= O.w - 1 || cy < 1 || cy >= O.h - 1) break;
            const float pxp = V.at(cl, cx + 1, cy), pxm = V.at(cl, cx - 1, cy);
            const float pyp = V.at(cl, cx, cy + 1), pym = V.at(cl, cx, cy - 1);
            const float nxt = V.at(cl + 1, cx, cy), prv = V.at(cl - 1, cx, cy);
            const float dx = (pxp - pxm) / 2.0f;
            const float dy = (pyp - pym) / 2.0f;
            const float ds = (nxt - prv) / 2.0f;
            const float cv = V.at(cl, cx, cy);
            const float dxx = pxp + pxm - 2.0f * cv;
            const float dyy = pyp + pym - 2.0f * cv;
            const float dss = nxt + prv - 2.0f * cv;
            const float dxy = (V.at(cl, cx + 1, cy + 1) - V.at(cl, cx - 1, cy + 1) -
                               V.at(cl, cx + 1, cy - 1) + V.at(cl, cx - 1, cy - 1)) / 4.0f;
            const float dxs = (V.at(cl + 1, cx + 1, cy) - V.at(cl + 1, cx - 1, cy) -
                               (V.at(cl - 1, cx + 1, cy) - V.at(cl - 1, cx - 1, cy))) / 4.0f;
            const float dys = (V.at(cl + 1, cx, cy + 1) - V.at(cl + 1, cx, cy - 1) -
                               (V.at(cl - 1, cx, cy + 1) - V.at(cl - 1, cx, cy - 1))) / 4.0f;
            // solve_linear_system (src/sift.rs:478-503), b = -g
            const float a00 = dxx, a01 = dxy, a02 = dxs, a10 = dxy, a11 = dyy, a12 = dys,
                        a20 = dxs, a21 = dys, a22 = dss;
            const float b0 = -dx, b1 = -dy, b2 = -ds;
            const float det_a = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) +
                                a02 * (a10 * a21 - a11 * a20);
            if (fabsf(det_a) < 1e-10f) break;
            const float det_x = b0 * (a11 * a22 - a12 * a21) - a01 * (b1 * a22 - a12 * b2) +
                                a02 * (b1 * a21 - a11 * b2);
            const float det_y = a00 * (b1 * a22 - a12 * b2) - b0 * (a10 * a22 - a12 * a20) +
                                a02 * (a10 * b2 - b1 * a20);
            const float det_z = a00 * (a11 * b2 - b1 * a21) - a01 * (a10 * b2 - b1 * a20) +
                                b0 * (a10 * a21 - a11 * a20);
            const float ox = det_x / det_a, oy = det_y / det_a, os = det_z / det_a;
            if (fabsf(ox) < 0.5f && fabsf(oy) < 0.5f && fabsf(os) < 0.5f) {
                const float val = cv + 0.5f * (dx * ox + dy * oy + ds * os);
                if (fabsf(val) < P.contrast_threshold) break;
                const float tr = dxx + dyy;
                const float trace_sq = tr * tr;
                const float det = dxx * dyy - dxy * dxy;
                if (det <= 0.0f) break;
                const float ratio = trace_sq / det;
                const float e1 = P.edge_threshold + 1.0f;
                const float thr = e1 * e1 / P.edge_threshold;
                if (ratio >= thr) break;
                const float scale = pow2f(O.octave);
                const float fl = (float)cl + os;
                const float e = (float)O.octave + fl / (float)P.num_intervals;
                // 2.0_f32.powf(e): correctly rounded via double exp2
                const float psig = P.sigma * (float)exp2((double)e);
                r.x = ((float)cx + ox) * scale;
                r.y = ((float)cy + oy) * scale;
                r.size = psig;
                r.angle = 0.0f;
                r.response = val;
                r.octave = O.octave;
                r.layer = f2i_sat(roundf(fl));
                ok = 1;
                break;
            } else {
                cx = f2i_sat(roundf((float)cx + ox));
                cy = f2i_sat(roundf((float)cy + oy));
                cl = f2i_sat(roundf((float)cl + os));
                if (cl < 0 || cl >= nd) break;
            }
        }
        valid[c] = ok;
        if (ok) out[c] = r;
    }
}

// ---------------------------------------------------------------------------
// k_orient: one wave (workgroup of 64) per refined keypoint, grid-stride
// (src/sift.rs:605-726). For R <= 16 the window's rows and columns of the
// Gaussian level (every pixel the gradient of a sample can touch) are staged
// in LDS as bytes; a sample reads q = v/255 of its 4 gradient texels from an
// LDS table filled with the exact f32 q255f. Lane l takes one contiguous chunk of the
// (2R+1)^2 window in raster order; each sample adds mag * exp(-(dx^2+dy^2)/denom)
// to bin floor(theta * 36 / 2pi) of a u64 fixed-point (2^40) histogram, one of 4
// copies by lane, with runs of same-bin samples summed in registers first
// (integer adds: order-independent, deterministic, and as fine as the f32
// contributions themselves). Lanes 0..35 then hold one bin
// each for the two [1,1,1]/3 smoothing passes, the max and the peak test; a
// ballot ranks the peaks in bin order.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Work queue of the keypoint kernels. With a queue, a launch has as many
// work-groups as fit on the GPU at once, and each takes chunks of kChunk
// consecutive keypoints until the count is reached. Work-group b uses counter
// b % kNumXcd (the XCD the dispatcher places it on) of kNumXcd counters, zeroed
// by a fill kernel earlier on the stream; counter x's chunk c holds keypoints
// [(c * kNumXcd + x) * kChunk, + kChunk). One counter taking every item
// serialises the atomics of the whole GPU on one address (measured: 10x the
// kernel's time); eight counters and chunks make them rare. Without a queue,
// work-group b takes keypoints b, b + grid, ...
// Thread 0 takes the next chunk at the start of a chunk's last item, into the
// slot the work-group is not reading, and everyone reads it after the barrier
// that ends that item: each slot is written before a barrier, read after it,
// and not written again until every thread has read it. All threads of a
// work-group thus see the same item sequence, every item's early exits
// (`continue`) still pass the barrier that ends it, and the loop test is
// uniform.
// ---------------------------------------------------------------------------
template <int kChunk>
struct WorkQueue {
    unsigned* heads;  // kNumXcd counters; null: static grid-stride assignment
    int slot_it, left, x;

    __device__ __forceinline__ int chunk_first(int c) const { return (c * kNumXcd + x) * kChunk; }
    __device__ __forceinline__ int first(int* slots) {
        slot_it = 0;
        left = kChunk - 1;
        x = (int)(blockIdx.x % kNumXcd);
        if (!heads) return (int)blockIdx.x;
        if (threadIdx.x == 0) slots[0] = (int)atomicAdd(&heads[x], 1u);
        __syncthreads();
        return chunk_first(slots[0]);
    }
    // at the start of an item: thread 0 takes the next chunk during the last
    // item of this one
    __device__ __forceinline__ void prefetch(int* slots) {
        if (heads && left == 0 && threadIdx.x == 0)
            slots[(slot_it + 1) & 1] = (int)atomicAdd(&heads[x], 1u);
    }
    __device__ __forceinline__ int next(int* slots, int cur) {
        if (!heads) return cur + (int)gridDim.x;
        __syncthreads();
        if (left > 0) {
            left--;
            return cur + 1;
        }
        slot_it++;
        left = kChunk - 1;
        return chunk_first(slots[slot_it & 1]);
    }
};
constexpr int kOrientChunk = 8;
constexpr int kDescChunk = 2;

constexpr int kOriMaxR = 16;
constexpr int kOriRows = 2 * kOriMaxR + 5;  // staged rows
constexpr int kOS = 2 * kOriMaxR + 8;       // staged row stride (bytes): 4 * ceil((3 + rows + 3) / 4)

__global__ __launch_bounds__(64) void k_orient(const OctaveDesc* __restrict__ oct,
                                               const ImageParams* __restrict__ params,
                                               const Cand* __restrict__ cand,
                                               const unsigned* __restrict__ total, int cap,
                                               const Keypoint* __restrict__ ref,
                                               const int* __restrict__ valid,
                                               unsigned* __restrict__ npeaks,
                                               float* __restrict__ angles,
                                               unsigned* __restrict__ queue) {
    // 4 histogram copies (lane & 3), the staged window and the q table: 4.2 KB of
    // LDS, so the 8 waves per SIMD that its 44 VGPRs allow fit (32 work-groups of one
    // wave per CU; 21 with the window staged as f32)
    __shared__ unsigned long long hist[4][36];
    // v / 255 for every byte (q255f, exact): 4 LDS reads per sample instead of 12 VALU
    // operations; written before the first keypoint's barrier, which orders it
    __shared__ float qt[256];
    for (int i = threadIdx.x; i < 256; i += 64) qt[i] = q255f((float)i);
    __shared__ __align__(16) uint8_t pq[kOriRows * kOS];
    // per column / row of the window: (pixel offset as int bits, Gaussian factor)
    __shared__ float2 ocx[2 * kOriMaxR + 1], ocy[2 * kOriMaxR + 1];
    const int lane = threadIdx.x;
    const int n = (int)min(*total, (unsigned)cap);
    __shared__ int qslot[2];
    WorkQueue<kOrientChunk> Q{queue, 0, 0, 0};
    for (int c = Q.first(qslot); c < n; c = Q.next(qslot, c)) {
        Q.prefetch(qslot);
        if (!valid[c]) {
            if (lane == 0) npeaks[c] = 0;
            continue;
        }
        const Keypoint kp = ref[c];
        const OctaveDesc& O = oct[cand[c].job];
        const ImageParams& P = params[O.img];
        const int S = P.num_intervals;
        const int gl = min(max(kp.layer, 0), S + 2);
        const uint8_t* g = O.levels + (long long)gl * O.level_stride;
        const int w = O.w, h = O.h, pitch = O.pitch;
        const float scale = pow2f(kp.octave);
        const float xo = kp.x / scale, yo = kp.y / scale, so = kp.size / scale;
        if (!(so > 0.0f)) {
            if (lane == 0) npeaks[c] = 0;
            continue;
        }
        const int R = f2i_sat(roundf(4.5f * so));
        const float ws = 1.5f * so;
        const float denom = 2.0f * ws * ws;
        const int side = 2 * R + 1, nsamp = side * side;
        const int iters = (nsamp + 63) >> 6;
        // weight * 2^40 = 2^(kq * (dx^2 + dy^2) + 40)
        const float kq = -1.44269504088896341f / denom;
        constexpr float lf = 40.0f;  // m = mag * weight * 2^40
        const bool lds = R <= kOriMaxR;
        for (int i = lane; i < 4 * 36; i += 64) (&hist[0][0])[i] = 0ull;
        // the samples' pixels (round(x_o + dx), round(y_o + dy)) and their 4-neighbours
        // lie in columns px0 .. px0 + 2R + 4 and rows py0 .. py0 + 2R + 4
        const int px0 = f2i_sat(roundf(xo)) - R - 2, py0 = f2i_sat(roundf(yo)) - R - 2;
        const int px0a = px0 & ~3;
        if (lds) {
            const int pside = 2 * R + 5;
            const int nw = (px0 - px0a + pside + 3) >> 2;  // <= kOS / 4
            const float inv_nw = 1.0f / (float)nw;
            for (int i = lane; i < nw * pside; i += 64) {
                const int v = (int)(((float)i + 0.5f) * inv_nw), wd = i - v * nw;
                const int yy = py0 + v, xw = px0a + 4 * wd;
                unsigned u = 0;
                if (yy >= 0 && yy < h) {
                    const uint8_t* row = g + (size_t)yy * pitch;
                    if (xw >= 0 && xw + 4 <= pitch) {
                        u = *reinterpret_cast<const unsigned*>(row + xw);
                    } else {
#pragma unroll
                        for (int b = 0; b < 4; b++)
                            if (xw + b >= 0 && xw + b < w) u |= (unsigned)row[xw + b] << (8 * b);
                    }
                }
                *reinterpret_cast<unsigned*>(&pq[v * kOS + 4 * wd]) = u;
            }
            // per column and row of the window: the sample pixel's offset in the
            // staged window, or -1 when it is outside [1, w-2] x [1, h-2], and the
            // Gaussian weight's factor 2^(kq d^2 + 20) of that column / row offset
            // d (the weight of sample (dx, dy) is the product of its column's and
            // its row's factors: one multiply per sample instead of exp2)
            if (lane < side) {
                const float d = (float)(lane - R);
                const float wf = __builtin_amdgcn_exp2f(__builtin_fmaf(kq, d * d, 0.5f * lf));
                const int xi = f2i_sat(roundf(xo + d));
                ocx[lane] = make_float2(__int_as_float((xi >= 1 && xi < w - 1) ? xi - px0a : -1), wf);
                const int yi = f2i_sat(roundf(yo + d));
                ocy[lane] = make_float2(__int_as_float((yi >= 1 && yi < h - 1) ? (yi - py0) * kOS : -1), wf);
            }
        }
        __syncthreads();
        // lane l takes the samples [l C, (l+1) C) in raster order, C = ceil(n / 64):
        // lanes sit far apart in the window, and a lane's consecutive samples are
        // neighbours that often share a bin, so their fixed-point values are
        // summed in registers (integer sums: the histogram is the same as with
        // one add per sample) and added to LDS once per run
        const int i0 = min(lane * iters, nsamp), i1 = min(i0 + iters, nsamp);
        const int iy0 = (int)(((float)i0 + 0.5f) * (1.0f / (float)side));
        int dy = iy0 - R, dx = i0 - iy0 * side - R;
        unsigned long long run = 0;
        int run_bin = -1;
        unsigned long long* const hrow = hist[lane & 3];
        for (int i = i0; i < i1; i++) {
            float gx = 0.0f, gy = 0.0f, wgt = 0.0f;
            bool in = true;
            if (lds) {
                const float2 cx = ocx[dx + R], cy = ocy[dy + R];
                const int ox = __float_as_int(cx.x), oy = __float_as_int(cy.x);
                wgt = cx.y * cy.y;
                in = (ox | oy) >= 0;
                if (in) {
                    const uint8_t* pp = &pq[oy + ox];
                    gx = qt[pp[1]] - qt[pp[-1]];
                    gy = qt[pp[kOS]] - qt[pp[-kOS]];
                }
            } else {
                const int yi = f2i_sat(roundf(yo + (float)dy));
                const int xi = f2i_sat(roundf(xo + (float)dx));
                in = yi >= 1 && yi < h - 1 && xi >= 1 && xi < w - 1;
                if (in) {
                    const uint8_t* p = g + (size_t)yi * pitch + xi;
                    gx = q255f((float)p[1]) - q255f((float)p[-1]);
                    gy = q255f((float)p[pitch]) - q255f((float)p[-pitch]);
                    const float fdx = (float)dx, fdy = (float)dy;
                    wgt = __builtin_amdgcn_exp2f(__builtin_fmaf(kq, __builtin_fmaf(fdx, fdx, fdy * fdy), lf));
                }
            }
            if (in) {
                const float m = __builtin_amdgcn_sqrtf(__builtin_fmaf(gx, gx, gy * gy)) * wgt;
                const float ang = atan2_grad(gy, gx);
                const float an = ang < 0.0f ? ang + kTwoPi : ang;
                // (an * 36) / 2pi with the reference's two roundings: gradients along the
                // axes land exactly on bin edges, where the rounding decides the bin
                int bin = (int)floorf(div_two_pi(an * 36.0f));
                bin = bin >= 36 ? bin - 36 : bin;
                // m * 2^40 as a u64: the high word is exact (power-of-two scaling and
                // floor), the remainder x - hi * 2^32 is exact (Sterbenz), both truncate
                const float hi = floorf(m * 2.3283064365386963e-10f);
                const float lo = __builtin_fmaf(-hi, 4294967296.0f, m);
                const unsigned long long v = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
                if (bin != run_bin) {
                    if (run_bin >= 0) atomicAdd(&hrow[run_bin], run);
                    run_bin = bin;
                    run = 0;
                }
                run += v;
            }
            if (++dx > R) {
                dx = -R;
                dy++;
            }
        }
        if (run_bin >= 0) atomicAdd(&hrow[run_bin], run);
        __syncthreads();
        float hv = 0.0f;
        if (lane < 36) {
            unsigned long long sum = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) sum += hist[k][lane];
            hv = (float)((double)sum * (1.0 / 1099511627776.0));
        }
        for (int it = 0; it < 2; it++) {
            const float pv = __shfl(hv, lane < 36 ? (lane + 35) % 36 : lane, 64);
            const float nv = __shfl(hv, lane < 36 ? (lane + 1) % 36 : lane, 64);
            if (lane < 36) hv = (pv + hv + nv) / 3.0f;
        }
        float mx = lane < 36 ? hv : 0.0f;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d, 64));
        mx = fmaxf(0.0f, mx);
        const float thr = mx * 0.8f;
        const float pv = __shfl(hv, lane < 36 ? (lane + 35) % 36 : lane, 64);
        const float nv = __shfl(hv, lane < 36 ? (lane + 1) % 36 : lane, 64);
        const bool peak = lane < 36 && hv >= thr && hv > pv && hv > nv;
        const unsigned long long m = ballot64(peak);
        if (peak) {
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            const float den = pv - 2.0f * hv + nv;
            const float off = fabsf(den) > 1e-5f ? 0.5f * (pv - nv) / den : 0.0f;
            const float step = kTwoPi / 36.0f;
            const float center = ((float)lane + 0.5f) * step;
            const float a = center + off * step;
            float fa = rem_euclid(a, kTwoPi);
            if (fa > kPi) fa = fa - kTwoPi;
            if (rank < kMaxPeaks) angles[(size_t)c * kMaxPeaks + rank] = fa;
        }
        if (lane == 0) npeaks[c] = (unsigned)min(__popcll(m), kMaxPeaks);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Descriptor helpers: clamped u8 sample, get_gauss_pixel_bilinear.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gauss_px(const uint8_t* g, int pitch, int w, int h, int x, int y,
                                          const float* q) {
    x = min(max(x, 0), w - 1);
    y = min(max(y, 0), h - 1);
    return q[g[(size_t)y * pitch + x]];
}

// get_gauss_pixel_bilinear (src/sift.rs:238-266)
__device__ __forceinline__ float bilinear(const uint8_t* g, int pitch, int w, int h, float x,
                                          float y, const float* q) {
    const float xf = floorf(x), yf = floorf(y);
    const float xc = xf + 1.0f, yc = yf + 1.0f;
    const float dx = x - xf, dy = y - yf;
    const int x0 = f2i_sat(xf), y0 = f2i_sat(yf), x1 = f2i_sat(xc), y1 = f2i_sat(yc);
    const float q11 = gauss_px(g, pitch, w, h, x0, y0, q);
    const float q21 = gauss_px(g, pitch, w, h, x1, y0, q);
    const float q12 = gauss_px(g, pitch, w, h, x0, y1, q);
    const float q22 = gauss_px(g, pitch, w, h, x1, y1, q);
    return q11 * (1.0f - dx) * (1.0f - dy) + q21 * dx * (1.0f - dy) + q12 * (1.0f - dx) * dy +
           q22 * dx * dy;
}

// Inclusive prefix sum over the 64 lanes of a fully active wave with DPP:
// shifts by 1, 2, 4, 8 inside each row of 16 lanes (lanes without a source add
// 0), then lane 15 of rows 0 and 2 is added to rows 1 and 3, and lane 31 to
// rows 2 and 3.
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// Sum over the 64 lanes of a fully active wave (DPP scan, then lane 63),
// returned to every lane; the order of the additions is fixed.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    int x = __float_as_int(v);
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false)));
    return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// Maximum over the 64 lanes of a fully active wave (DPP, as wave_incl_scan),
// returned to every lane.
__device__ __forceinline__ unsigned wave_max_u32(unsigned x) {
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true));  // row_shr:1
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true));  // row_shr:2
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true));  // row_shr:4
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true));  // row_shr:8
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false)); // row_bcast:15
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false)); // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

// Maximum over the 64 lanes of a fully active wave, returned to every lane.
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

// Per-keypoint gain of the descriptor's fixed-point scale: the largest power of
// two 2^e (0 <= e <= 40) with M * 2^e <= lim, where M bounds |gx| and |gy| of
// every sample of the window and lim is their largest possible value (255 in
// texel units, 1 in [0, 1] units). Scaled by it, a sample's contribution
// mag * weight * fix is still at most sqrt(2) * fix, so the u32 sum bound holds,
// while a low-contrast window's sums use the high bits of the words instead of
// a few units of 2^-20 (whose rounding, relative to the window's small norm,
// reached 7e-4 of L2 at 3840x2160, contrast 0.03).
// desc_fix_gain returns 2^e, desc_fix_exp e.
__device__ __forceinline__ int desc_fix_exp(float M, float lim) {
    if (!(M > 0.0f)) return 0;
    int x;
    (void)frexpf(M, &x);                  // M in [2^(x-1), 2^x)
    int e = (lim > 1.0f ? 8 : 0) - x;     // M * 2^e in [128, 256), or [0.5, 1)
    if (ldexpf(M, e) > lim) e--;
    return min(max(e, 0), 40);
}
__device__ __forceinline__ float desc_fix_gain(float M, float lim) {
    return ldexpf(1.0f, desc_fix_exp(M, lim));
}

// k_descriptor's fixed-point gain exponent for one keypoint: that of its
// patch's texel range (max - min over the waves' (max << 8) | (255 - min) words),
// the same value in every lane.
__device__ __forceinline__ int desc_range_exp(const unsigned* wrange) {
    unsigned hi = 0u, lo = 0u;
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) {
        hi = max(hi, wrange[w] >> 8);
        lo = max(lo, wrange[w] & 255u);
    }
    const int range = (int)hi - (int)(255u - lo);
    return __builtin_amdgcn_readfirstlane(desc_fix_exp((float)range, 255.0f));
}

constexpr int kDescMaxR = 31;    // LDS path for sigma_octave <= 3.65
// The instantiation for windows up to R = 28 (99.5% of the keypoints of the
// default schedule, 20.1 KB of LDS, 64 VGPRs) fits 8 work-groups per CU, the R = 31 one 6.
constexpr int kDescSmallR = 28;
// Candidate capacities up to this (single frames up to 2048 x 1536, small
// batches) take every descriptor in one launch of the R <= kDescMaxR
// instantiation.
constexpr int kDescOneLaunchCap = 131072;
// LDS sizes of the descriptor instantiation for windows up to MR.
template <int MR>
struct DescDims {
    static constexpr int GS = 2 * MR + 3;         // bilinear grid row stride (floats)
    static constexpr int PatchRows = 2 * MR + 6;  // staged patch rows
    // staged patch row stride (bytes): the aligned words of a row, up to 3 + rows + 3 bytes
    static constexpr int PS = (2 * MR + 12) / 4 * 4;
};
// One histogram cell: the even-pair table (bins 0..7 in words 0..7) and the
// odd-pair table (bin b in word 9 + b, bin 0 again in word 17); 18 words keep
// every cell, and with it every bin pair, 8-byte aligned.
constexpr int kCellWords = 18;
constexpr int kOddBase = 9;                     // word of bin 0 in the odd-pair table
constexpr int kHistWords = 36 * kCellWords;     // one copy: 6 x 6 padded cells
constexpr int kHistCopies = 2;                  // even and odd lanes scatter into separate copies
constexpr int kMaxCopySamples = 2048;           // per-copy sample bound that keeps u32 sums exact

// Trilinear scatter of one sample (src/sift.rs:1012-1049), contribution
// ((m * wx) * wy) * wa in u32 fixed point, m = weighted magnitude * fix.
// The histogram is padded to 6 x 6 spatial cells (cells -1..4 in x and y): the
// reference drops neighbours outside the 4 x 4 grid, here they land in border
// cells that are never read, so there is no clamping and the four spatial
// neighbours are constant offsets. The two angle bins (a, a+1 mod 8) of each
// spatial neighbour go in ONE 64-bit LDS add: a cell's even-pair table holds
// bins 0..7 in its words 0..7 (pairs (0,1), (2,3), ... are 8-byte aligned); its
// odd-pair table holds bin b in word 9 + b and bin 0 again in word 17, so pairs
// (1,2), (3,4), (5,6), (7,0) are aligned too. Every bin sum of a copy stays
// below 2^32 (at most kMaxCopySamples samples of at most sqrt(2) * fix + 1/2
// each), so the low word never carries into the high word.
//
// desc_bins: the bin pair address of one sample (its spatial neighbour 0), its
// four spatial weights t_k and its angle fraction da; ab, the angle in bins, may
// be anywhere in (-8, 8): its integer part is taken mod 8.
__device__ __forceinline__ unsigned long long* desc_bins(unsigned* hcopy, float xb, float yb,
                                                         float ab, float m, float (&tt)[4],
                                                         float& da) {
    const float xf = floorf(xb), yf = floorf(yb), af = floorf(ab);
    const unsigned ai = ((unsigned)(int)af) & 7u;
    const float dx = xb - xf, dy = yb - yf;
    da = ab - af;
    // cell (xf, yf) of the padded 6 x 6 grid: 7 * kCellWords + the offset below
    const unsigned cw =
        (unsigned)(int)__builtin_fmaf(yf, 6.0f * kCellWords, xf * (float)kCellWords);
    const float t0 = m * (1.0f - dx), t1 = m * dx;
    tt[0] = t0 * (1.0f - dy);
    tt[1] = t1 * (1.0f - dy);
    tt[2] = t0 * dy;
    tt[3] = t1 * dy;
    return reinterpret_cast<unsigned long long*>(hcopy + 7 * kCellWords +
                                                 (cw + ai + __umul24(ai & 1u, (unsigned)kOddBase)));
}

// Adds four pairs of fixed-point sums (a[2k], a[2k+1]) at the spatial
// neighbours of a bin pair address, each rounded as (unsigned)(a + bias):
// bias 0.5 for plain sums, 0 for sums that started at 0.5.
template <int kHalfBiased = 0>
__device__ __forceinline__ void desc_add_pairs(unsigned long long* base, const float (&a)[8]) {
    constexpr int off[4] = {0, kCellWords / 2, 6 * kCellWords / 2, 7 * kCellWords / 2};
    constexpr float bias = kHalfBiased ? 0.0f : 0.5f;
#pragma unroll
    for (int k = 0; k < 4; k++)
        atomicAdd(base + off[k], ((unsigned long long)(unsigned)(a[2 * k + 1] + bias) << 32) |
                                     (unsigned)(a[2 * k] + bias));
}

__device__ __forceinline__ void desc_scatter(unsigned* hcopy, float xb, float yb, float ab,
                                             float m) {
    float tt[4], da;
    unsigned long long* const base = desc_bins(hcopy, xb, yb, ab, m, tt, da);
    const float a[8] = {tt[0] * (1.0f - da), tt[0] * da, tt[1] * (1.0f - da), tt[1] * da,
                        tt[2] * (1.0f - da), tt[2] * da, tt[3] * (1.0f - da), tt[3] * da};
    desc_add_pairs(base, a);
}

// Interval of px where c * px + d lies in (-B, B) in real arithmetic; the
// caller widens it by one sample and the per-sample window test decides.
// A coefficient too small to move c * px by 0.1% of B over the window keeps the
// whole row or none of it.
__device__ __forceinline__ void desc_row_interval(float c, float d, float B, int R, float& lo,
                                                  float& hi) {
    if (fabsf(c) * (float)(R + 1) <= 1e-3f * B) {
        const bool keep = fabsf(d) < B * 1.002f;
        lo = keep ? -1e6f : 1e6f;
        hi = keep ? 1e6f : -1e6f;
        return;
    }
    const float rc = __builtin_amdgcn_rcpf(c);  // endpoints need far less than one sample of accuracy
    const float e1 = (-B - d) * rc, e2 = (B - d) * rc;
    lo = fminf(e1, e2);
    hi = fmaxf(e1, e2);
}

// One patch row of the descriptor's bilinear grid: texels (a, a+1) of the u8
// patch at byte offset a, interpolated along x in texel units (the x half of
// get_gauss_pixel_bilinear, src/sift.rs:238-266, as one lerp). The two texels
// come from aligned words: a byte-aligned 16-bit LDS read would be an
// unaligned access.
__device__ __forceinline__ float desc_grid_row(const unsigned* pu, int a, float dx) {
    const unsigned* wp = pu + ((unsigned)a >> 2);
    const unsigned r = __builtin_amdgcn_alignbyte(wp[1], wp[0], (unsigned)a & 3u);
    const float p1 = (float)(r & 255u), p2 = (float)((r >> 8) & 255u);
    return __builtin_fmaf(p2 - p1, dx, p1);
}

// Per-keypoint geometry of compute (src/sift.rs:940-960).
struct DescGeom {
    Keypoint base;
    const uint8_t* g;
    int w, h, pitch, R, side, np;
    unsigned k0;
    float xo, yo, so, bw, denom;
};

__device__ __forceinline__ DescGeom desc_geom(int c, const OctaveDesc* __restrict__ oct,
                                              const ImageParams* __restrict__ params,
                                              const Cand* __restrict__ cand,
                                              const Keypoint* __restrict__ ref,
                                              const unsigned* __restrict__ npeaks,
                                              const unsigned* __restrict__ koff) {
    DescGeom G;
    G.np = (int)npeaks[c];
    G.base = ref[c];
    G.k0 = koff[c];
    const OctaveDesc& O = oct[cand[c].job];
    const int S = params[O.img].num_intervals;
    const int gl = min(max(G.base.layer, 0), S + 2);
    G.g = O.levels + (long long)gl * O.level_stride;
    G.w = O.w;
    G.h = O.h;
    G.pitch = O.pitch;
    const float scale = pow2f(G.base.octave);
    G.xo = G.base.x / scale;
    G.yo = G.base.y / scale;
    G.so = G.base.size / scale;
    G.bw = 3.0f * G.so;
    const float win = G.bw * 4.0f;
    const float wsig = 0.5f * win;
    G.denom = 2.0f * wsig * wsig;
    G.R = G.so > 0.0f ? f2i_sat(ceilf(win * sqrtf(2.0f) * 0.5f)) : 0;
    G.side = 2 * G.R + 1;
    return G;
}

// (ix, iy) range whose x_o + (ix - R) lies in [0, w-1) and y_o + (iy - R) in
// [0, h-1) (src/sift.rs:981-987), evaluated with the reference's f32 test at
// the candidate integers.
__device__ __forceinline__ void desc_bounds(const DescGeom& G, int& bx0, int& bx1, int& by0,
                                            int& by1) {
    const float wm1 = (float)(G.w - 1), hm1 = (float)(G.h - 1);
    bx0 = 0;
    bx1 = G.side - 1;
    by0 = 0;
    by1 = G.side - 1;
    while (bx0 <= bx1 && G.xo + (float)(bx0 - G.R) < 0.0f) bx0++;
    while (bx1 >= bx0 && G.xo + (float)(bx1 - G.R) >= wm1) bx1--;
    while (by0 <= by1 && G.yo + (float)(by0 - G.R) < 0.0f) by0++;
    while (by1 >= by0 && G.yo + (float)(by1 - G.R) >= hm1) by1--;
}

// Wave 0: sums the histogram copies of one peak, normalises, clips at 0.2,
// renormalises (src/sift.rs:889-913) and writes the descriptor and keypoint.
// The sums are divided by the fixed-point scale `fix`, or (kDivide false)
// multiplied by its exact reciprocal inv_fix when fix is a power of two.
template <bool kDivide>
__device__ __forceinline__ void desc_finish(const unsigned* hist, float fix, double inv_fix, int l,
                                            const Keypoint& base, float angle, unsigned kidx,
                                            int kcap, Keypoint* __restrict__ kps_out,
                                            float* __restrict__ desc_out) {
    float v[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int dd = 2 * l + e;  // (iy * 4 + ix) * 8 + ia
        const int a = dd & 7;
        const int cell = ((dd >> 5) + 1) * 6 + (((dd >> 3) & 3) + 1);
        unsigned long long sum = 0;
#pragma unroll
        for (int cp = 0; cp < kHistCopies; cp++) {
            const unsigned* t = hist + cp * kHistWords + cell * kCellWords;
            sum += t[a];                // even-pair table
            sum += t[kOddBase + a];     // odd-pair table
            if (a == 0) sum += t[kOddBase + 8];
        }
        v[e] = kDivide ? (float)((double)sum / (double)fix) : (float)((double)sum * inv_fix);
    }
    float v0 = v[0], v1 = v[1];
    const float nrm = sqrtf(wave_sum_dpp(v0 * v0 + v1 * v1));
    if (nrm >= 1e-8f) {
        const float inv = 1.0f / nrm;
        v0 = fminf(v0 * inv, 0.2f);
        v1 = fminf(v1 * inv, 0.2f);
        const float n2 = sqrtf(wave_sum_dpp(v0 * v0 + v1 * v1));
        if (n2 >= 1e-8f) {
            const float inv2 = 1.0f / n2;
            v0 *= inv2;
            v1 *= inv2;
        }
    }
    if (kidx < (unsigned)kcap) {
        float2* dd = (float2*)(desc_out + (size_t)kidx * 128);
        dd[l] = make_float2(v0, v1);
        if (l == 0) {
            Keypoint o = base;
            o.angle = angle;
            kps_out[kidx] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// k_descriptor<MR>: one workgroup (256) per refined keypoint with R <= MR
// (sigma_octave <= 3.65), looping over its orientation peaks
// (src/sift.rs:917-1060).
// Per keypoint: the clamped u8 patch of the Gaussian level around (x_o, y_o) --
// every texel the reference's get_gauss_pixel_bilinear can touch in the
// window, with its clamp-to-edge rule -- is staged in LDS (aligned 32-bit
// copies when the patch is inside the level, clamped byte loads at the
// borders), and the bilinear samples on the shifted grid (x_o + i, y_o + j),
// |i|,|j| <= R+1, are formed once from it, in texel units (0..255): gradients
// are differences of grid values, so their angle is the same and their
// magnitude is scaled by 1/255 with the Gaussian weight. A window sample (px, py) takes its
// gradient from the four grid neighbours (the reference's bilinear(x_s +- 1, y_s)
// and bilinear(x_s, y_s +- 1)).
// Per peak, the samples of each row whose rotated bin coordinates can fall in
// (-1, 4) form one interval (the rotated window is convex); the intervals,
// widened by one sample and clipped to the in-image rectangle, are cut into at
// most 256 segments of at most `per` consecutive samples of one row, one per
// thread, so nearly every lane works on in-window samples of a single row and
// the lanes of a wave sit far apart in the window (different histogram cells). Each sample still applies the window
// test xb, yb in (-1, 4), then computes mag * exp(-(px^2+py^2)/(2 (6 sigma_o)^2))
// and atan2 and scatters trilinearly into its copy of the histogram (even and
// odd lanes use separate copies, which halves same-address collisions of the
// LDS adds; integer adds: independent of the order). The fixed-point scale is
// 2^20 times the largest power of two that keeps the patch's texel range times
// it within 255 (desc_range_exp). Wave 0 sums the copies and normalises.
// ---------------------------------------------------------------------------
template <int MR, bool kFromList>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(MR <= kDescSmallR ? 8 : 6))) void k_descriptor(
    const OctaveDesc* __restrict__ oct, const ImageParams* __restrict__ params,
    const Cand* __restrict__ cand, const unsigned* __restrict__ total, int cap,
    const Keypoint* __restrict__ ref, const unsigned* __restrict__ npeaks,
    const float* __restrict__ angles, const unsigned* __restrict__ koff, int kcap,
    Keypoint* __restrict__ kps_out, float* __restrict__ desc_out,
    const unsigned* __restrict__ in_count, const int* __restrict__ in_list,
    unsigned* __restrict__ out_count, int* __restrict__ out_list, unsigned* __restrict__ queue) {
    using D = DescDims<MR>;
    constexpr int kGS = D::GS, kPatchRows = D::PatchRows, kPS = D::PS;
    static_assert(kHistCopies * kHistWords * 4 >= kPatchRows * kPS + 4, "patch shares the histograms");
    // The u8 patch (and, when they fit beside it, the grid's column and row
    // terms) are dead once the grid is built; the histograms live after that.
    constexpr int kPatchBytes = (kPatchRows * kPS + 15) / 16 * 16;
    constexpr bool kTermsInPatch = kPatchBytes + 2 * kGS * (int)sizeof(float2) <= kHistCopies * kHistWords * 4;
    __shared__ __align__(16) unsigned pu[kHistCopies * kHistWords];
    __shared__ float grid[kGS * kGS];
    __shared__ float exs[2 * MR + 1];
    __shared__ int rspan[64];  // per row: first and last column of its samples (16 bits each)
    __shared__ int rpre[65];
    __shared__ int seg_len;
    // per wave: (max << 8) | (255 - min) of the patch texels it staged
    __shared__ unsigned wrange[kThreads / 64];
    __shared__ float2 terms_own[kTermsInPatch ? 1 : 2 * kGS];
    __shared__ float pang[kMaxPeaks], pcos[kMaxPeaks], psin[kMaxPeaks];
    uint8_t* const patch = reinterpret_cast<uint8_t*>(pu);
    unsigned* const hist = pu;
    // grid columns / rows: texel offset, fraction
    float2* const gcol = kTermsInPatch ? reinterpret_cast<float2*>(patch + kPatchBytes) : terms_own;
    float2* const grw = gcol + kGS;
    // all candidates, or the ones an instantiation for smaller windows listed
    const int n = (int)min(kFromList ? *in_count : *total, (unsigned)cap);
    __shared__ int qslot[2];
    WorkQueue<kDescChunk> Q{queue, 0, 0, 0};
    for (int idx = Q.first(qslot); idx < n; idx = Q.next(qslot, idx)) {
        Q.prefetch(qslot);
        // the thread index passes through an empty asm in every iteration, so
        // the many cheap values derived from it are formed where they are used
        // instead of being hoisted out of this loop and kept live across it
        // (spilled to scratch at 7 waves per SIMD)
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int wave = tid >> 6, lane = tid & 63;
        const int c = kFromList ? in_list[idx] : idx;
        const DescGeom G = desc_geom(c, oct, params, cand, ref, npeaks, koff);
        if (G.np == 0) continue;
        if (G.R > MR) {  // wider windows: listed for the next instantiation
            if (threadIdx.x == 0) out_list[atomicAdd(out_count, 1u)] = c;
            continue;
        }
        const int R = G.R, side = G.side, gside = 2 * R + 3;
        const float xo = G.xo, yo = G.yo, bw = G.bw, denom = G.denom;
        if (tid < G.np) {  // the peaks' angles and their cos / sin, once per keypoint
            const float a = angles[(size_t)c * kMaxPeaks + tid];
            pang[tid] = a;
            pcos[tid] = cosf(a);
            psin[tid] = sinf(a);
        }
        {
            const int pside = 2 * R + 6;
            const int px0 = f2i_sat(floorf(xo)) - R - 2, py0 = f2i_sat(floorf(yo)) - R - 2;
            const int px0a = px0 & ~3, poff = px0 - px0a;
            unsigned tmax = 0u, tmin = 255u;
            if (px0 >= 0 && py0 >= 0 && px0 + pside <= G.w && py0 + pside <= G.h) {
                // interior: aligned words [px0a, round_up(px0 + pside, 4)) of each row,
                // which end within the row's pitch
                const int nw = (poff + pside + 3) >> 2;  // <= kPS / 4
                const int tot = nw * pside;              // <= 18 * 68 < 5 * 256
                const float inv_nw = 1.0f / (float)nw;
                const uint8_t* rowp = G.g + (size_t)py0 * G.pitch + px0a;
                unsigned wv[5];
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const int i = tid + k * kThreads;
                    wv[k] = 0u;  // defined on every path: nothing carried across keypoints
                    if (i < tot) {
                        const int v = (int)(((float)i + 0.5f) * inv_nw);
                        const int wd = i - v * nw;
                        wv[k] = *reinterpret_cast<const unsigned*>(rowp + (size_t)v * G.pitch + 4 * wd);
                    }
                }
                // texel range of the staged words (a few texels left and right of
                // the patch included: a wider range is still a bound)
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    if (tid + k * kThreads < tot) {
                        const unsigned b0 = wv[k] & 255u, b1 = (wv[k] >> 8) & 255u;
                        const unsigned b2 = (wv[k] >> 16) & 255u, b3 = wv[k] >> 24;
                        tmax = max(max(tmax, max(b0, b1)), max(b2, b3));
                        tmin = min(min(tmin, min(b0, b1)), min(b2, b3));
                    }
                }
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const int i = tid + k * kThreads;
                    if (i < tot) {
                        const int v = (int)(((float)i + 0.5f) * inv_nw);
                        pu[v * (kPS / 4) + (i - v * nw)] = wv[k];
                    }
                }
            } else {
                const float inv_ps = 1.0f / (float)pside;
                for (int i = tid; i < pside * pside; i += kThreads) {
                    const int v = (int)(((float)i + 0.5f) * inv_ps), u = i - v * pside;
                    const int xx = min(max(px0 + u, 0), G.w - 1), yy = min(max(py0 + v, 0), G.h - 1);
                    const unsigned t = G.g[(size_t)yy * G.pitch + xx];
                    patch[v * kPS + poff + u] = (uint8_t)t;
                    tmax = max(tmax, t);
                    tmin = min(tmin, t);
                }
            }
            // Every grid value is a convex combination of patch texels, so every
            // sample's |gx| and |gy| is at most the patch's texel range: it sets
            // the keypoint's fixed-point gain (read after the barrier below).
            tmax = wave_max_u32(tmax);
            tmin = 255u - wave_max_u32(255u - tmin);
            if (lane == 0) wrange[wave] = (tmax << 8) | (255u - tmin);
            // the grid's column and row terms: x = x_o + i (and y = y_o + j) are the
            // same f32 sums for every point of a column (row), so their texel offset
            // and fraction are formed once
            if (tid < gside) {
                const float x = xo + (float)(tid - R - 1);
                const float xf = floorf(x);
                gcol[tid] = make_float2(__int_as_float(min(max(f2i_sat(xf) - px0, 0), pside - 2)), x - xf);
            } else if (tid >= 128 && tid < 128 + gside) {
                const int gy = tid - 128;
                const float y = yo + (float)(gy - R - 1);
                const float yf = floorf(y);
                const int v0 = min(max(f2i_sat(yf) - py0, 0), pside - 2);
                grw[gy] = make_float2(__int_as_float(v0 * kPS + poff), y - yf);
            }
            __syncthreads();
            {
                // Thread (band, gx) walks grid column gx down one band of rows.
                // Grid row gy interpolates patch rows v0 and v0 + 1 along x
                // ("top", "bot") and then along y; the next grid row's v0 is
                // v0 + 1 except where the clamp to the patch (or the f32 sum
                // y_o + j) makes it otherwise, so its "top" is this row's "bot"
                // and each patch row is interpolated once per band.
                const float inv_gs = 1.0f / (float)gside;
                const int nb = kThreads / gside;  // bands: >= 3 (gside <= 65)
                const int band = (int)(((float)tid + 0.5f) * inv_gs), gx = tid - band * gside;
                if (band < nb) {
                    const int rows = (gside + nb - 1) / nb;
                    const int gy1 = min((band + 1) * rows, gside);
                    const float2 cxd = gcol[gx];
                    const int cxo = __float_as_int(cxd.x);
                    const float dx = cxd.y;
                    float top = 0.0f;
                    // patch offset of the previous row's v0; offsets are >= 0, so
                    // the first row of the band never matches it
                    int prev = -(1 << 30);
                    for (int gy = band * rows; gy < gy1; gy++) {
                        const float2 ryd = grw[gy];
                        const int ro = __float_as_int(ryd.x);
                        if (ro != prev + kPS) top = desc_grid_row(pu, ro + cxo, dx);
                        const float bot = desc_grid_row(pu, ro + kPS + cxo, dx);
                        grid[gy * kGS + gx] = __builtin_fmaf(bot - top, ryd.y, top);
                        top = bot;
                        prev = ro;
                    }
                }
            }
            // Gaussian weight exp(-(px^2+py^2)/denom) as exp(-px^2/denom) * exp(-py^2/denom)
            for (int i = tid; i < side; i += kThreads) exs[i] = expf(-(float)((i - R) * (i - R)) / denom);
        }
        __syncthreads();  // patch reads done before the histograms reuse its storage
        int bx0, bx1, by0, by1;
        desc_bounds(G, bx0, bx1, by0, by1);
        const int nrows = G.so > 0.0f ? max(by1 - by0 + 1, 0) : 0;  // <= 63
        for (int pk = 0; pk < G.np; pk++) {
            const float angle = uniform_f(pang[pk]);
            for (int i = tid; i < kHistCopies * kHistWords / 4; i += kThreads)
                reinterpret_cast<uint4*>(hist)[i] = make_uint4(0u, 0u, 0u, 0u);
            const float ca = uniform_f(pcos[pk]), sa = uniform_f(psin[pk]);
            const float inv_bw = uniform_f(1.0f / bw);
            const float msa = -sa;  // -sin * px is computed as (-sin) * px, the same value
            const float abin = uniform_f(angle * (4.0f / kPi));  // the peak angle in bins
            if (tid < 64) {
                // per-row sample intervals of the rotated window, prefix-summed
                int lo = 0, cnt = 0;
                if (tid < nrows) {
                    const int cy = by0 + tid;
                    const float py = (float)(cy - R);
                    const float B = 2.5f * bw;
                    float l1, h1, l2, h2;
                    desc_row_interval(ca, sa * py, B, R, l1, h1);
                    desc_row_interval(msa, ca * py, B, R, l2, h2);
                    const float lof = fmaxf(fmaxf(l1, l2), -1e5f), hif = fminf(fminf(h1, h2), 1e5f);
                    const int a = max((int)ceilf(lof) + R - 1, bx0);
                    const int b = min((int)floorf(hif) + R + 1, bx1);
                    lo = a;
                    cnt = max(b - a + 1, 0);
                    rspan[tid] = lo | ((lo + cnt - 1) << 16);  // columns < 64, last >= lo - 1
                }
                // Each row is cut into segments of at most `per` consecutive samples,
                // one segment per thread: per is the smallest value >= T / 256 that
                // gives at most 256 segments (per <= 16: a row holds at most 63
                // samples, so per = 16 gives at most 4 * 63 segments).
                const int T = __builtin_amdgcn_readlane(wave_incl_scan(cnt), 63);
                int per = max((T + kThreads - 1) / kThreads, 1);
                int segs = wave_incl_scan((cnt + per - 1) / per);
                while (__builtin_amdgcn_readlane(segs, 63) > kThreads) {
                    per++;
                    segs = wave_incl_scan((cnt + per - 1) / per);
                }
                rpre[tid + 1] = segs;  // segment prefix over the rows
                // the fixed-point scale 2^(20 + fe), passed with the segment length
                const int fe = desc_range_exp(wrange);
                if (tid == 0) {
                    rpre[0] = 0;
                    seg_len = per | (fe << 16);
                }
            }
            __syncthreads();
            const int per = seg_len & 0xffff, fe = seg_len >> 16;
            const int sg = lane * 4 + wave;  // lanes of a wave take segments far apart
            if (sg < rpre[nrows]) {
                unsigned* hw = hist + (lane & 1) * kHistWords;  // neighbouring lanes: different copies
                int r0 = 0, r1 = nrows;  // rpre[r0] <= sg < rpre[r1]: row r0 holds the segment
                while (r1 - r0 > 1) {
                    const int mid = (r0 + r1) >> 1;
                    if (rpre[mid] <= sg) r0 = mid;
                    else r1 = mid;
                }
                const int span = rspan[r0];
                int cx = (span & 0xffff) + (sg - rpre[r0]) * per;
                const int len = min(per, (span >> 16) - cx + 1);
                // the row's terms (src/sift.rs:967-968 at px = 0, and the row's
                // Gaussian factor), formed as the row-interval pass forms them
                const int cy = by0 + r0;
                const float py = (float)(cy - R);
                const float rz = sa * py * inv_bw, rw = ca * py * inv_bw;
                const float wyf = exs[cy] * ldexpf(kDescFix / 255.0f, fe);
                // grid point (cx, cy) of sample (cx, cy) -- the up-left neighbour -- so that
                // all four gradient neighbours sit at non-negative offsets
                const float* grow = grid + (by0 + r0) * kGS;
                // rotated bin coordinates (rx / bw, ry / bw) (src/sift.rs:967-968), advanced
                // by one sample step along the row
                const float cai = ca * inv_bw, msai = msa * inv_bw;
                // consecutive samples of a thread often share their bin pair: their
                // contributions are summed in registers, starting from 0.5 so that the
                // rounding to fixed point is a plain conversion, and added to LDS once
                // per run. The sums are defined from the start (each run resets them
                // anyway): left undefined, they were carried through the keypoint
                // loop and spilled to scratch once per keypoint.
                unsigned long long* acc_base = nullptr;
                float acc[8] = {0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f};
                float u = __builtin_fmaf(cai, (float)(cx - R), rz);
                float v = __builtin_fmaf(msai, (float)(cx - R), rw);
                for (int i = 0; i < len; i++, cx++, u += cai, v += msai) {
                    // window: bin coordinates u + 1.5, v + 1.5 in (-1, 4) (src/sift.rs:969-977)
                    if (!(fabsf(u) < 2.5f && fabsf(v) < 2.5f)) continue;
                    const float* gq = grow + cx;
                    const float gx = gq[kGS + 2] - gq[kGS];
                    const float gyv = gq[2 * kGS + 1] - gq[1];
                    const float mag = __builtin_amdgcn_sqrtf(__builtin_fmaf(gx, gx, gyv * gyv));
                    const float m = mag * (exs[cx] * wyf);
                    // relative angle rem_euclid(atan2 - angle, 2pi) in bins (src/sift.rs:1009):
                    // ab in (-8, 8); the bin index wraps in desc_bins (& 7 of the
                    // integer floor) and the fraction ab - floor(ab) needs no wrap
                    const float ab = atan2_bins(gyv, gx) - abin;
                    float tt[4], da;
                    unsigned long long* const base = desc_bins(hw, u + 1.5f, v + 1.5f, ab, m, tt, da);
                    if (base != acc_base) {  // a new bin pair: add the finished sums
                        if (acc_base) desc_add_pairs<1>(acc_base, acc);
                        acc_base = base;
#pragma unroll
                        for (int q8 = 0; q8 < 8; q8++) acc[q8] = 0.5f;
                    }
                    const float wa0 = 1.0f - da;
#pragma unroll
                    for (int q4 = 0; q4 < 4; q4++) {
                        acc[2 * q4] = __builtin_fmaf(tt[q4], wa0, acc[2 * q4]);
                        acc[2 * q4 + 1] = __builtin_fmaf(tt[q4], da, acc[2 * q4 + 1]);
                    }
                }
                if (acc_base) desc_add_pairs<1>(acc_base, acc);
            }
            __syncthreads();
            if (tid < 64) {
                // the keypoint and its output slot are read again here rather than
                // kept in registers through the peak loop (the index goes through an
                // empty asm so the compiler does not reuse the first reads)
                int cr = c;
                asm volatile("" : "+v"(cr));
                desc_finish<false>(hist, 0.0f, ldexp(1.0 / (double)kDescFix, -(seg_len >> 16)), tid,
                                   ref[cr], angle,
                                   koff[cr] + (unsigned)pk, kcap, kps_out,
                            desc_out);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// k_descriptor_wide: keypoints whose window exceeds the LDS patch
// (R > 31, sigma_octave > 3.65; none with the default sigma and 3 intervals).
// Direct per-sample path with global reads and the reference's f32 arithmetic;
// 256 contiguous chunks of the (2R+1)^2 window as in k_descriptor. The fixed-point
// scale drops below 2^20 when a copy would see more than kMaxCopySamples samples,
// and is raised by the gain of the largest |gx|, |gy| of the window's samples.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_descriptor_wide(
    const OctaveDesc* __restrict__ oct, const ImageParams* __restrict__ params,
    const Cand* __restrict__ cand, const unsigned* __restrict__ total, int cap,
    const Keypoint* __restrict__ ref, const unsigned* __restrict__ npeaks,
    const float* __restrict__ angles, const unsigned* __restrict__ koff, int kcap,
    Keypoint* __restrict__ kps_out, float* __restrict__ desc_out,
    const unsigned* __restrict__ wide_count, const int* __restrict__ wide_list) {
    __shared__ float q[256];
    __shared__ __align__(16) unsigned hist[kHistCopies * kHistWords];
    __shared__ float wave_gmax[kThreads / 64];
    fill_q(q);
    __syncthreads();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    (void)total;
    (void)cap;
    const int nw = (int)*wide_count;  // keypoints k_descriptor listed
    for (int i = blockIdx.x; i < nw; i += gridDim.x) {
        const int c = wide_list[i];
        const DescGeom G = desc_geom(c, oct, params, cand, ref, npeaks, koff);
        const int R = G.R, side = G.side, nsamp = side * side;
        const float wm1 = (float)(G.w - 1), hm1 = (float)(G.h - 1);
        const int per = (nsamp + kThreads - 1) / kThreads;
        const int i0 = (lane * 4 + wave) * per;
        const int iend = min(i0 + per, nsamp);
        {
            // the largest |gx|, |gy| of the in-image samples sets the keypoint's
            // fixed-point gain (written per wave, read after the first barrier of
            // the peak loop)
            float gm = 0.0f;
            for (int i = i0; i < iend && G.np > 0; i++) {
                const int cy = i / side, cx = i - cy * side;
                const float xs = G.xo + (float)(cx - R), ys = G.yo + (float)(cy - R);
                if (xs < 0.0f || xs >= wm1 || ys < 0.0f || ys >= hm1) continue;
                const float gx = bilinear(G.g, G.pitch, G.w, G.h, xs + 1.0f, ys, q) -
                                 bilinear(G.g, G.pitch, G.w, G.h, xs - 1.0f, ys, q);
                const float gy = bilinear(G.g, G.pitch, G.w, G.h, xs, ys + 1.0f, q) -
                                 bilinear(G.g, G.pitch, G.w, G.h, xs, ys - 1.0f, q);
                gm = fmaxf(gm, fmaxf(fabsf(gx), fabsf(gy)));
            }
            gm = wave_max(gm);
            if (lane == 0) wave_gmax[wave] = gm;
        }
        for (int pk = 0; pk < G.np; pk++) {
            const float angle = angles[(size_t)c * kMaxPeaks + pk];
            for (int i = tid; i < kHistCopies * kHistWords / 4; i += kThreads)
                reinterpret_cast<uint4*>(hist)[i] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
            const float ca = cosf(angle), sa = sinf(angle);
            unsigned* hw = hist + (lane & 1) * kHistWords;
            const float gmax = fmaxf(fmaxf(wave_gmax[0], wave_gmax[1]), fmaxf(wave_gmax[2], wave_gmax[3]));
            const float fix = kDescFix * fminf(1.0f, (float)kMaxCopySamples / (float)(128 * per)) *
                              desc_fix_gain(gmax, 1.0f);
            for (int i = i0; i < iend; i++) {
                const int cy = i / side, cx = i - cy * side;
                const float py = (float)(cy - R), px = (float)(cx - R);
                const float rx = ca * px + sa * py;
                const float ry = -sa * px + ca * py;
                const float xb = rx / G.bw + 4.0f / 2.0f - 0.5f;
                const float yb = ry / G.bw + 4.0f / 2.0f - 0.5f;
                if (!(xb > -1.0f && xb < 4.0f && yb > -1.0f && yb < 4.0f)) continue;
                const float xs = G.xo + px, ys = G.yo + py;
                if (xs < 0.0f || xs >= wm1 || ys < 0.0f || ys >= hm1) continue;
                const float gx = bilinear(G.g, G.pitch, G.w, G.h, xs + 1.0f, ys, q) -
                                 bilinear(G.g, G.pitch, G.w, G.h, xs - 1.0f, ys, q);
                const float gy = bilinear(G.g, G.pitch, G.w, G.h, xs, ys + 1.0f, q) -
                                 bilinear(G.g, G.pitch, G.w, G.h, xs, ys - 1.0f, q);
                const float wt = expf(-(px * px + py * py) / G.denom);
                const float mag = sqrtf(gx * gx + gy * gy);
                const float pa = atan2f(gy, gx);
                const float ar = rem_euclid(pa - angle, kTwoPi);
                desc_scatter(hw, xb, yb, ar * 8.0f / kTwoPi, mag * wt * fix);
            }
            __syncthreads();
            if (tid < 64)
                desc_finish<true>(hist, fix, 0.0, tid, G.base, angle, G.k0 + (unsigned)pk, kcap, kps_out,
                            desc_out);
            __syncthreads();
        }
    }
}

// Per-image keypoint ranges: summary[2+i] = koff[first candidate of image i]
// (and out_offsets[i], when given); summary[0] = total candidates (may exceed
// cap), summary[1] = total keypoints.
__global__ void k_finalize(const unsigned* __restrict__ slot_offsets,
                           const int* __restrict__ img_slot_begin, int nimg,
                           const unsigned* __restrict__ koff, const unsigned* __restrict__ total,
                           int cap, unsigned* __restrict__ summary, unsigned* __restrict__ out_offsets) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nimg) return;
    const unsigned n = min(*total, (unsigned)cap);
    const unsigned cb = min(slot_offsets[img_slot_begin[i]], n);
    summary[2 + i] = koff[cb];
    if (out_offsets) out_offsets[i] = koff[cb];
    if (i == 0) {
        summary[0] = *total;
        summary[1] = koff[n];
    }
}

// u8 descriptors of the WebGPU backends (src/shaders/descriptor.wgsl:256-262):
// trunc(clamp(d * 512, 0, 255)) of each value, 4 values per lane.
__device__ __forceinline__ unsigned desc_u8_of(float d) {
    return (unsigned)fminf(fmaxf(d * 512.0f, 0.0f), 255.0f);
}

__global__ __launch_bounds__(256) void k_desc_u8(const float4* __restrict__ desc,
                                                 const unsigned* __restrict__ count, int kcap,
                                                 unsigned* __restrict__ out) {
    const size_t words = (size_t)min(*count, (unsigned)kcap) * 32;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = desc[i];
        out[i] = desc_u8_of(v.x) | (desc_u8_of(v.y) << 8) | (desc_u8_of(v.z) << 16) |
                 (desc_u8_of(v.w) << 24);
    }
}

}  // namespace

void launch_extrema(const OctaveDesc* d_oct, int noct, int total_rows, const ImageParams* d_params,
                    unsigned long long* d_bits, unsigned* d_counts, int nslots, hipStream_t s) {
    // the counters after the row-slot counts (the descriptor's lists, the work queues)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, d_counts + nslots, kCountExtra, 0u);
    if (total_rows <= 0) return;
    hipLaunchKernelGGL(k_extrema, dim3(total_rows), dim3(kThreads), 0, s, d_oct, noct, d_params,
                       d_bits, d_counts);
}

bool extrema_tile_supported(int S) { return S >= 1 && S <= 6; }

void launch_extrema_tile(int S, int strip_rows, const OctaveDesc* d_oct, int noct, int total_tiles,
                         const ImageParams* d_params, unsigned long long* d_bits, unsigned* d_counts,
                         int nslots, hipStream_t s) {
    // the row-slot counts and the counters after them (the descriptor's lists, the work queues)
    hipLaunchKernelGGL(k_fill_u32, dim3((nslots + kCountExtra + 255) / 256), dim3(256), 0, s, d_counts,
                       nslots + kCountExtra, 0u);
    if (total_tiles <= 0) return;
    switch (S) {
#define SH_EXT(SS) \
        case SS: hipLaunchKernelGGL(k_extrema_tile<SS>, dim3(total_tiles), dim3(kThreads), 0, s, d_oct, \
                                    noct, d_params, d_bits, d_counts, strip_rows); break;
        SH_EXT(1) SH_EXT(2) SH_EXT(3) SH_EXT(4) SH_EXT(5) SH_EXT(6)
#undef SH_EXT
        default: break;
    }
}

// n_dev may be null (then n_max is the count). out needs n_max+1 entries;
// tmp needs ceil(n_max/4096)+1 entries.
void launch_scan_dev(const unsigned* d_in, int n_max, const unsigned* n_dev, unsigned* d_out,
                     unsigned* d_tmp, hipStream_t s) {
    if (n_max <= kScanSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanSingleThreads), 0, s, d_in, n_max, n_dev,
                           d_out);
        return;
    }
    const int tiles = n_max > 0 ? (n_max + kScanTile - 1) / kScanTile : 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kThreads), 0, s, d_in, n_max, n_dev, d_tmp);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, s, d_tmp, tiles);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kThreads), 0, s, d_in, n_max, n_dev, d_tmp,
                       d_out);
}

void launch_compact(const OctaveDesc* d_oct, int noct, int nslots, const unsigned long long* d_bits,
                    const unsigned* d_counts, const unsigned* d_offsets, Cand* d_cand, int cap,
                    hipStream_t s) {
    if (nslots <= 0) return;
    const int spw = nslots >= 65536 ? kCompactSlots : 1;  // slots per wave
    const int per = (kThreads / 64) * spw;                 // slots per work-group
    hipLaunchKernelGGL(k_compact, dim3((nslots + per - 1) / per), dim3(kThreads), 0, s, d_oct, noct,
                       nslots, d_bits, d_counts, d_offsets, d_cand, cap, spw);
}

static int grid_for(int cap, int per_block, int max_blocks) {
    int g = (cap + per_block - 1) / per_block;
    return g < 1 ? 1 : (g > max_blocks ? max_blocks : g);
}

// Work-groups of `kernel` that fit on the current device at once (its
// occupancy times the CU count), computed once per kernel.
static int resident_groups(const void* kernel, int block) {
    static std::mutex mu;
    static std::map<const void*, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu <= 0 || cus <= 0) {
        per_cu = 4;
        cus = 256;
    }
    const int g = per_cu * cus;
    cache[kernel] = g;
    return g;
}

// The keypoint kernels take their keypoints by a static grid-stride assignment;
// SIFT_HIP_QUEUE=1 (read on every launch) gives them the device work queue
// instead (the same results; on 64 x 1920x1080 frames it was 45% slower for
// k_orient and even for k_descriptor).
static bool queue_enabled() {
    const char* e = getenv("SIFT_HIP_QUEUE");
    return e && e[0] == '1';
}

void launch_refine(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, Keypoint* d_ref, int* d_valid,
                   float* d_cand_resp, hipStream_t s) {
    hipLaunchKernelGGL(k_refine, dim3(grid_for(cap, kThreads, 2048)), dim3(kThreads), 0, s, d_oct,
                       d_params, d_cand, d_total, cap, d_ref, d_valid, d_cand_resp);
}

void launch_orient(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, const Keypoint* d_ref, const int* d_valid,
                   unsigned* d_npeaks, float* d_angles, unsigned* d_queue, hipStream_t s) {
    // work queue: as many work-groups as fit, each taking keypoints until none
    // are left; or grid-stride over the candidates with many more work-groups
    // than fit at once (each takes only a few keypoints, so the launch ends evenly)
    const bool q = queue_enabled();
    const int grid = q ? std::min(resident_groups((const void*)k_orient, 64), grid_for(cap, 1, 65536))
                       : grid_for(cap, 1, 65536);
    // (d_queue: kNumXcd counters)
    hipLaunchKernelGGL(k_orient, dim3(grid), dim3(64), 0, s, d_oct, d_params, d_cand, d_total, cap,
                       d_ref, d_valid, d_npeaks, d_angles, q ? d_queue : nullptr);
}

void launch_descriptor(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                       const unsigned* d_total, int cap, const Keypoint* d_ref,
                       const unsigned* d_npeaks, const float* d_angles, const unsigned* d_koff,
                       int kcap, Keypoint* d_kps, float* d_desc, unsigned* d_wide_count,
                       int* d_wide_list, unsigned* d_mid_count, int* d_mid_list, unsigned* d_queues,
                       hipStream_t s) {
    // as k_orient: a work queue, or few keypoints per work-group. Windows up to
    // kDescSmallR run at 8 work-groups per CU and list the few others for the
    // kDescMaxR instantiation (6 per CU), which lists still wider ones for
    // k_descriptor_wide. d_queues: the two instantiations' queue counters
    // (kNumXcd each).
    const bool q = queue_enabled();
    if (cap <= kDescOneLaunchCap) {
        // few candidates (small frames): one launch of the wide instantiation
        // over all of them; the GPU is far from full, so its lower occupancy
        // costs less than the second launch's latency
        const void* kall = (const void*)k_descriptor<kDescMaxR, false>;
        const int g = q ? std::min(resident_groups(kall, kThreads), grid_for(cap, 1, 49152))
                        : grid_for(cap, 1, 49152);
        hipLaunchKernelGGL((k_descriptor<kDescMaxR, false>), dim3(g), dim3(kThreads), 0, s, d_oct,
                           d_params, d_cand, d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap,
                           d_kps, d_desc, nullptr, nullptr, d_wide_count, d_wide_list,
                           q ? d_queues + kNumXcd : nullptr);
        hipLaunchKernelGGL(k_descriptor_wide, dim3(256), dim3(kThreads), 0, s, d_oct, d_params, d_cand,
                           d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap, d_kps, d_desc,
                           d_wide_count, d_wide_list);
        return;
    }
    const void* k28 = (const void*)k_descriptor<kDescSmallR, false>;
    const void* k31 = (const void*)k_descriptor<kDescMaxR, true>;
    const int g28 = q ? std::min(resident_groups(k28, kThreads), grid_for(cap, 1, 49152)) : grid_for(cap, 1, 49152);
    const int g31 = q ? std::min(resident_groups(k31, kThreads), grid_for(cap, 1, 49152)) : grid_for(cap, 1, 49152);
    hipLaunchKernelGGL((k_descriptor<kDescSmallR, false>), dim3(g28), dim3(kThreads), 0, s, d_oct,
                       d_params, d_cand, d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap,
                       d_kps, d_desc, nullptr, nullptr, d_mid_count, d_mid_list, q ? d_queues : nullptr);
    hipLaunchKernelGGL((k_descriptor<kDescMaxR, true>), dim3(g31), dim3(kThreads), 0, s, d_oct,
                       d_params, d_cand, d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap,
                       d_kps, d_desc, d_mid_count, d_mid_list, d_wide_count, d_wide_list,
                       q ? d_queues + kNumXcd : nullptr);
    hipLaunchKernelGGL(k_descriptor_wide, dim3(256), dim3(kThreads), 0, s, d_oct, d_params, d_cand,
                       d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap, d_kps, d_desc,
                       d_wide_count, d_wide_list);
}

void launch_finalize(const unsigned* d_slot_offsets, const int* d_img_slot_begin, int nimg,
                     const unsigned* d_koff, const unsigned* d_total, int cap, unsigned* d_summary,
                     unsigned* d_out_offsets, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3((nimg + 1 + 255) / 256), dim3(256), 0, s, d_slot_offsets,
                       d_img_slot_begin, nimg, d_koff, d_total, cap, d_summary, d_out_offsets);
}

void launch_desc_u8(const float* d_desc, const unsigned* d_count, int kcap, uint8_t* d_out,
                    hipStream_t s) {
    if (kcap <= 0) return;
    const int blocks = (int)std::min<long long>(1024, ((long long)kcap * 32 + 255) / 256);
    hipLaunchKernelGGL(k_desc_u8, dim3(blocks), dim3(256), 0, s, (const float4*)d_desc, d_count, kcap,
                       (unsigned*)d_out);
}

}  // namespace sh
