// kernels_detect.hip -- gfx950 kernels for everything after the pyramid:
//   k_extrema     find_scale_space_extrema      src/sift.rs:510-603
//   k_scan_*      ordered compaction offsets (exclusive scan)
//   k_compact     candidate list in the reference order (octave, layer, row, column)
//   k_refine      refine_and_filter_extrema     src/sift.rs:299-474 (+ solve 478-503)
//   k_orient      assign_orientations           src/sift.rs:605-726
//   k_descriptor  compute + normalize           src/sift.rs:917-1060, 889-913
//
// DoG values are never stored: every DoG sample is recomputed from the u8
// Gaussian levels as q[G_{s+1}] - q[G_s] with q[v] = v / 255 (IEEE division,
// kept in an LDS table), which is bit-identical to the reference's
// (G_{s+1} as f32 / 255.0) - (G_s as f32 / 255.0) (src/sift.rs:97,120).
#include "sift_internal.h"

namespace sh {

namespace {

constexpr float kPi = 3.14159274101257324219f;  // std::f32::consts::PI
constexpr float kTwoPi = 2.0f * kPi;
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

// ---------------------------------------------------------------------------
// k_compact: one wave per (octave, layer, row) slot. Reads the row's bitmask
// words, ranks the set bits with popcounts and a wave prefix sum and writes the
// candidates at offsets[slot] + rank, i.e. in the reference's order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_compact(const OctaveDesc* __restrict__ oct, int noct,
                                                      int nslots,
                                                      const unsigned long long* __restrict__ bits,
                                                      const unsigned* __restrict__ counts,
                                                      const unsigned* __restrict__ offsets,
                                                      Cand* __restrict__ cand, int cap) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (kThreads / 64) + wave;
    if (slot >= nslots) return;
    if (counts[slot] == 0) return;
    const int j = find_by(oct, noct, slot, &OctaveDesc::slot_begin);
    const OctaveDesc& O = oct[j];
    const int local = slot - O.slot_begin;
    const int s = local / O.h + 1, y = local - (s - 1) * O.h;
    const unsigned long long* wrow = bits + O.bits_begin + (long long)local * O.nwords;
    unsigned base = offsets[slot];
    for (int w0 = 0; w0 < O.nwords; w0 += 64) {
        const int wi = w0 + lane;
        unsigned long long m = wi < O.nwords ? wrow[wi] : 0ull;
        unsigned c = (unsigned)__popcll(m);
        unsigned x = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            unsigned t = __shfl_up(x, d, 64);
            if (lane >= d) x += t;
        }
        unsigned pre = base + x - c;
        unsigned tot = __shfl(x, 63, 64);
        while (m) {
            int b = __ffsll((long long)m) - 1;
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
}

// ---------------------------------------------------------------------------
// k_refine: one thread per candidate; the reference's arithmetic in the same
// f32 operation order (compiled without FMA contraction).
// ---------------------------------------------------------------------------
struct DogView {
    const uint8_t* levels;
    long long stride;
    int pitch, w, h;
    const float* q;
    __device__ __forceinline__ float at(int l, int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        const uint8_t* g0 = levels + (long long)l * stride + (size_t)y * pitch + x;
        return q[g0[stride]] - q[g0[0]];
    }
};

__global__ __launch_bounds__(kThreads) void k_refine(const OctaveDesc* __restrict__ oct,
                                                     const ImageParams* __restrict__ params,
                                                     const Cand* __restrict__ cand,
                                                     const unsigned* __restrict__ total, int cap,
                                                     Keypoint* __restrict__ out,
                                                     int* __restrict__ valid,
                                                     float* __restrict__ cand_resp) {
    __shared__ float q[256];
    fill_q(q);
    __syncthreads();
    const int n = (int)min(*total, (unsigned)cap);
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
        const Cand cd = cand[c];
        const OctaveDesc& O = oct[cd.job];
        const ImageParams& P = params[O.img];
        DogView V{O.levels, O.level_stride, O.pitch, O.w, O.h, q};
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
// k_orient: one wave (workgroup of 64) per refined keypoint, grid-stride.
// Lanes sweep the (2R+1)^2 window; each sample adds mag * weight to its bin in
// an LDS int64 fixed-point histogram (order-independent, deterministic). Lanes
// 0..35 then hold one bin each for the two [1,1,1]/3 smoothing passes, the max,
// and the peak test; a ballot ranks the peaks in bin order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_orient(const OctaveDesc* __restrict__ oct,
                                               const ImageParams* __restrict__ params,
                                               const Cand* __restrict__ cand,
                                               const unsigned* __restrict__ total, int cap,
                                               const Keypoint* __restrict__ ref,
                                               const int* __restrict__ valid,
                                               unsigned* __restrict__ npeaks,
                                               float* __restrict__ angles) {
    __shared__ float q[256];
    __shared__ unsigned long long hist[36];
    fill_q(q);
    __syncthreads();
    const int lane = threadIdx.x;
    const int n = (int)min(*total, (unsigned)cap);
    for (int c = blockIdx.x; c < n; c += gridDim.x) {
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
        if (lane < 36) hist[lane] = 0ull;
        __syncthreads();
        const int side = 2 * R + 1;
        const int nsamp = side * side;
        for (int idx = lane; idx < nsamp; idx += 64) {
            const int iy = idx / side;
            const int dy = iy - R, dx = idx - iy * side - R;
            const int xi = f2i_sat(roundf(xo + (float)dx));
            const int yi = f2i_sat(roundf(yo + (float)dy));
            if (xi < 1 || xi >= w - 1 || yi < 1 || yi >= h - 1) continue;
            const uint8_t* p = g + (size_t)yi * pitch + xi;
            const float gx = q[p[1]] - q[p[-1]];
            const float gy = q[p[pitch]] - q[p[-pitch]];
            const float mag = sqrtf(gx * gx + gy * gy);
            const float ang = atan2f(gy, gx);
            const float fdx = (float)dx, fdy = (float)dy;
            const float wt = expf(-(fdx * fdx + fdy * fdy) / denom);
            const float an = ang < 0.0f ? ang + kTwoPi : ang;
            const float bf = an * 36.0f / kTwoPi;
            const int bin = f2i_sat(floorf(bf)) % 36;
            const float contrib = mag * wt;
            atomicAdd(&hist[bin], (unsigned long long)(double)(contrib * (float)kFixScale));
        }
        __syncthreads();
        float hv = 0.0f;
        if (lane < 36) hv = (float)((double)hist[lane] * (1.0 / kFixScale));
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
// k_descriptor: one workgroup (256) per refined keypoint, looping over its
// orientation peaks. The (2R+1)^2 image-grid samples are spread over the
// threads; each in-window sample adds its 8 trilinear contributions to an LDS
// int64 fixed-point 4x4x8 histogram. Wave 0 normalises (norm, clip 0.2,
// renormalise) and writes the keypoint and its 128 floats.
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

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__global__ __launch_bounds__(kThreads) void k_descriptor(
    const OctaveDesc* __restrict__ oct, const ImageParams* __restrict__ params,
    const Cand* __restrict__ cand, const unsigned* __restrict__ total, int cap,
    const Keypoint* __restrict__ ref, const unsigned* __restrict__ npeaks,
    const float* __restrict__ angles, const unsigned* __restrict__ koff, int kcap,
    Keypoint* __restrict__ kps_out, float* __restrict__ desc_out) {
    __shared__ float q[256];
    __shared__ unsigned long long hist[128];
    fill_q(q);
    __syncthreads();
    const int n = (int)min(*total, (unsigned)cap);
    for (int c = blockIdx.x; c < n; c += gridDim.x) {
        const int np = (int)npeaks[c];
        if (np == 0) continue;
        const Keypoint base = ref[c];
        const OctaveDesc& O = oct[cand[c].job];
        const ImageParams& P = params[O.img];
        const int S = P.num_intervals;
        const int gl = min(max(base.layer, 0), S + 2);
        const uint8_t* g = O.levels + (long long)gl * O.level_stride;
        const int w = O.w, h = O.h, pitch = O.pitch;
        const float scale = pow2f(base.octave);
        const float xo = base.x / scale, yo = base.y / scale, so = base.size / scale;
        const unsigned k0 = koff[c];
        for (int pk = 0; pk < np; pk++) {
            const float angle = angles[(size_t)c * kMaxPeaks + pk];
            if (threadIdx.x < 128) hist[threadIdx.x] = 0ull;
            __syncthreads();
            if (so > 0.0f) {
                const float ca = cosf(angle), sa = sinf(angle);
                const float bw = 3.0f * so;
                const float win = bw * 4.0f;
                const float wsig = 0.5f * win;
                const float denom = 2.0f * wsig * wsig;
                const int R = f2i_sat(ceilf(win * sqrtf(2.0f) * 0.5f));
                const float wm1 = (float)(w - 1), hm1 = (float)(h - 1);
                const int side = 2 * R + 1;
                const int nsamp = side * side;
                for (int idx = threadIdx.x; idx < nsamp; idx += kThreads) {
                    const int iy = idx / side;
                    const float py = (float)(iy - R), px = (float)(idx - iy * side - R);
                    const float rx = ca * px + sa * py;
                    const float ry = -sa * px + ca * py;
                    const float xb = rx / bw + 4.0f / 2.0f - 0.5f;
                    const float yb = ry / bw + 4.0f / 2.0f - 0.5f;
                    if (!(xb > -1.0f && xb < 4.0f && yb > -1.0f && yb < 4.0f)) continue;
                    const float xs = xo + px, ys = yo + py;
                    if (xs < 0.0f || xs >= wm1 || ys < 0.0f || ys >= hm1) continue;
                    const float gx = bilinear(g, pitch, w, h, xs + 1.0f, ys, q) -
                                     bilinear(g, pitch, w, h, xs - 1.0f, ys, q);
                    const float gy = bilinear(g, pitch, w, h, xs, ys + 1.0f, q) -
                                     bilinear(g, pitch, w, h, xs, ys - 1.0f, q);
                    const float mag = sqrtf(gx * gx + gy * gy);
                    const float pa = atan2f(gy, gx);
                    const float ar = rem_euclid(pa - angle, kTwoPi);
                    const float wt = expf(-(px * px + py * py) / denom);
                    const float wm = mag * wt;
                    const float ab = ar * 8.0f / kTwoPi;
                    const int xbi = f2i_sat(floorf(xb)), ybi = f2i_sat(floorf(yb)),
                              abi = f2i_sat(floorf(ab));
                    const float dxi = xb - (float)xbi, dyi = yb - (float)ybi, dai = ab - (float)abi;
#pragma unroll
                    for (int i = 0; i < 2; i++) {
#pragma unroll
                        for (int jj = 0; jj < 2; jj++) {
#pragma unroll
                            for (int k = 0; k < 2; k++) {
                                const int ix = xbi + i, iy2 = ybi + jj;
                                int ia = (abi + k) % 8;
                                if (ia < 0) ia += 8;
                                if (ix >= 0 && ix < 4 && iy2 >= 0 && iy2 < 4) {
                                    const float wx = i == 0 ? 1.0f - dxi : dxi;
                                    const float wy = jj == 0 ? 1.0f - dyi : dyi;
                                    const float wa = k == 0 ? 1.0f - dai : dai;
                                    const float cc = wm * wx * wy * wa;
                                    atomicAdd(&hist[(iy2 * 4 + ix) * 8 + ia],
                                              (unsigned long long)(double)(cc * (float)kFixScale));
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x < 64) {
                const int l = threadIdx.x;
                float v0 = (float)((double)hist[2 * l] * (1.0 / kFixScale));
                float v1 = (float)((double)hist[2 * l + 1] * (1.0 / kFixScale));
                const float nrm = sqrtf(wave_sum(v0 * v0 + v1 * v1));
                if (nrm >= 1e-8f) {
                    const float inv = 1.0f / nrm;
                    v0 = fminf(v0 * inv, 0.2f);
                    v1 = fminf(v1 * inv, 0.2f);
                    const float n2 = sqrtf(wave_sum(v0 * v0 + v1 * v1));
                    if (n2 >= 1e-8f) {
                        const float inv2 = 1.0f / n2;
                        v0 *= inv2;
                        v1 *= inv2;
                    }
                }
                const unsigned kidx = k0 + (unsigned)pk;
                if (kidx < (unsigned)kcap) {
                    float2* d = (float2*)(desc_out + (size_t)kidx * 128);
                    d[l] = make_float2(v0, v1);
                    if (l == 0) {
                        Keypoint o = base;
                        o.angle = angle;
                        kps_out[kidx] = o;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Per-image keypoint ranges: summary[2+i] = koff[first candidate of image i];
// summary[0] = total candidates (may exceed cap), summary[1] = total keypoints.
__global__ void k_finalize(const unsigned* __restrict__ slot_offsets,
                           const int* __restrict__ img_slot_begin, int nimg,
                           const unsigned* __restrict__ koff, const unsigned* __restrict__ total,
                           int cap, unsigned* __restrict__ summary) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nimg) return;
    const unsigned n = min(*total, (unsigned)cap);
    const unsigned cb = min(slot_offsets[img_slot_begin[i]], n);
    summary[2 + i] = koff[cb];
    if (i == 0) {
        summary[0] = *total;
        summary[1] = koff[n];
    }
}

}  // namespace

void launch_extrema(const OctaveDesc* d_oct, int noct, int total_rows, const ImageParams* d_params,
                    unsigned long long* d_bits, unsigned* d_counts, hipStream_t s) {
    if (total_rows <= 0) return;
    hipLaunchKernelGGL(k_extrema, dim3(total_rows), dim3(kThreads), 0, s, d_oct, noct, d_params,
                       d_bits, d_counts);
}

// n_dev may be null (then n_max is the count). out needs n_max+1 entries;
// tmp needs ceil(n_max/4096)+1 entries.
void launch_scan_dev(const unsigned* d_in, int n_max, const unsigned* n_dev, unsigned* d_out,
                     unsigned* d_tmp, hipStream_t s) {
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
    const int per = kThreads / 64;
    hipLaunchKernelGGL(k_compact, dim3((nslots + per - 1) / per), dim3(kThreads), 0, s, d_oct, noct,
                       nslots, d_bits, d_counts, d_offsets, d_cand, cap);
}

static int grid_for(int cap, int per_block, int max_blocks) {
    int g = (cap + per_block - 1) / per_block;
    return g < 1 ? 1 : (g > max_blocks ? max_blocks : g);
}

void launch_refine(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, Keypoint* d_ref, int* d_valid,
                   float* d_cand_resp, hipStream_t s) {
    hipLaunchKernelGGL(k_refine, dim3(grid_for(cap, kThreads, 2048)), dim3(kThreads), 0, s, d_oct,
                       d_params, d_cand, d_total, cap, d_ref, d_valid, d_cand_resp);
}

void launch_orient(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                   const unsigned* d_total, int cap, const Keypoint* d_ref, const int* d_valid,
                   unsigned* d_npeaks, float* d_angles, hipStream_t s) {
    hipLaunchKernelGGL(k_orient, dim3(grid_for(cap, 1, 16384)), dim3(64), 0, s, d_oct, d_params,
                       d_cand, d_total, cap, d_ref, d_valid, d_npeaks, d_angles);
}

void launch_descriptor(const OctaveDesc* d_oct, const ImageParams* d_params, const Cand* d_cand,
                       const unsigned* d_total, int cap, const Keypoint* d_ref,
                       const unsigned* d_npeaks, const float* d_angles, const unsigned* d_koff,
                       int kcap, Keypoint* d_kps, float* d_desc, hipStream_t s) {
    hipLaunchKernelGGL(k_descriptor, dim3(grid_for(cap, 1, 8192)), dim3(kThreads), 0, s, d_oct,
                       d_params, d_cand, d_total, cap, d_ref, d_npeaks, d_angles, d_koff, kcap,
                       d_kps, d_desc);
}

void launch_finalize(const unsigned* d_slot_offsets, const int* d_img_slot_begin, int nimg,
                     const unsigned* d_koff, const unsigned* d_total, int cap, unsigned* d_summary,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3((nimg + 1 + 255) / 256), dim3(256), 0, s, d_slot_offsets,
                       d_img_slot_begin, nimg, d_koff, d_total, cap, d_summary);
}

}  // namespace sh
