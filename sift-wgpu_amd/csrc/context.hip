// context.hip -- host side of libsift_hip.so: parameter math, device memory,
// batch planning, the launch sequence and the C ABI of include/sift_hip.h.
//
// Host f32 arithmetic here mirrors the reference's operation order (compiled
// with -ffp-contract=off, SSE f32, glibc libm): level sigmas
// (src/sift.rs:132-147), initial blur (src/sift.rs:1205-1209), imageproc's
// Gaussian taps and image's Lanczos3 weights. These feed the kernels as tables.
#include <dlfcn.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "sift_hip.h"
#include "sift_internal.h"
#include "sift_context.h"

using namespace sh;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SIFT_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr float kPiF = 3.14159274101257324219f;

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// f32::powi with a run-time exponent (compiler-rt __powisf2: square-and-multiply).
float powi_f(float a, int b) {
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

// imageproc gaussian_kernel_f32: r = ceil(2 sigma); taps = (sqrt(2 pi) sigma)^-1 *
// exp(-i^2 / (2 sigma^2)); not normalised.
int gaussian_taps(float sigma, std::vector<float>& taps) {
    int r = (int)ceilf(2.0f * sigma);
    int n = 2 * r + 1;
    taps.assign(n, 0.0f);
    float norm = 1.0f / (sqrtf(2.0f * kPiF) * sigma);
    float two_s2 = 2.0f * (sigma * sigma);
    for (int i = 0; i <= r; i++) {
        float x = (float)i;
        float v = norm * expf(-(x * x) / two_s2);
        taps[r + i] = v;
        taps[r - i] = v;
    }
    return n;
}

// image 0.25.6 Lanczos3: sinc(t) = sin(t*PI)/(t*PI), window 3*sratio, weights
// divided by their f32 sum.
float sinc_f(float t) {
    float a = t * kPiF;
    if (t == 0.0f) return 1.0f;
    return sinf(a) / a;
}
float lanczos3_f(float x) { return fabsf(x) < 3.0f ? sinc_f(x) * sinc_f(x / 3.0f) : 0.0f; }

}  // namespace

namespace sh {

int set_error(int code, const char* msg) { return fail(code, msg); }

// ---------------------------------------------------------------------------
// Table store: taps and Lanczos tables in one growing device buffer.
// ---------------------------------------------------------------------------
int TableStore::ensure_uploaded(hipStream_t stream) {
    if (host.size() == uploaded) return SIFT_OK;
    if (host.size() > dev_cap) {
        // rare: the first 2 MiB hold the tables of many parameter sets and sizes
        size_t ncap = std::max(host.size() * 2, (size_t)2 << 20);
        void* nd = nullptr;
        HIP_TRY(hipMalloc(&nd, ncap));
        if (dev) {
            HIP_TRY(hipStreamSynchronize(stream));
            (void)hipFree(dev);
        }
        dev = (char*)nd;
        dev_cap = ncap;
        uploaded = 0;
    }
    const size_t n = host.size() - uploaded;
    if (n > bounce_cap) {
        HIP_TRY(hipStreamSynchronize(stream));
        if (bounce) (void)hipHostFree(bounce);
        bounce = nullptr;
        bounce_cap = 0;
        size_t ncap = std::max(n * 2, (size_t)1 << 20);
        HIP_TRY(hipHostMalloc((void**)&bounce, ncap, hipHostMallocDefault));
        bounce_cap = ncap;
    }
    memcpy(bounce, host.data() + uploaded, n);
    HIP_TRY(hipMemcpyAsync(dev + uploaded, bounce, n, hipMemcpyHostToDevice, stream));
    uploaded = host.size();
    return SIFT_OK;
}

size_t TableStore::append(const void* p, size_t bytes) {
    size_t off = round_up(host.size(), 256);
    host.resize(off + bytes);
    memcpy(host.data() + off, p, bytes);
    return off;
}

const ParamTables& Context::tables_for(const sift_params_t& p) {
    ParamKey key{p.sigma, p.assumed_blur, p.num_intervals, p.num_octaves};
    auto it = param_tables.find(key);
    if (it != param_tables.end()) return it->second;
    ParamTables T;
    const int S = (int)p.num_intervals, L = S + 3, O = (int)p.num_octaves;
    auto add_taps = [&](float sigma, TapsRef& out) {
        std::vector<float> taps;
        int n = gaussian_taps(sigma, taps);
        int npad = (int)round_up((size_t)n, kTapBlock);
        taps.resize(npad, 0.0f);
        out.ntaps = n;
        out.npad = npad;
        out.off = tables.append(taps.data(), taps.size() * sizeof(float));
        out.copy = false;
    };
    // initial blur (src/sift.rs:1205-1214)
    float s0 = 0.0f;
    if (p.sigma > p.assumed_blur) s0 = sqrtf(p.sigma * p.sigma - p.assumed_blur * p.assumed_blur);
    if (s0 > 1e-4f) add_taps(s0, T.initial);
    else T.initial.copy = true;
    // level sigmas (src/sift.rs:132-147)
    float k = powf(2.0f, 1.0f / (float)S);
    std::vector<float> tgt(L);
    for (int s = 0; s < L; s++) tgt[s] = p.sigma * powi_f(k, s);
    T.level.resize((size_t)O * L);
    for (int o = 0; o < O; o++) {
        float prev = p.sigma * powi_f(2.0f, o);
        for (int s = 1; s < L; s++) {
            float t = tgt[s] * powi_f(2.0f, o);
            float b = sqrtf(t * t - prev * prev);
            TapsRef& R = T.level[(size_t)o * L + s];
            if (b < 1e-4f || !(b == b)) R.copy = true;  // src/sift.rs:149-150
            else add_taps(b, R);
            prev = t;
        }
    }
    return param_tables.emplace(key, std::move(T)).first->second;
}

const LanczosRef& Context::lanczos_for(int in, int out) {
    auto key = std::make_pair(in, out);
    auto it = lanczos.find(key);
    if (it != lanczos.end()) return it->second;
    float ratio = (float)in / (float)out;
    float sratio = ratio < 1.0f ? 1.0f : ratio;
    float support = 3.0f * sratio;
    std::vector<int> left(out), cnt(out);
    int maxt = 0;
    for (int j = 0; j < out; j++) {
        float c = ((float)j + 0.5f) * ratio;
        long long l = (long long)floorf(c - support);
        l = std::min(std::max(l, 0LL), (long long)in - 1);
        long long r = (long long)ceilf(c + support);
        r = std::min(std::max(r, l + 1), (long long)in);
        left[j] = (int)l;
        cnt[j] = (int)(r - l);
        maxt = std::max(maxt, cnt[j]);
    }
    std::vector<float> wts((size_t)out * maxt, 0.0f);
    for (int j = 0; j < out; j++) {
        float c = ((float)j + 0.5f) * ratio - 0.5f;
        float sum = 0.0f;
        float* wr = &wts[(size_t)j * maxt];
        for (int i = 0; i < cnt[j]; i++) {
            float wv = lanczos3_f(((float)(left[j] + i) - c) / sratio);
            wr[i] = wv;
            sum += wv;
        }
        for (int i = 0; i < cnt[j]; i++) wr[i] /= sum;
    }
    LanczosRef R;
    R.maxtaps = maxt;
    R.left_off = tables.append(left.data(), left.size() * sizeof(int));
    R.count_off = tables.append(cnt.data(), cnt.size() * sizeof(int));
    R.w_off = tables.append(wts.data(), wts.size() * sizeof(float));
    R.left = std::move(left);
    R.count = std::move(cnt);
    return lanczos.emplace(key, R).first->second;
}

// ---------------------------------------------------------------------------
// Graph cache
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kMaxGraphs = 32;  // per context

bool graphs_enabled_by_env() {
    static const bool on = [] {
        const char* e = getenv("SIFT_HIP_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}
}  // namespace

// The fused Lanczos3 octave step is used wherever its LDS span fits;
// SIFT_HIP_RESIZE_FUSED=0 (read on every call) selects the two-pass form, so
// that tests can compare the two.
static bool resize_fusion_enabled() {
    const char* e = getenv("SIFT_HIP_RESIZE_FUSED");
    return !(e && e[0] == '0');
}

// Initial candidate capacity max(min, pixel-layers / divisor): 4096 and 128 by
// default; SIFT_HIP_CAP_MIN / SIFT_HIP_CAP_DIVISOR override them (read on every
// call) so that tests can force the overflow-and-retry path.
static size_t initial_cap_min() {
    const char* e = getenv("SIFT_HIP_CAP_MIN");
    const long v = e ? atol(e) : 0;
    return v > 0 ? (size_t)v : 4096;
}
static size_t initial_cap_divisor() {
    const char* e = getenv("SIFT_HIP_CAP_DIVISOR");
    const long v = e ? atol(e) : 0;
    return v > 0 ? (size_t)v : 128;
}

void Context::clear_graphs() {
    for (GraphEntry& e : graphs) (void)hipGraphExecDestroy(e.exec);
    graphs.clear();
    seen_once.clear();
}

// ---------------------------------------------------------------------------
// Arena: one device allocation, carved per call; grows on demand.
// ---------------------------------------------------------------------------
int Context::ensure_arena(size_t bytes) {
    if (bytes <= arena_cap) return SIFT_OK;
    if (arena_limit && bytes > arena_limit)
        return fail(SIFT_ERR_NOMEM, "batch needs " + std::to_string(bytes) +
                                        " device bytes, more than this lane's share of the HBM budget (" +
                                        std::to_string(arena_limit) + ")");
    size_t ncap = std::max(bytes + bytes / 4, arena_floor);
    if (arena_limit) ncap = std::min(ncap, arena_limit);
    if (arena) {
        HIP_TRY(hipStreamSynchronize(stream));
        clear_graphs();  // they address the old arena
        (void)hipFree(arena);
        arena = nullptr;
        arena_cap = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, ncap);
    if (e != hipSuccess) return fail(SIFT_ERR_NOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
    arena = (char*)p;
    arena_cap = ncap;
    return SIFT_OK;
}

int Context::ensure_pinned_out(size_t bytes) {
    if (bytes <= pinned_out_cap) return SIFT_OK;
    // geometric growth: a stream of growing batches reallocates O(log) times
    size_t ncap = std::max(std::max(bytes + bytes / 4, 2 * pinned_out_cap), (size_t)1 << 20);
    if (pinned_out) {
        HIP_TRY(hipStreamSynchronize(stream));
        (void)hipHostFree(pinned_out);
        pinned_out = nullptr;
        pinned_out_cap = 0;
    }
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, ncap, hipHostMallocDefault);
    if (e != hipSuccess) return fail(SIFT_ERR_NOMEM, "pinned host allocation failed");
    pinned_out = (char*)p;
    pinned_out_cap = ncap;
    return SIFT_OK;
}

void Context::release() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    clear_graphs();
    prof_resolve();
    for (hipEvent_t e : prof_events) (void)hipEventDestroy(e);
    prof_events.clear();
    if (arena) (void)hipFree(arena);
    if (tables.dev) (void)hipFree(tables.dev);
    if (tables.bounce) (void)hipHostFree(tables.bounce);
    if (pinned) (void)hipHostFree(pinned);
    if (pinned_out) (void)hipHostFree(pinned_out);
    if (ext_pinned) (void)hipHostFree(ext_pinned);
    if (stream) (void)hipStreamDestroy(stream);
    arena = nullptr;
    tables.dev = nullptr;
    tables.bounce = nullptr;
    pinned = pinned_out = ext_pinned = nullptr;
    ext_pinned_cap = 0;
    stream = nullptr;
    arena_cap = tables.dev_cap = tables.bounce_cap = pinned_cap = pinned_out_cap = 0;
}

int Context::ensure_pinned(size_t bytes) {
    if (bytes <= pinned_cap) return SIFT_OK;
    // geometric growth: a stream of growing batches reallocates O(log) times
    size_t ncap = std::max(std::max(bytes + bytes / 4, 2 * pinned_cap), (size_t)1 << 20);
    if (pinned) {
        HIP_TRY(hipStreamSynchronize(stream));
        (void)hipHostFree(pinned);
        pinned = nullptr;
        pinned_cap = 0;
    }
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, ncap, hipHostMallocDefault);
    if (e != hipSuccess) return fail(SIFT_ERR_NOMEM, "pinned host allocation failed");
    pinned = (char*)p;
    pinned_cap = ncap;
    return SIFT_OK;
}

// ---------------------------------------------------------------------------
// Planning and launch
// ---------------------------------------------------------------------------
namespace {

struct Carver {
    size_t off = 0;
    size_t take(size_t bytes, size_t align = 256) {
        off = round_up(off, align);
        size_t r = off;
        off += bytes;
        return r;
    }
};

struct LaunchRec {
    int kind;        // KernelId
    size_t job_off;  // offset in the job blob
    int njobs, tiles, max_pad;
    int variant;     // k_blur_h: outputs per lane; k_blur_v_narrow: staged rows
    double bytes;    // algorithmic HBM bytes of the launch
    double ops;      // f32 operations the reference's arithmetic prescribes (profile ops)
};

enum KernelId {
    K_BLUR_H = 0, K_BLUR_V, K_COPY, K_BLUR_H_GENERIC, K_RESIZE_V, K_RESIZE_H, K_EXTREMA, K_SCAN,
    K_COMPACT, K_REFINE, K_ORIENT, K_DESCRIPTOR, K_FINALIZE, K_BLUR_V_NARROW, K_BLUR_V_GENERIC,
    K_BLUR_H2, K_BLUR_H_SMALL, K_BLUR_V_SMALL, K_LUMA, K_DESC_U8, K_RESIZE_FUSED, K_BLUR_V_SMALL8,
    K_BLUR_H16, K_BLUR_V8, K_COUNT
};
// Names as rocprofv3 reports the kernels (template arguments included).
const char* const kKernelNames[K_COUNT] = {
    "k_blur_h<8>", "k_blur_v<4>", "k_copy_level", "k_blur_h_generic", "k_resize_v", "k_resize_h",
    "k_extrema_tile", "k_scan", "k_compact", "k_refine", "k_orient", "k_descriptor",
    "k_finalize", "k_blur_v_narrow", "k_blur_v_generic", "k_blur_h<2>", "k_blur_h_small",
    "k_blur_v_small", "k_luma", "k_desc_u8", "k_resize_fused", "k_blur_v_small8", "k_blur_h<16>",
    "k_blur_v<8>"};


double blur_job_bytes(const BlurJob& j, int kind) {
    double a = (double)j.w * j.h;
    if ((kind == K_BLUR_V || kind == K_BLUR_V_NARROW || kind == K_BLUR_V_GENERIC ||
         kind == K_BLUR_V_SMALL || kind == K_BLUR_V_SMALL8 || kind == K_BLUR_V8) && j.dog)
        return a * (2.0 + 1.0 + 4.0);
    if (kind == K_COPY && j.dog) return a * (2.0 + 4.0);
    return a * 2.0;
}
// A multiply and an add per Gaussian tap (Lanczos weight) and output pixel:
// the reference's f32 arithmetic, which the kernels keep without contraction.
double blur_job_ops(const BlurJob& j, int kind) {
    if (kind == K_COPY) return 0.0;
    return 2.0 * (double)j.w * j.h * j.ntaps;
}
double resize_job_ops(const ResizeJob& j, int kind) {
    const double v = 2.0 * (double)j.in_w * j.out_h * (kind == K_RESIZE_FUSED ? j.vmaxtaps : j.maxtaps);
    const double h = 2.0 * (double)j.out_w * j.out_h * j.maxtaps;
    if (kind == K_RESIZE_FUSED) return v + h;
    return kind == K_RESIZE_V ? v : h;
}

double luma_job_bytes(const LumaJob& j) {
    return (double)j.w * j.h * (pixel_bytes((uint32_t)j.format) + 1);
}
double resize_job_bytes(const ResizeJob& j, int kind) {
    if (kind == K_RESIZE_FUSED) return (double)j.in_w * j.in_h + (double)j.out_w * j.out_h;
    if (kind == K_RESIZE_V) return (double)j.in_w * j.in_h + (double)j.in_w * j.out_h * 4.0;
    return (double)j.in_w * j.out_h * 4.0 + (double)j.out_w * j.out_h;
}

// ---------------------------------------------------------------------------
// Blur pass forms: the choices of the tuning table
// ---------------------------------------------------------------------------
// Horizontal: row tiles of 512 or 128 columns x 4 rows, 8 or 2 outputs per
// lane (k_blur_h<8> / <2>); 64 x 4 tiles with one output per lane
// (k_blur_h_small); one row of 256 outputs per work-group (k_blur_h_generic).
// Vertical: 256 x 16 tiles, 4 x 4 outputs per lane (k_blur_v); 64 x 16 tiles
// with the input rows staged in LDS (k_blur_v_narrow); 64 x 4 tiles, one output
// per lane (k_blur_v_small); one row per work-group (k_blur_v_generic). Every
// form computes the same exact sums, so any valid choice gives the same level.
const char* const kBlurHNames[BH_COUNT] = {"h8", "h2", "hsmall", "hgeneric", "h16"};
const char* const kBlurVNames[BV_COUNT] = {"v", "vnarrow", "vsmall", "vgeneric", "vsmall8", "v8"};

int blur_form_id(const char* name, bool vertical) {
    if (!name) return -1;
    const char* const* names = vertical ? kBlurVNames : kBlurHNames;
    const int n = vertical ? BV_COUNT : BH_COUNT;
    for (int i = 0; i < n; i++)
        if (strcmp(names[i], name) == 0) return i;
    return -1;
}

// LDS of the 4 staged rows of k_blur_h<px> (padded as it pads them)
int blur_h_row_lds(int px, int npad) {
    const int hrow = 64 * px + npad + 8;
    return (hrow + 8 + ((hrow + 8) >> (px == 16 ? 4 : px == 8 ? 3 : 5)) + 1) * 4 * (int)sizeof(float);
}

bool blur_h_valid(int form, int w, int h, int ntaps, int npad) {
    (void)w;
    (void)h;
    const int n4 = (ntaps + 3) & ~3;
    switch (form) {
        case BH_W8: return blur_h_row_lds(8, npad) <= 65536;
        case BH_W16: return blur_h_row_lds(16, npad) <= 65536;
        case BH_W2: return blur_h_row_lds(2, npad) <= 65536;
        case BH_SMALL: return n4 <= kSmallMaxTaps4;
        case BH_GENERIC: return true;
        default: return false;
    }
}

bool blur_v_valid(int form, int w, int h, int ntaps) {
    (void)w;
    const int n4 = (ntaps + 3) & ~3, r = ntaps / 2;
    switch (form) {
        case BV_WIDE: return true;
        case BV_NARROW: return std::min(h, 16 + 2 * r) * 64 <= 65536;
        case BV_SMALL: return h <= kSmallMaxRows && n4 <= kSmallMaxTaps4;
        case BV_SMALL8: return n4 <= kSmallMaxTaps4 && std::min(h, 4 + 2 * r) <= kSmallMaxRows8;
        case BV_WIDE8: return ntaps == 7 || ntaps == 9 || ntaps == 11 || ntaps == 13 || ntaps == 15;
        case BV_GENERIC: return true;
        default: return false;
    }
}

// The built-in choice: small levels and long kernels on narrow levels take one
// output per lane; wide levels the 16- or 8-output row tiles (16 when the last
// 1024-column tile of a row is at least 7/8 full) and the 4 x 4 column tiles.
std::pair<int, int> blur_heuristic(int w, int h, int ntaps, int npad) {
    const bool full16 = w >= 1536 && ((w + 1023) / 1024 * 1024 - w) * 8 <= w;
    const int px = full16 ? 16 : w >= 384 ? 8 : 2;
    const int n4 = (ntaps + 3) & ~3, r = ntaps / 2;
    const bool small_level = (long long)w * h <= 65536 || r >= 24;
    int hf, vf;
    if (w < 384 && n4 <= kSmallMaxTaps4 && small_level) hf = BH_SMALL;
    else if (blur_h_row_lds(px, npad) <= 65536) hf = px == 16 ? BH_W16 : px == 8 ? BH_W8 : BH_W2;
    else if (blur_h_row_lds(8, npad) <= 65536) hf = BH_W8;
    else hf = BH_GENERIC;
    if (h <= kSmallMaxRows && n4 <= kSmallMaxTaps4 && (small_level || w < 192 || ntaps > 64)) vf = BV_SMALL;
    else if (w >= 192 && ntaps <= 64) vf = BV_WIDE;
    else if (blur_v_valid(BV_NARROW, w, h, ntaps)) vf = BV_NARROW;
    else vf = BV_GENERIC;
    return {hf, vf};
}

// Tile counts of a job in a form.
void blur_h_tiles(BlurJob& J, int form) {
    const int w = J.w, h = J.h;
    switch (form) {
        case BH_W8: J.tiles_x = (w + 511) / 512; J.tiles = J.tiles_x * ((h + 3) / 4); break;
        case BH_W16: J.tiles_x = (w + 1023) / 1024; J.tiles = J.tiles_x * ((h + 3) / 4); break;
        case BH_W2: J.tiles_x = (w + 127) / 128; J.tiles = J.tiles_x * ((h + 3) / 4); break;
        case BH_SMALL: J.tiles_x = (w + 63) / 64; J.tiles = J.tiles_x * ((h + 3) / 4); break;
        default: J.tiles_x = (w + 255) / 256; J.tiles = J.tiles_x * h; break;
    }
}
void blur_v_tiles(BlurJob& J, int form) {
    const int w = J.w, h = J.h;
    switch (form) {
        case BV_WIDE: J.tiles_x = (w + 255) / 256; J.tiles = J.tiles_x * ((h + 15) / 16); break;
        case BV_WIDE8: J.tiles_x = (w + 255) / 256; J.tiles = J.tiles_x * ((h + 31) / 32); break;
        case BV_NARROW: J.tiles_x = (w + 63) / 64; J.tiles = J.tiles_x * ((h + 15) / 16); break;
        case BV_SMALL:
        case BV_SMALL8: J.tiles_x = (w + 63) / 64; J.tiles = J.tiles_x * ((h + 3) / 4); break;
        default: J.tiles_x = (w + 255) / 256; J.tiles = J.tiles_x * h; break;
    }
}

// SIFT_HIP_BLUR_FORCE="<h form>,<v form>" (read on every call): forms used for
// every job they are valid for, so that tests can compare them.
std::pair<int, int> blur_forced() {
    const char* e = getenv("SIFT_HIP_BLUR_FORCE");
    if (!e || !e[0]) return {-1, -1};
    std::string s(e);
    const size_t c = s.find(',');
    const std::string a = s.substr(0, c), b = c == std::string::npos ? "" : s.substr(c + 1);
    return {blur_form_id(a.c_str(), false), blur_form_id(b.c_str(), true)};
}

}  // namespace

std::pair<int, int> Context::blur_choice(int w, int h, int ntaps, int npad, bool many,
                                         std::pair<int, int> forced) const {
    std::pair<int, int> c = blur_heuristic(w, h, ntaps, npad);
    auto it = tune.find(TuneKey{w, h, ntaps, many ? 1 : 0});
    if (it != tune.end()) {
        if (blur_h_valid(it->second.first, w, h, ntaps, npad)) c.first = it->second.first;
        if (blur_v_valid(it->second.second, w, h, ntaps)) c.second = it->second.second;
    }
    if (forced.first >= 0 && blur_h_valid(forced.first, w, h, ntaps, npad)) c.first = forced.first;
    if (forced.second >= 0 && blur_v_valid(forced.second, w, h, ntaps)) c.second = forced.second;
    return c;
}

hipEvent_t Context::prof_event() {
    if (!prof_events.empty()) {
        hipEvent_t e = prof_events.back();
        prof_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void Context::prof_begin(int kernel, double bytes, hipEvent_t* a) {
    (void)kernel;
    (void)bytes;
    *a = nullptr;
    if (!profiling || prof_blocked || (prof_only >= 0 && kernel != prof_only)) return;
    *a = prof_event();
    if (*a) (void)hipEventRecord(*a, stream);
}

void Context::prof_end(int kernel, double bytes, hipEvent_t a, double ops) {
    if (!profiling || !a) return;
    hipEvent_t b = prof_event();
    if (!b) return;
    (void)hipEventRecord(b, stream);
    prof_pending.push_back({kernel, bytes, ops, a, b});
}

void Context::prof_resolve() {
    for (const ProfPending& p : prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ProfStat& S = prof_stats[p.kernel];
            S.launches++;
            S.total_ms += ms;
            S.bytes += p.bytes;
            S.ops += p.ops;
        }
        prof_events.push_back(p.a);
        prof_events.push_back(p.b);
    }
    prof_pending.clear();
}

int Context::run(const Batch& B, RunOutputs* outs) {
    const int n = (int)B.images.size();
    // every public entry point synchronises before returning, so the previous
    // call's events are complete
    prof_resolve();
    if (n == 0) return fail(SIFT_ERR_INVALID, "empty batch");
    // --- tables (host) ---
    std::vector<const ParamTables*> ptab(n);
    for (int i = 0; i < n; i++) ptab[i] = &tables_for(B.params[i]);
    struct OctShape { int w, h, pitch; };
    std::vector<std::vector<OctShape>> shapes(n);
    for (int i = 0; i < n; i++) {
        int w = (int)B.images[i].width, h = (int)B.images[i].height;
        for (uint32_t o = 0; o < B.params[i].num_octaves; o++) {
            shapes[i].push_back({w, h, (int)round_up((size_t)std::max(w, 1), 64)});
            w /= 2;
            h /= 2;
        }
        for (uint32_t o = 0; o + 1 < B.params[i].num_octaves; o++) {
            const OctShape& a = shapes[i][o];
            const OctShape& b = shapes[i][o + 1];
            if (a.w > 0 && a.h > 0 && b.w > 0 && b.h > 0) {
                lanczos_for(a.h, b.h);
                lanczos_for(a.w, b.w);
            }
        }
    }
    // the previous call synchronised the stream before returning, so the table
    // bounce buffer and the pinned staging are free to reuse
    // (a call in external memory copies the whole table store to its base instead)
    int rc = B.ext ? SIFT_OK : tables.ensure_uploaded(stream);
    if (rc) return rc;

    // --- device layout ---
    Carver C;
    if (B.ext) C.off = round_up(tables.host.size(), 256);  // the tables come first
    std::vector<size_t> in_off(n), rtmp_off(n);
    std::vector<std::vector<size_t>> lev_off(n), dog_off(n), htemp_off(n);
    std::vector<int> in_pitch(n);
    // colour inputs: raw pixels (host inputs are staged at color_off) and the
    // grey image k_luma writes at in_off
    std::vector<size_t> color_off(n, 0);
    std::vector<int> color_pitch(n, 0);
    for (int i = 0; i < n; i++) {
        const sift_image_t& im = B.images[i];
        const uint32_t fmt = B.format(i);
        int S = (int)B.params[i].num_intervals, L = S + 3;
        if (B.device_input && fmt == SIFT_PIXEL_GRAY8) {
            in_pitch[i] = (int)(im.stride ? im.stride : im.width);
        } else {
            in_pitch[i] = (int)round_up(im.width, 64);
            in_off[i] = C.take((size_t)in_pitch[i] * im.height);
        }
        if (fmt != SIFT_PIXEL_GRAY8) {
            const int bpp = pixel_bytes(fmt);
            if (B.device_input) {
                color_pitch[i] = (int)(im.stride ? im.stride : (size_t)im.width * bpp);
            } else {
                color_pitch[i] = (int)round_up((size_t)im.width * bpp, 64);
                color_off[i] = C.take((size_t)color_pitch[i] * im.height);
            }
        }
        lev_off[i].resize(shapes[i].size());
        dog_off[i].resize(shapes[i].size());
        htemp_off[i].resize(shapes[i].size());
        for (size_t o = 0; o < shapes[i].size(); o++) {
            const OctShape& s = shapes[i][o];
            lev_off[i][o] = C.take((size_t)L * s.pitch * s.h + 64);
            if (B.want_dog) dog_off[i][o] = C.take((size_t)(L - 1) * s.w * s.h * sizeof(float));
            // an H-pass temporary per octave: two octaves of an image can blur in one step
            htemp_off[i][o] = C.take((size_t)s.pitch * s.h + 64);
        }
        const OctShape& s0 = shapes[i][0];
        rtmp_off[i] = C.take((size_t)s0.pitch * std::max(s0.h / 2, 1) * sizeof(float));
    }

    // octave table for detection; the tiled extremum kernel needs one S per batch
    int S0 = (int)B.params[0].num_intervals;
    bool tile_ext = extrema_tile_supported(S0);
    for (int i = 1; i < n; i++) tile_ext = tile_ext && (int)B.params[i].num_intervals == S0;
    // strips of kExtStripRows rows when the batch has enough 16-row bands to fill
    // the GPU with them; single bands otherwise (single frames, small batches)
    long long bands16 = 0;
    for (int i = 0; i < n; i++)
        for (const OctShape& s : shapes[i])
            if (s.w > 0 && s.h > 0) bands16 += (long long)((s.w + 63) / 64) * ((s.h + 15) / 16);
    const int strip_rows = bands16 >= kExtFillBands ? kExtStripRows : 16;
    int total_tiles = 0;
    std::vector<OctaveDesc> octs;
    std::vector<int> img_slot_begin(n + 1);
    int nslots = 0, total_rows = 0;
    long long nbits = 0;
    size_t pix_layers = 0;
    for (int i = 0; i < n; i++) {
        img_slot_begin[i] = nslots;
        int S = (int)B.params[i].num_intervals, L = S + 3;
        for (size_t o = 0; o < shapes[i].size(); o++) {
            const OctShape& s = shapes[i][o];
            if (s.w <= 0 || s.h <= 0) continue;
            OctaveDesc d{};
            d.level_stride = (long long)s.pitch * s.h;
            d.w = s.w;
            d.h = s.h;
            d.pitch = s.pitch;
            d.octave = (int)o;
            d.img = i;
            d.nlevels = L;
            d.slot_begin = nslots;
            d.bits_begin = nbits;
            d.nwords = (s.w + 63) / 64;
            d.row_begin = tile_ext ? total_tiles : total_rows;
            total_tiles += ((s.w + 63) / 64) * ((s.h + strip_rows - 1) / strip_rows);
            nslots += S * s.h;
            nbits += (long long)S * s.h * d.nwords;
            total_rows += s.h;
            pix_layers += (size_t)S * s.w * s.h;
            octs.push_back(d);
        }
    }
    img_slot_begin[n] = nslots;
    const int noct = (int)octs.size();

    // capacities (grow-and-retry on overflow)
    size_t want_cap = std::max<size_t>(initial_cap_min(), pix_layers / initial_cap_divisor());
    if (!B.ext) {
        if (cand_cap < want_cap) cand_cap = want_cap;
        if (kp_cap < cand_cap + cand_cap / 2) kp_cap = cand_cap + cand_cap / 2;
    }
    // external memory: the call's own candidate capacity, the caller's output capacity
    const int cap = B.ext ? (int)std::min<size_t>(std::max(B.ext->cand_cap, want_cap), INT_MAX / 256)
                          : (int)cand_cap;
    const int kcap = B.ext ? (int)std::min<size_t>(B.ext->kcap, INT_MAX / 512) : (int)kp_cap;

    const size_t bits_off = C.take((size_t)std::max<long long>(nbits, 1) * 8);
    const size_t counts_off = C.take((size_t)(nslots + kCountExtra) * 4);  // + lists and queue heads
    const size_t offs_off = C.take((size_t)(nslots + 2) * 4);
    const size_t scantmp1_off = C.take((size_t)(nslots / kScanTileElems + 4) * 4);
    const size_t cand_off = C.take((size_t)cap * sizeof(Cand));
    const size_t candresp_off = C.take((size_t)cap * 4);
    const size_t ref_off = C.take((size_t)cap * sizeof(Keypoint));
    const size_t valid_off = C.take((size_t)cap * 4);
    const size_t npk_off = C.take((size_t)(cap + 1) * 4);
    const size_t ang_off = C.take((size_t)cap * kMaxPeaks * 4);
    const size_t koff_off = C.take((size_t)(cap + 2) * 4);
    const size_t scantmp2_off = C.take((size_t)(cap / kScanTileElems + 4) * 4);
    const size_t kps_off = B.ext ? 0 : C.take((size_t)kcap * sizeof(Keypoint));
    const size_t desc_off = B.ext ? 0 : C.take((size_t)kcap * 128 * 4);
    const size_t summary_off = C.take((size_t)(n + 3) * 4);
    const size_t wide_off = C.take((size_t)cap * 4);  // keypoints for the wide-window descriptor
    const size_t mid_off = C.take((size_t)cap * 4);   // keypoints for the R <= 31 descriptor
    const size_t descu8_off = B.desc_u8 ? C.take((size_t)kcap * 128) : 0;

    // --- job blob (host), uploaded to the end of the arena ---
    std::vector<char> blob;
    auto blob_take = [&](size_t bytes) {
        size_t off = round_up(blob.size(), 256);
        blob.resize(off + bytes);
        return off;
    };
    std::vector<LaunchRec> launches;
    int maxO = 0, Lmax = 0;
    for (int i = 0; i < n; i++) {
        maxO = std::max(maxO, (int)shapes[i].size());
        Lmax = std::max(Lmax, (int)B.params[i].num_intervals + 3);
    }
    // per blur step up to 12 blur launches (five H forms, six V forms, the copy)
    // and 3 Lanczos3 launches; at most 1 + maxO * Lmax steps
    const size_t max_launches = (size_t)(1 + maxO * Lmax) * 15 + (size_t)maxO * 3 + 8;
    const size_t max_jobs = (size_t)n * ((1 + (size_t)maxO * Lmax) * 2 + (size_t)maxO * 2);
    const size_t blob_bound = max_launches * 256 + max_jobs * std::max(sizeof(BlurJob), sizeof(ResizeJob)) +
                              octs.size() * sizeof(OctaveDesc) + (size_t)n * sizeof(ImageParams) +
                              (size_t)(n + 1) * sizeof(int) + 4096;
    const size_t blob_dev_off = round_up(C.off, 256);
    if (B.ext && blob_dev_off + blob_bound > B.ext->cap)
        return fail(SIFT_ERR_NOMEM, "batch needs " + std::to_string(blob_dev_off + blob_bound) +
                                        " bytes of scratch; the no-allocation entry point has " +
                                        std::to_string(B.ext->cap) + " per GPU");
    if (!B.ext && arena_limit && blob_dev_off + blob_bound > arena_limit &&
        (cand_cap > want_cap || kp_cap > want_cap + want_cap / 2)) {
        // capacities grown by an earlier, larger batch do not fit this lane's
        // share: lay this batch out with its own
        cand_cap = want_cap;
        kp_cap = want_cap + want_cap / 2;
        return run(B, outs);
    }
    if (!B.ext) {
        rc = ensure_arena(blob_dev_off + blob_bound);
        if (rc) return rc;
    }
    char* A = B.ext ? B.ext->base : arena;
    char* T = B.ext ? B.ext->base : tables.dev;
    auto AP = [&](size_t off) { return A + off; };

    std::vector<BlurJob> hj, hj2, hj16, hgj, hsj, vj, v8j, vnj, vgj, vsj, vs8j, cj;
    const std::pair<int, int> forced = blur_forced();
    int vn_rows = 0, hs_taps4 = 0, vs_taps4 = 0, vs_rows = 0, vs8_taps4 = 0, vs8_rows = 0;

    auto job_bytes = [](const auto& j, int kind) -> double {
        using J = std::decay_t<decltype(j)>;
        if constexpr (std::is_same_v<J, BlurJob>) return blur_job_bytes(j, kind);
        else if constexpr (std::is_same_v<J, LumaJob>) return luma_job_bytes(j);
        else return resize_job_bytes(j, kind);
    };
    auto job_ops = [](const auto& j, int kind) -> double {
        using J = std::decay_t<decltype(j)>;
        if constexpr (std::is_same_v<J, BlurJob>) return blur_job_ops(j, kind);
        else if constexpr (std::is_same_v<J, LumaJob>) return 0.0;
        else return resize_job_ops(j, kind);
    };
    auto push_launch = [&](int kind, auto& jobs, int max_pad, int variant) {
        if (jobs.empty()) return;
        int tiles = 0;
        double bytes = 0.0, ops = 0.0;
        for (auto& j : jobs) {
            j.tile_begin = tiles;
            tiles += j.tiles;
            bytes += job_bytes(j, kind);
            ops += job_ops(j, kind);
        }
        if (tiles == 0) return;
        size_t off = blob_take(jobs.size() * sizeof(jobs[0]));
        memcpy(blob.data() + off, jobs.data(), jobs.size() * sizeof(jobs[0]));
        launches.push_back({kind, off, (int)jobs.size(), tiles, max_pad, variant, bytes, ops});
    };

    auto make_blur_pair = [&](const uint8_t* src, int src_pitch, uint8_t* tmp, int tmp_pitch,
                              uint8_t* dst, int dst_pitch, int w, int h, const TapsRef& R,
                              float* dog, const uint8_t* prev, int dog_pitch) {
        if (w <= 0 || h <= 0) return;
        BlurJob J{};
        J.w = w;
        J.h = h;
        if (R.copy) {
            J.src = src;
            J.src_pitch = src_pitch;
            J.dst = dst;
            J.dst_pitch = dst_pitch;
            J.dog = dog;
            J.dog_pitch = dog_pitch;
            J.tiles_x = (w + 255) / 256;
            J.tiles = J.tiles_x * h;
            cj.push_back(J);
            return;
        }
        J.taps = (const float*)(T + R.off);
        J.ntaps = R.ntaps;
        J.ntaps_pad = R.npad;
        const std::pair<int, int> form = blur_choice(w, h, R.ntaps, R.npad, n >= kTuneManyJobs, forced);
        const int n4 = (R.ntaps + 3) & ~3, r = R.ntaps / 2;
        BlurJob H = J;
        H.src = src;
        H.src_pitch = src_pitch;
        H.dst = tmp;
        H.dst_pitch = tmp_pitch;
        blur_h_tiles(H, form.first);
        switch (form.first) {
            case BH_W8: hj.push_back(H); break;
            case BH_W16: hj16.push_back(H); break;
            case BH_W2: hj2.push_back(H); break;
            case BH_SMALL:
                hsj.push_back(H);
                hs_taps4 = std::max(hs_taps4, n4);
                break;
            default: hgj.push_back(H); break;
        }
        BlurJob V = J;
        V.src = tmp;
        V.src_pitch = tmp_pitch;
        V.dst = dst;
        V.dst_pitch = dst_pitch;
        V.dog = dog;
        V.prev = prev;
        V.dog_pitch = dog_pitch;
        blur_v_tiles(V, form.second);
        switch (form.second) {
            case BV_WIDE: vj.push_back(V); break;
            case BV_WIDE8: v8j.push_back(V); break;
            case BV_NARROW:
                vnj.push_back(V);
                vn_rows = std::max(vn_rows, std::min(h, 16 + 2 * r));
                break;
            case BV_SMALL:
                vsj.push_back(V);
                vs_taps4 = std::max(vs_taps4, n4);
                vs_rows = std::max(vs_rows, std::min(h, 4 + 2 * r));
                break;
            case BV_SMALL8:
                vs8j.push_back(V);
                vs8_taps4 = std::max(vs8_taps4, n4);
                vs8_rows = std::max(vs8_rows, std::min(h, 4 + 2 * r));
                break;
            default: vgj.push_back(V); break;
        }
    };

    auto flush_blur = [&]() {
        int maxpad = 0, maxpad2 = 0, maxpad16 = 0;
        for (auto& j : hj) maxpad = std::max(maxpad, j.ntaps_pad);
        for (auto& j : hj2) maxpad2 = std::max(maxpad2, j.ntaps_pad);
        for (auto& j : hj16) maxpad16 = std::max(maxpad16, j.ntaps_pad);
        push_launch(K_BLUR_H16, hj16, maxpad16, 16);
        push_launch(K_BLUR_H, hj, maxpad, 8);
        push_launch(K_BLUR_H2, hj2, maxpad2, 2);
        push_launch(K_BLUR_H_GENERIC, hgj, 0, 0);
        push_launch(K_BLUR_H_SMALL, hsj, hs_taps4, 0);
        push_launch(K_BLUR_V, vj, 0, 0);
        push_launch(K_BLUR_V8, v8j, 0, 0);
        push_launch(K_BLUR_V_NARROW, vnj, 0, vn_rows);
        push_launch(K_BLUR_V_GENERIC, vgj, 0, 0);
        push_launch(K_BLUR_V_SMALL, vsj, vs_taps4, vs_rows);
        push_launch(K_BLUR_V_SMALL8, vs8j, vs8_taps4, vs8_rows);
        push_launch(K_COPY, cj, 0, 0);
        hj.clear(); hj2.clear(); hgj.clear(); hsj.clear(); vj.clear(); vnj.clear(); vgj.clear(); vsj.clear();
        vs8j.clear();
        hj16.clear();
        v8j.clear();
        cj.clear();
        vn_rows = hs_taps4 = vs_taps4 = vs_rows = vs8_taps4 = vs8_rows = 0;
    };

    // colour -> grey (to_luma8 / the canvas rule)
    {
        std::vector<LumaJob> lj;
        for (int i = 0; i < n; i++) {
            const uint32_t fmt = B.format(i);
            if (fmt == SIFT_PIXEL_GRAY8) continue;
            const sift_image_t& im = B.images[i];
            LumaJob J{};
            J.src = B.device_input ? im.data : (const uint8_t*)AP(color_off[i]);
            J.src_pitch = color_pitch[i];
            J.dst = (uint8_t*)AP(in_off[i]);
            J.dst_pitch = in_pitch[i];
            J.w = (int)im.width;
            J.h = (int)im.height;
            J.format = (int)fmt;
            J.tiles_x = (J.w + kLumaTileW - 1) / kLumaTileW;
            J.tiles = J.tiles_x * J.h;
            lj.push_back(J);
        }
        push_launch(K_LUMA, lj, 0, 0);
    }

    // initial blur into octave 0 level 0
    for (int i = 0; i < n; i++) {
        const OctShape& s = shapes[i][0];
        const uint8_t* src = B.device_input && B.format(i) == SIFT_PIXEL_GRAY8
                                 ? B.images[i].data
                                 : (const uint8_t*)AP(in_off[i]);
        make_blur_pair(src, in_pitch[i], (uint8_t*)AP(htemp_off[i][0]), s.pitch,
                       (uint8_t*)AP(lev_off[i][0]), s.pitch, s.w, s.h, ptab[i]->initial, nullptr,
                       nullptr, 0);
    }
    flush_blur();
    // The levels as a wavefront: level s >= 1 of octave o of image i is blurred in
    // step o * S_i + s. Octave o + 1 needs only level S of octave o (its Lanczos3
    // base, made at the end of step (o + 1) * S_i), so its first levels share
    // launches with levels S + 1, S + 2 of octave o: a 5-octave, S = 3 frame takes
    // 18 blur steps instead of 26. Every level still reads the finished level
    // before it; only the order of independent launches changes.
    int steps = 0;
    for (int i = 0; i < n; i++) {
        const int S = (int)B.params[i].num_intervals;
        steps = std::max(steps, ((int)shapes[i].size() - 1) * S + S + 2);
    }
    for (int t = 1; t <= steps; t++) {
        for (int i = 0; i < n; i++) {
            const int S = (int)B.params[i].num_intervals, L = S + 3;
            for (int o = 0; o < (int)shapes[i].size(); o++) {
                const int s = t - o * S;
                if (s < 1 || s >= L) continue;
                const OctShape& sh = shapes[i][o];
                const size_t lsz = (size_t)sh.pitch * sh.h;
                uint8_t* base = (uint8_t*)AP(lev_off[i][o]);
                float* dog = B.want_dog ? (float*)AP(dog_off[i][o]) + (size_t)(s - 1) * sh.w * sh.h : nullptr;
                make_blur_pair(base + (s - 1) * lsz, sh.pitch, (uint8_t*)AP(htemp_off[i][o]), sh.pitch,
                               base + s * lsz, sh.pitch, sh.w, sh.h, ptab[i]->level[(size_t)o * L + s],
                               dog, base + (s - 1) * lsz, sh.w);
            }
        }
        flush_blur();
        // Lanczos3 to the next octave base (src/sift.rs:166-173), for every octave
        // whose level S was made in this step
        std::vector<ResizeJob> rv, rh, rf;
        const bool fuse_resize = resize_fusion_enabled();
        for (int i = 0; i < n; i++) {
            const int Si = (int)B.params[i].num_intervals;
            if (t % Si != 0) continue;
            const int o = t / Si - 1;
            if (o < 0 || (int)shapes[i].size() <= o + 1) continue;
            const OctShape& a = shapes[i][o];
            const OctShape& b = shapes[i][o + 1];
            if (a.w <= 0 || a.h <= 0 || b.w <= 0 || b.h <= 0) continue;
            const int S = (int)B.params[i].num_intervals;
            const LanczosRef& LV = lanczos_for(a.h, b.h);
            const LanczosRef& LH = lanczos_for(a.w, b.w);
            ResizeJob J{};
            J.src_u8 = (uint8_t*)AP(lev_off[i][o]) + (size_t)S * a.pitch * a.h;
            J.tmp = (float*)AP(rtmp_off[i]);
            J.dst = (uint8_t*)AP(lev_off[i][o + 1]);
            J.in_w = a.w;
            J.in_h = a.h;
            J.out_w = b.w;
            J.out_h = b.h;
            J.src_pitch = a.pitch;
            J.tmp_pitch = a.pitch;
            J.dst_pitch = b.pitch;
            // both passes in one kernel when every tile's input span fits its LDS
            bool fuse = fuse_resize;
            for (int ox0 = 0; ox0 < b.w && fuse; ox0 += 256) {
                const int oxl = std::min(ox0 + 256, b.w) - 1;
                const int lo4 = LH.left[ox0] & ~3;
                const int hi = LH.left[oxl] + LH.count[oxl];
                fuse = (int)round_up((size_t)(hi - lo4), 4) <= kResizeFusedSpan;
            }
            if (fuse) {
                ResizeJob F = J;
                F.left = (const int*)(T + LH.left_off);
                F.count = (const int*)(T + LH.count_off);
                F.weights = (const float*)(T + LH.w_off);
                F.maxtaps = LH.maxtaps;
                F.vleft = (const int*)(T + LV.left_off);
                F.vcount = (const int*)(T + LV.count_off);
                F.vweights = (const float*)(T + LV.w_off);
                F.vmaxtaps = LV.maxtaps;
                F.tiles_x = (b.w + 255) / 256;
                F.tiles = F.tiles_x * ((b.h + kResizeRows - 1) / kResizeRows);
                rf.push_back(F);
                continue;
            }
            ResizeJob V = J;
            V.left = (const int*)(T + LV.left_off);
            V.count = (const int*)(T + LV.count_off);
            V.weights = (const float*)(T + LV.w_off);
            V.maxtaps = LV.maxtaps;
            V.tiles_x = (a.w + 4 * 256 - 1) / (4 * 256);
            V.tiles = V.tiles_x * b.h;
            rv.push_back(V);
            ResizeJob H = J;
            H.left = (const int*)(T + LH.left_off);
            H.count = (const int*)(T + LH.count_off);
            H.weights = (const float*)(T + LH.w_off);
            H.maxtaps = LH.maxtaps;
            H.tiles_x = (b.w + 255) / 256;
            H.tiles = H.tiles_x * ((b.h + kResizeRows - 1) / kResizeRows);
            rh.push_back(H);
        }
        push_launch(K_RESIZE_FUSED, rf, 0, 0);
        push_launch(K_RESIZE_V, rv, 0, 0);
        push_launch(K_RESIZE_H, rh, 0, 0);
    }

    // octave descriptors and per-image params
    for (auto& d : octs) d.levels = (const uint8_t*)AP(lev_off[d.img][d.octave]);
    std::vector<ImageParams> iparams(n);
    for (int i = 0; i < n; i++) {
        const sift_params_t& p = B.params[i];
        iparams[i] = {p.sigma, p.contrast_threshold, p.edge_threshold, (int)p.num_intervals,
                      (int)p.image_border_width, (int)p.num_octaves};
    }
    const size_t oct_blob = blob_take(std::max<size_t>(octs.size(), 1) * sizeof(OctaveDesc));
    if (!octs.empty()) memcpy(blob.data() + oct_blob, octs.data(), octs.size() * sizeof(OctaveDesc));
    const size_t par_blob = blob_take(n * sizeof(ImageParams));
    memcpy(blob.data() + par_blob, iparams.data(), n * sizeof(ImageParams));
    const size_t isb_blob = blob_take((n + 1) * sizeof(int));
    memcpy(blob.data() + isb_blob, img_slot_begin.data(), (n + 1) * sizeof(int));

    if (blob.size() > blob_bound) return fail(SIFT_ERR_DEVICE, "internal: job blob bound exceeded");

    hipStream_t st = stream;
    if (B.ext) {
        // the table store and the blob through the staging allocated with the
        // context (the previous call synchronised before returning)
        const size_t tb = round_up(tables.host.size(), 256);
        if (tb + blob.size() + kExtSummaryBytes > ext_pinned_cap)
            return fail(SIFT_ERR_NOMEM, "weight tables and job records (" + std::to_string(tb + blob.size()) +
                                            " bytes) exceed the no-allocation staging");
        memcpy(ext_pinned, tables.host.data(), tables.host.size());
        memcpy(ext_pinned + tb, blob.data(), blob.size());
        if (!tables.host.empty())
            HIP_TRY(hipMemcpyAsync(B.ext->base, ext_pinned, tables.host.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(A + blob_dev_off, ext_pinned + tb, blob.size(), hipMemcpyHostToDevice, st));
    }
    // staging: inputs + blob in pinned memory (inputs that are pinned already
    // are copied from where they are)
    size_t in_bytes = 0;
    if (!B.device_input && !B.host_pinned)
        for (int i = 0; i < n; i++)
            in_bytes += (size_t)B.images[i].width * B.images[i].height * pixel_bytes(B.format(i));
    if (!B.ext) {
        rc = ensure_pinned(round_up(blob.size(), 256) + in_bytes + 256);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(stream));  // pinned staging is reused across calls
        memcpy(pinned, blob.data(), blob.size());
        HIP_TRY(hipMemcpyAsync(A + blob_dev_off, pinned, blob.size(), hipMemcpyHostToDevice, st));
    }
    if (!B.device_input) {
        size_t po = round_up(blob.size(), 256);
        for (int i = 0; i < n; i++) {
            const sift_image_t& im = B.images[i];
            const uint32_t fmt = B.format(i);
            const size_t row = (size_t)im.width * pixel_bytes(fmt);
            const size_t stride = im.stride ? im.stride : row;
            const uint8_t* src = im.data;
            size_t spitch = stride;
            if (!B.host_pinned) {
                for (uint32_t y = 0; y < im.height; y++)
                    memcpy(pinned + po + (size_t)y * row, im.data + (size_t)y * stride, row);
                src = (const uint8_t*)pinned + po;
                spitch = row;
                po += row * im.height;
            }
            if (fmt == SIFT_PIXEL_GRAY8)
                HIP_TRY(hipMemcpy2DAsync(A + in_off[i], in_pitch[i], src, spitch, row, im.height,
                                         hipMemcpyHostToDevice, st));
            else
                HIP_TRY(hipMemcpy2DAsync(A + color_off[i], color_pitch[i], src, spitch, row,
                                         im.height, hipMemcpyHostToDevice, st));
        }
    }
    char* D = A + blob_dev_off;
    const OctaveDesc* d_oct = (const OctaveDesc*)(D + oct_blob);
    const ImageParams* d_par = (const ImageParams*)(D + par_blob);
    unsigned* d_counts = (unsigned*)AP(counts_off);
    unsigned* d_offs = (unsigned*)AP(offs_off);
    Cand* d_cand = (Cand*)AP(cand_off);
    Keypoint* d_ref = (Keypoint*)AP(ref_off);
    int* d_valid = (int*)AP(valid_off);
    unsigned* d_npk = (unsigned*)AP(npk_off);
    float* d_ang = (float*)AP(ang_off);
    unsigned* d_koff = (unsigned*)AP(koff_off);
    Keypoint* d_kps = B.ext ? B.ext->kps : (Keypoint*)AP(kps_off);
    float* d_desc = B.ext ? B.ext->desc : (float*)AP(desc_off);
    unsigned* d_sum = (unsigned*)AP(summary_off);
    unsigned* d_total = d_offs + nslots;

    // Every launch of the batch, in order. Kernel arguments are arena addresses
    // and scalars of this layout; the per-call data (images, taps) is read from
    // the job blob uploaded above.
    auto issue = [&](hipStream_t st) -> int {
        // --- pyramid ---
        for (const LaunchRec& L : launches) {
            hipEvent_t ev = nullptr;
            prof_begin(L.kind, L.bytes, &ev);
            switch (L.kind) {
                case K_BLUR_H:
                case K_BLUR_H2:
                case K_BLUR_H16:
                    launch_blur_h((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, L.max_pad, L.variant, st);
                    break;
                case K_BLUR_V: launch_blur_v((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, 0, st); break;
                case K_BLUR_V8: launch_blur_v8((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, st); break;
                case K_BLUR_V_NARROW:
                    launch_blur_v((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, L.variant, st);
                    break;
                case K_BLUR_V_GENERIC:
                    launch_blur_v_generic((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, st);
                    break;
                case K_COPY: launch_copy_level((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, st); break;
                case K_BLUR_H_GENERIC:
                    launch_blur_h_generic((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, st);
                    break;
                case K_BLUR_H_SMALL:
                    launch_blur_h_small((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, L.max_pad, st);
                    break;
                case K_BLUR_V_SMALL:
                    launch_blur_v_small((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, L.max_pad,
                                        L.variant, st);
                    break;
                case K_BLUR_V_SMALL8:
                    launch_blur_v_small8((const BlurJob*)(D + L.job_off), L.njobs, L.tiles, L.max_pad,
                                         L.variant, st);
                    break;
                case K_LUMA: launch_luma((const LumaJob*)(D + L.job_off), L.njobs, L.tiles, st); break;
                case K_RESIZE_V: launch_resize_v((const ResizeJob*)(D + L.job_off), L.njobs, L.tiles, st); break;
                case K_RESIZE_H: launch_resize_h((const ResizeJob*)(D + L.job_off), L.njobs, L.tiles, st); break;
                case K_RESIZE_FUSED:
                    launch_resize_fused((const ResizeJob*)(D + L.job_off), L.njobs, L.tiles, st);
                    break;
            }
            prof_end(L.kind, L.bytes, ev, L.ops);
        }
        HIP_TRY(hipGetLastError());
        if (B.pyramid_only) return SIFT_OK;
        // --- detection ---
        double ext_bytes = 0.0;
        for (const auto& d : octs)
            ext_bytes += (double)d.nlevels * d.w * d.h + (double)(d.nlevels - 3) * d.h * d.nwords * 8.0;
        hipEvent_t ev = nullptr;
        prof_begin(K_EXTREMA, ext_bytes, &ev);
        if (tile_ext)
            launch_extrema_tile(S0, strip_rows, d_oct, noct, total_tiles, d_par,
                                (unsigned long long*)AP(bits_off),
                                d_counts, nslots, st);
        else
            launch_extrema(d_oct, noct, total_rows, d_par, (unsigned long long*)AP(bits_off), d_counts,
                           nslots, st);
        prof_end(K_EXTREMA, ext_bytes, ev);
        prof_begin(K_SCAN, 0.0, &ev);
        launch_scan_dev(d_counts, nslots, nullptr, d_offs, (unsigned*)AP(scantmp1_off), st);
        prof_end(K_SCAN, 0.0, ev);
        prof_begin(K_COMPACT, 0.0, &ev);
        launch_compact(d_oct, noct, nslots, (const unsigned long long*)AP(bits_off), d_counts, d_offs,
                       d_cand, cap, st);
        prof_end(K_COMPACT, 0.0, ev);
        prof_begin(K_REFINE, 0.0, &ev);
        launch_refine(d_oct, d_par, d_cand, d_total, cap, d_ref, d_valid,
                      B.want_stages ? (float*)AP(candresp_off) : nullptr, st);
        prof_end(K_REFINE, 0.0, ev);
        prof_begin(K_ORIENT, 0.0, &ev);
        launch_orient(d_oct, d_par, d_cand, d_total, cap, d_ref, d_valid, d_npk, d_ang,
                      d_counts + nslots + 2, st);  // kNumXcd queue counters
        prof_end(K_ORIENT, 0.0, ev);
        prof_begin(K_SCAN, 0.0, &ev);
        launch_scan_dev(d_npk, cap, d_total, d_koff, (unsigned*)AP(scantmp2_off), st);
        prof_end(K_SCAN, 0.0, ev);
        prof_begin(K_DESCRIPTOR, 0.0, &ev);
        launch_descriptor(d_oct, d_par, d_cand, d_total, cap, d_ref, d_npk, d_ang, d_koff, kcap, d_kps,
                          d_desc, d_counts + nslots, (int*)AP(wide_off), d_counts + nslots + 1,
                          (int*)AP(mid_off), d_counts + nslots + 2 + kNumXcd, st);
        prof_end(K_DESCRIPTOR, 0.0, ev);
        prof_begin(K_FINALIZE, 0.0, &ev);
        launch_finalize(d_offs, (const int*)(D + isb_blob), n, d_koff, d_total, cap, d_sum,
                        B.ext ? B.ext->offsets : nullptr, st);
        prof_end(K_FINALIZE, 0.0, ev);
        if (B.desc_u8) {
            prof_begin(K_DESC_U8, 0.0, &ev);
            launch_desc_u8(d_desc, d_sum + 1, kcap, (uint8_t*)AP(descu8_off), st);
            prof_end(K_DESC_U8, 0.0, ev);
        }
        HIP_TRY(hipGetLastError());
        return SIFT_OK;
    };

    // (a capture and its instantiation allocate: external-memory calls launch eagerly)
    if (B.ext || !use_graphs || profiling || !graphs_enabled_by_env()) {
        rc = issue(st);
        if (rc) return rc;
    } else {
        // everything the launch sequence depends on
        std::vector<uint64_t> key = {(uint64_t)(uintptr_t)A, blob_dev_off, (uint64_t)B.pyramid_only,
                                     (uint64_t)B.want_stages, (uint64_t)tile_ext, (uint64_t)S0,
                                     (uint64_t)strip_rows,
                                     (uint64_t)noct, (uint64_t)total_tiles, (uint64_t)total_rows,
                                     (uint64_t)nslots, (uint64_t)cap, (uint64_t)kcap, (uint64_t)n,
                                     oct_blob, par_blob, isb_blob, bits_off, counts_off, offs_off,
                                     scantmp1_off, cand_off, candresp_off, ref_off, valid_off, npk_off,
                                     ang_off, koff_off, scantmp2_off, kps_off, desc_off, summary_off,
                                     wide_off, mid_off, (uint64_t)B.desc_u8, descu8_off};
        for (const LaunchRec& L : launches) {
            key.push_back((uint64_t)L.kind << 32 | (uint32_t)L.njobs);
            key.push_back(L.job_off);
            key.push_back((uint64_t)(uint32_t)L.tiles << 32 | (uint32_t)L.max_pad);
            key.push_back((uint64_t)(uint32_t)L.variant);
        }
        GraphEntry* hit = nullptr;
        for (GraphEntry& e : graphs)
            if (e.key == key) {
                hit = &e;
                break;
            }
        // a layout is captured the second time it is seen: one-off layouts (mixed
        // serving batches) run eagerly instead of paying for a capture
        bool seen = hit != nullptr;
        if (!hit) {
            uint64_t hsh = 1469598103934665603ull;  // FNV-1a over the key words
            for (uint64_t v : key) hsh = (hsh ^ v) * 1099511628211ull;
            for (uint64_t v : seen_once)
                if (v == hsh) seen = true;
            if (!seen) {
                if (seen_once.size() >= 4 * kMaxGraphs) seen_once.erase(seen_once.begin());
                seen_once.push_back(hsh);
            }
        }
        if (!seen) {
            rc = issue(st);
            if (rc) return rc;
        } else {
            if (!hit) {
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                const int irc = issue(st);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(st, &g);
                if (irc || ce != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    if (irc) return irc;
                    return fail(SIFT_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(ce));
                }
                hipGraphExec_t ex = nullptr;
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess)
                    return fail(SIFT_ERR_DEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(ie));
                if (graphs.size() >= kMaxGraphs) {  // evict the least recently used
                    size_t lru = 0;
                    for (size_t i = 1; i < graphs.size(); i++)
                        if (graphs[i].used < graphs[lru].used) lru = i;
                    (void)hipGraphExecDestroy(graphs[lru].exec);
                    graphs.erase(graphs.begin() + lru);
                }
                graphs.push_back(GraphEntry{std::move(key), ex, 0});
                hit = &graphs.back();
            }
            hit->used = ++graph_clock;
            HIP_TRY(hipGraphLaunch(hit->exec, st));
        }
    }

    if (outs) {
        outs->d_kps = d_kps;
        outs->d_desc = d_desc;
        outs->d_desc_u8 = B.desc_u8 ? (uint8_t*)AP(descu8_off) : nullptr;
        outs->d_summary = d_sum;
        outs->d_cand = d_cand;
        outs->d_cand_resp = (float*)AP(candresp_off);
        outs->d_ref = d_ref;
        outs->d_valid = d_valid;
        outs->d_octs = d_oct;
        outs->octs = octs;
        outs->cap = cap;
        outs->kcap = kcap;
        outs->lev_off = lev_off;
        outs->dog_off = dog_off;
        outs->arena = A;
    }
    return SIFT_OK;
}

}  // namespace sh

// ===========================================================================
// Tuning table
// ===========================================================================
namespace {

// The table file: SIFT_HIP_TUNE_TABLE (a path; "0" = none), else blur_tune.json
// next to this library.
std::string tune_table_path() {
    const char* e = getenv("SIFT_HIP_TUNE_TABLE");
    if (e) return (e[0] == '0' && e[1] == 0) ? std::string() : std::string(e);
    Dl_info info{};
    if (!dladdr((const void*)&tune_table_path, &info) || !info.dli_fname) return std::string();
    std::string lib(info.dli_fname);
    const size_t slash = lib.rfind('/');
    return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/blur_tune.json";
}

// Reads the "entries" of the table: arrays [w, h, taps, many, "<h form>", "<v form>"]
// (every innermost array of the file with those six fields). Unknown forms and
// malformed arrays are skipped.
void load_tune_table(std::map<sh::TuneKey, std::pair<int, int>>& out) {
    const std::string path = tune_table_path();
    if (path.empty()) return;
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return;
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    fclose(f);
    const size_t at = text.find("\"entries\"");
    if (at == std::string::npos) return;
    size_t i = at;
    while (i < text.size()) {
        const size_t open = text.find('[', i);
        if (open == std::string::npos) break;
        const size_t close = text.find(']', open + 1);
        if (close == std::string::npos) break;
        const size_t inner = text.find('[', open + 1);
        if (inner != std::string::npos && inner < close) {  // not innermost: descend
            i = inner;
            continue;
        }
        std::vector<std::string> fields;
        std::string cur;
        bool quoted = false;
        for (size_t p = open + 1; p < close; p++) {
            const char c = text[p];
            if (c == '"') quoted = !quoted;
            else if (c == ',' && !quoted) {
                fields.push_back(cur);
                cur.clear();
            } else if (quoted || (c != ' ' && c != '\n' && c != '\t' && c != '\r')) cur += c;
        }
        fields.push_back(cur);
        if (fields.size() == 6) {
            const int hf = sh::blur_form_id(fields[4].c_str(), false);
            const int vf = sh::blur_form_id(fields[5].c_str(), true);
            const int w = atoi(fields[0].c_str()), h = atoi(fields[1].c_str());
            const int nt = atoi(fields[2].c_str()), many = atoi(fields[3].c_str()) ? 1 : 0;
            if (hf >= 0 && vf >= 0 && w > 0 && h > 0 && nt > 0) out[sh::TuneKey{w, h, nt, many}] = {hf, vf};
        }
        i = close + 1;
    }
}

}  // namespace

// ===========================================================================
// Scratch of the no-allocation entry point
// ===========================================================================
// sift_hip_detect_batch_device_noalloc works in this array instead of a
// context's arena. Each GPU holds its own copy, at its own address, from the
// library's first use on it until the process ends. In host code `scratch`
// only names the symbol: its device address on each GPU is looked up once,
// with that GPU current, when the first context on it is created, and only
// that address (g_scratch) is handed to kernels and copies.
static __device__ __align__(256) char scratch[SIFT_HIP_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxScratchDevices] = {};
hipError_t g_scratch_err[kMaxScratchDevices] = {};

// Looks up this device's scratch address (the device is current). A failed
// lookup does not fail context creation: only the no-allocation entry point
// needs the address, and it reports the error.
void scratch_setup(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_scratch[device] = (char*)p;
    else g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
}

char* scratch_of(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

int check_params(const sift_params_t* p) {
    if (!p) return fail(SIFT_ERR_INVALID, "params is null");
    if (p->num_intervals < 1 || p->num_intervals > 13) return fail(SIFT_ERR_INVALID, "num_intervals must be in [1, 13]");
    if (p->num_octaves < 1 || p->num_octaves > 16) return fail(SIFT_ERR_INVALID, "num_octaves must be in [1, 16]");
    if (!(p->sigma > 0.0f)) return fail(SIFT_ERR_INVALID, "sigma must be > 0");
    return SIFT_OK;
}

int check_image(const sift_image_t* im) {
    if (!im || !im->data) return fail(SIFT_ERR_INVALID, "image or image data is null");
    if (im->width == 0 || im->height == 0) return fail(SIFT_ERR_INVALID, "image has zero size");
    if (im->width > 65536 || im->height > 65536) return fail(SIFT_ERR_INVALID, "image too large");
    if (im->stride && im->stride < im->width) return fail(SIFT_ERR_INVALID, "stride < width");
    return SIFT_OK;
}

int check_color_image(const sift_color_image_t* im) {
    if (!im || !im->data) return fail(SIFT_ERR_INVALID, "image or image data is null");
    const int bpp = sh::pixel_bytes(im->format);
    if (bpp == 0) return fail(SIFT_ERR_INVALID, "unknown pixel format");
    if (im->width == 0 || im->height == 0) return fail(SIFT_ERR_INVALID, "image has zero size");
    if (im->width > 65536 || im->height > 65536) return fail(SIFT_ERR_INVALID, "image too large");
    if (im->stride && im->stride < (size_t)im->width * bpp)
        return fail(SIFT_ERR_INVALID, "stride < width * bytes per pixel");
    return SIFT_OK;
}

// Copies the results out: f32 descriptors into `out`, or (when out8 is set)
// the u8 descriptors into `out8`, with the keypoint convention kp_conv.
// The f32 results of a finished run (summary `sum` read): each image's
// keypoints and descriptors in one pinned pool block, [keypoints | pad |
// descriptors], filled by two DMAs from the device; images the pool cannot
// serve get malloc'd buffers, filled through the context's pinned staging.
int collect_pinned(sh::Context* C, int n, sh::RunOutputs& R, const std::vector<unsigned>& sum,
                   sift_result_t* out) {
    constexpr size_t kDesc = 128 * sizeof(float);
    std::vector<char> pooled(n, 0);
    size_t staged = 0;  // keypoints of the images that go through the staging
    for (int i = 0; i < n; i++) {
        const size_t c = sum[3 + i] - sum[2 + i];
        out[i] = sift_result_t{nullptr, nullptr, c};
        if (!c) continue;
        const size_t kb = (c * sizeof(sift_keypoint_t) + 255) / 256 * 256;
        char* blk = (char*)sh::pinned_alloc(kb + c * kDesc);
        if (blk) {
            out[i].keypoints = (sift_keypoint_t*)blk;
            out[i].descriptors = (float*)(blk + kb);
            pooled[i] = 1;
            continue;
        }
        out[i].keypoints = (sift_keypoint_t*)malloc(c * sizeof(sift_keypoint_t));
        out[i].descriptors = (float*)malloc(c * kDesc);
        if (!out[i].keypoints || !out[i].descriptors) return fail(SIFT_ERR_NOMEM, "host allocation failed");
        staged += c;
    }
    // staging slots of the malloc'd images, in image order
    sh::Keypoint* skps = nullptr;
    char* sdesc = nullptr;
    if (staged) {
        const size_t kbytes = staged * sizeof(sh::Keypoint);
        int rc = C->ensure_pinned_out(kbytes + staged * kDesc + 256);
        if (rc) return rc;
        skps = (sh::Keypoint*)C->pinned_out;
        sdesc = C->pinned_out + (kbytes + 255) / 256 * 256;
    }
    size_t so = 0;
    for (int i = 0; i < n; i++) {
        const size_t b = sum[2 + i], c = out[i].count;
        if (!c) continue;
        void* kd = pooled[i] ? (void*)out[i].keypoints : (void*)(skps + so);
        void* dd = pooled[i] ? (void*)out[i].descriptors : (void*)(sdesc + so * kDesc);
        hipError_t e = hipMemcpyAsync(kd, R.d_kps + b, c * sizeof(sh::Keypoint), hipMemcpyDeviceToHost, C->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dd, R.d_desc + b * 128, c * kDesc, hipMemcpyDeviceToHost, C->stream);
        if (e != hipSuccess) {
            // copies already queued must land before the caller frees (and the
            // pool reuses) the blocks they write
            (void)hipStreamSynchronize(C->stream);
            return fail(SIFT_ERR_DEVICE, std::string("hipMemcpyAsync (results) failed: ") + hipGetErrorString(e));
        }
        if (!pooled[i]) so += c;
    }
    HIP_TRY(hipStreamSynchronize(C->stream));
    so = 0;
    for (int i = 0; i < n; i++) {
        const size_t c = out[i].count;
        if (!c || pooled[i]) continue;
        memcpy(out[i].keypoints, skps + so, c * sizeof(sift_keypoint_t));
        memcpy(out[i].descriptors, sdesc + so * kDesc, c * kDesc);
        so += c;
    }
    return SIFT_OK;
}

int collect(sh::Context* C, int n, sh::RunOutputs& R, sift_result_t* out, bool* retry,
            sift_result_u8_t* out8 = nullptr, uint32_t kp_conv = SIFT_KEYPOINTS_CPU) {
    std::vector<unsigned> sum(n + 3);
    HIP_TRY(hipMemcpyAsync(sum.data(), R.d_summary, (n + 3) * 4, hipMemcpyDeviceToHost, C->stream));
    HIP_TRY(hipStreamSynchronize(C->stream));
    const size_t ncand = sum[0], nkp = sum[1];
    *retry = false;
    if (ncand > (size_t)R.cap || nkp > (size_t)R.kcap) {
        C->cand_cap = std::max(C->cand_cap, ncand + ncand / 4 + 1024);
        C->kp_cap = std::max(C->kp_cap, std::max(C->cand_cap + C->cand_cap / 2, nkp + nkp / 4 + 1024));
        *retry = true;
        return SIFT_OK;
    }
    if (!out && !out8) return SIFT_OK;
    if (!out8) return collect_pinned(C, n, R, sum, out);
    const size_t dsz = out8 ? 128 : 128 * 4;
    const size_t kbytes = nkp * sizeof(sh::Keypoint), dbytes = nkp * dsz;
    int rc = C->ensure_pinned_out(kbytes + dbytes + 256);
    if (rc) return rc;
    sh::Keypoint* kps = (sh::Keypoint*)C->pinned_out;
    char* desc = C->pinned_out + ((kbytes + 255) / 256) * 256;
    if (nkp) {
        HIP_TRY(hipMemcpyAsync((void*)kps, R.d_kps, kbytes, hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipMemcpyAsync((void*)desc, out8 ? (const void*)R.d_desc_u8 : (const void*)R.d_desc,
                               dbytes, hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
    }
    if (kp_conv == SIFT_KEYPOINTS_WGPU_V2)
        for (size_t k = 0; k < nkp; k++) {
            kps[k].size = kps[k].size * 2.0f;
            kps[k].response = 0.0f;
            kps[k].octave = 0;
            kps[k].layer = 0;
        }
    if (out8) {
        for (int i = 0; i < n; i++) {
            size_t b = sum[2 + i], c = sum[3 + i] - b;
            out8[i] = sift_result_u8_t{nullptr, nullptr, c};
            if (c) {
                out8[i].keypoints = (sift_keypoint_t*)malloc(c * sizeof(sift_keypoint_t));
                out8[i].descriptors = (uint8_t*)malloc(c * 128);
                if (!out8[i].keypoints || !out8[i].descriptors) return fail(SIFT_ERR_NOMEM, "host allocation failed");
                memcpy(out8[i].keypoints, kps + b, c * sizeof(sift_keypoint_t));
                memcpy(out8[i].descriptors, desc + b * 128, c * 128);
            }
        }
        return SIFT_OK;
    }
    for (int i = 0; i < n; i++) {
        size_t b = sum[2 + i], e = sum[3 + i];
        size_t c = e - b;
        out[i].count = c;
        out[i].keypoints = nullptr;
        out[i].descriptors = nullptr;
        if (c) {
            out[i].keypoints = (sift_keypoint_t*)malloc(c * sizeof(sift_keypoint_t));
            out[i].descriptors = (float*)malloc(c * 128 * sizeof(float));
            if (!out[i].keypoints || !out[i].descriptors) return fail(SIFT_ERR_NOMEM, "host allocation failed");
            memcpy(out[i].keypoints, kps + b, c * sizeof(sift_keypoint_t));
            memcpy(out[i].descriptors, desc + b * 128 * 4, c * 128 * sizeof(float));
        }
    }
    return SIFT_OK;
}

}  // namespace

namespace sh {
// Runs a host-input batch and copies the results out, growing the candidate
// and keypoint capacities and re-running when a count overflowed them.
int run_and_collect(Context* C, const Batch& B, sift_result_t* out, sift_result_u8_t* out8,
                    uint32_t kp_conv) {
    const int n = (int)B.images.size();
    for (int i = 0; i < n; i++) {
        if (out) out[i] = sift_result_t{nullptr, nullptr, 0};
        if (out8) out8[i] = sift_result_u8_t{nullptr, nullptr, 0};
    }
    // on an error, entries collect() already filled are freed: the caller gets
    // empty results with the error code
    auto drop = [&](int rc) {
        for (int i = 0; i < n; i++) {
            if (out) sift_result_free(&out[i]);
            if (out8) sift_result_u8_free(&out8[i]);
        }
        return rc;
    };
    for (int attempt = 0; attempt < 4; attempt++) {
        RunOutputs R;
        int rc = C->run(B, &R);
        if (rc) return drop(rc);
        bool retry = false;
        rc = collect(C, n, R, out, &retry, out8, kp_conv);
        if (rc) return drop(rc);
        if (!retry) return SIFT_OK;
    }
    return drop(fail(SIFT_ERR_NOMEM, "keypoint capacity did not converge"));
}
}  // namespace sh

extern "C" {

void sift_params_default(sift_params_t* out) {
    if (!out) return;
    sift_params_new(out, 1.6f, 4, 3, 0.5f, 0.04f, 10.0f);
}

void sift_params_new(sift_params_t* out, float sigma, uint32_t num_octaves, uint32_t num_intervals,
                     float assumed_blur, float contrast_threshold, float edge_threshold) {
    if (!out) return;
    out->sigma = sigma;
    out->num_octaves = num_octaves;
    out->num_intervals = num_intervals;
    out->assumed_blur = assumed_blur;
    out->contrast_threshold = contrast_threshold / (float)num_intervals;
    out->edge_threshold = edge_threshold;
    out->image_border_width = 5;
}

const char* sift_hip_last_error(void) { return g_last_error.c_str(); }

int sift_hip_create(const sift_params_t* params, int device, sift_hip_ctx** out) {
    if (!out) return fail(SIFT_ERR_INVALID, "out is null");
    *out = nullptr;
    int rc = check_params(params);
    if (rc) return rc;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(SIFT_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(SIFT_ERR_DEVICE, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    sh::Context* C = new (std::nothrow) sh::Context();
    if (!C) return fail(SIFT_ERR_NOMEM, "context allocation failed");
    C->device = device;
    C->params = *params;
    load_tune_table(C->tune);
    scratch_setup(device);
    // staging of the no-allocation entry point (host memory; a failure here only
    // disables that entry point)
    if (hipHostMalloc((void**)&C->ext_pinned, kExtStagingBytes, hipHostMallocDefault) == hipSuccess)
        C->ext_pinned_cap = kExtStagingBytes;
    else
        C->ext_pinned = nullptr;
    e = hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete C;
        return fail(SIFT_ERR_DEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    *out = reinterpret_cast<sift_hip_ctx*>(C);
    return SIFT_OK;
}

void sift_hip_destroy(sift_hip_ctx* ctx) {
    if (!ctx) return;
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    sift_hip_server_stop(ctx);
    C->release();
    delete C;
}

int sift_hip_detect_batch(sift_hip_ctx* ctx, const sift_image_t* images, const sift_params_t* params,
                          size_t n, sift_result_t* out) {
    if (!ctx || !images || !out || n == 0) return fail(SIFT_ERR_INVALID, "null argument or empty batch");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    for (size_t i = 0; i < n; i++) {
        out[i] = sift_result_t{nullptr, nullptr, 0};
        int rc = check_image(&images[i]);
        if (rc) return rc;
        if (params && (rc = check_params(&params[i]))) return rc;
    }
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    sh::Batch B;
    B.images.assign(images, images + n);
    B.params.assign(n, C->params);
    if (params) B.params.assign(params, params + n);
    return sh::run_and_collect(C, B, out);
}

int sift_hip_detect(sift_hip_ctx* ctx, const sift_image_t* image, sift_result_t* out) {
    return sift_hip_detect_batch(ctx, image, nullptr, 1, out);
}

namespace {
// Checks colour images and params and fills the batch (images, formats, params).
int color_batch(sh::Context* C, const sift_color_image_t* images, const sift_params_t* params,
                size_t n, sh::Batch* B) {
    B->images.resize(n);
    B->formats.resize(n);
    B->params.assign(n, C->params);
    for (size_t i = 0; i < n; i++) {
        int rc = check_color_image(&images[i]);
        if (rc) return rc;
        if (params && (rc = check_params(&params[i]))) return rc;
        const sift_color_image_t& im = images[i];
        B->images[i] = sift_image_t{im.data, im.width, im.height, im.stride};
        B->formats[i] = im.format;
        if (params) B->params[i] = params[i];
    }
    return SIFT_OK;
}
}  // namespace

int sift_hip_detect_batch_color(sift_hip_ctx* ctx, const sift_color_image_t* images,
                                const sift_params_t* params, size_t n, sift_result_t* out) {
    if (!ctx || !images || !out || n == 0) return fail(SIFT_ERR_INVALID, "null argument or empty batch");
    for (size_t i = 0; i < n; i++) out[i] = sift_result_t{nullptr, nullptr, 0};
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    sh::Batch B;
    int rc = color_batch(C, images, params, n, &B);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    return sh::run_and_collect(C, B, out);
}

int sift_hip_detect_color(sift_hip_ctx* ctx, const sift_color_image_t* image, sift_result_t* out) {
    return sift_hip_detect_batch_color(ctx, image, nullptr, 1, out);
}

int sift_hip_detect_batch_u8(sift_hip_ctx* ctx, const sift_color_image_t* images,
                             const sift_params_t* params, size_t n, uint32_t keypoint_convention,
                             sift_result_u8_t* out) {
    if (!ctx || !images || !out || n == 0) return fail(SIFT_ERR_INVALID, "null argument or empty batch");
    for (size_t i = 0; i < n; i++) out[i] = sift_result_u8_t{nullptr, nullptr, 0};
    if (keypoint_convention != SIFT_KEYPOINTS_CPU && keypoint_convention != SIFT_KEYPOINTS_WGPU_V2)
        return fail(SIFT_ERR_INVALID, "unknown keypoint convention");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    sh::Batch B;
    int rc = color_batch(C, images, params, n, &B);
    if (rc) return rc;
    B.desc_u8 = true;
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    return sh::run_and_collect(C, B, nullptr, out, keypoint_convention);
}

void sift_result_u8_free(sift_result_u8_t* r) {
    if (!r) return;
    free(r->keypoints);
    free(r->descriptors);
    r->keypoints = nullptr;
    r->descriptors = nullptr;
    r->count = 0;
}

int sift_hip_to_luma8(sift_hip_ctx* ctx, const sift_color_image_t* image, uint8_t* out) {
    if (!ctx || !out) return fail(SIFT_ERR_INVALID, "null argument");
    int rc = check_color_image(image);
    if (rc) return rc;
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    const uint32_t w = image->width, h = image->height;
    const size_t row = (size_t)w * sh::pixel_bytes(image->format);
    const size_t stride = image->stride ? image->stride : row;
    const size_t cpitch = round_up(row, 64), gpitch = round_up(w, 64);
    sh::Carver L;
    const size_t job_off = L.take(sizeof(sh::LumaJob));
    const size_t color_off = L.take(cpitch * h);
    const size_t gray_off = L.take(gpitch * h);
    if ((rc = C->ensure_arena(L.off))) return rc;
    if ((rc = C->ensure_pinned(256 + row * h))) return rc;
    if ((rc = C->ensure_pinned_out((size_t)w * h))) return rc;
    HIP_TRY(hipStreamSynchronize(C->stream));  // pinned staging is reused across calls
    sh::LumaJob J{};
    J.src = (const uint8_t*)C->arena + color_off;
    J.dst = (uint8_t*)C->arena + gray_off;
    J.w = (int)w;
    J.h = (int)h;
    J.src_pitch = (int)cpitch;
    J.dst_pitch = (int)gpitch;
    J.format = (int)image->format;
    J.tiles_x = (J.w + sh::kLumaTileW - 1) / sh::kLumaTileW;
    J.tiles = J.tiles_x * J.h;
    J.tile_begin = 0;
    memcpy(C->pinned, &J, sizeof(J));
    for (uint32_t y = 0; y < h; y++) memcpy(C->pinned + 256 + y * row, image->data + y * stride, row);
    HIP_TRY(hipMemcpyAsync(C->arena + job_off, C->pinned, sizeof(J), hipMemcpyHostToDevice, C->stream));
    HIP_TRY(hipMemcpy2DAsync(C->arena + color_off, cpitch, C->pinned + 256, row, row, h,
                             hipMemcpyHostToDevice, C->stream));
    sh::launch_luma((const sh::LumaJob*)(C->arena + job_off), 1, J.tiles, C->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy2DAsync(C->pinned_out, w, C->arena + gray_off, gpitch, w, h, hipMemcpyDeviceToHost,
                             C->stream));
    HIP_TRY(hipStreamSynchronize(C->stream));
    memcpy(out, C->pinned_out, (size_t)w * h);
    return SIFT_OK;
}

int sift_hip_detect_batch_device(sift_hip_ctx* ctx, const sift_image_t* images,
                                 const sift_params_t* params, size_t n, size_t* out_counts,
                                 const sift_keypoint_t** d_keypoints, const float** d_descriptors,
                                 const uint32_t** d_offsets) {
    if (!ctx || !images || n == 0) return fail(SIFT_ERR_INVALID, "null argument or empty batch");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    for (size_t i = 0; i < n; i++) {
        int rc = check_image(&images[i]);
        if (rc) return rc;
        if (params && (rc = check_params(&params[i]))) return rc;
    }
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    sh::Batch B;
    B.images.assign(images, images + n);
    B.params.assign(n, C->params);
    if (params) B.params.assign(params, params + n);
    B.device_input = true;
    for (int attempt = 0; attempt < 4; attempt++) {
        sh::RunOutputs R;
        int rc = C->run(B, &R);
        if (rc) return rc;
        std::vector<unsigned> sum(n + 3);
        HIP_TRY(hipMemcpyAsync(sum.data(), R.d_summary, (n + 3) * 4, hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
        if (sum[0] > (unsigned)R.cap || sum[1] > (unsigned)R.kcap) {
            C->cand_cap = std::max(C->cand_cap, (size_t)sum[0] + sum[0] / 4 + 1024);
            C->kp_cap = std::max(C->kp_cap, std::max(C->cand_cap + C->cand_cap / 2, (size_t)sum[1] + sum[1] / 4 + 1024));
            continue;
        }
        if (out_counts)
            for (size_t i = 0; i < n; i++) out_counts[i] = sum[3 + i] - sum[2 + i];
        if (d_keypoints) *d_keypoints = (const sift_keypoint_t*)R.d_kps;
        if (d_descriptors) *d_descriptors = R.d_desc;
        if (d_offsets) *d_offsets = R.d_summary + 2;
        return SIFT_OK;
    }
    return fail(SIFT_ERR_NOMEM, "keypoint capacity did not converge");
}

int sift_hip_detect_batch_device_noalloc(sift_hip_ctx* ctx, const sift_image_t* images,
                                         const sift_params_t* params, size_t n,
                                         sift_keypoint_t* d_keypoints, float* d_descriptors,
                                         size_t keypoint_capacity, uint32_t* d_offsets,
                                         size_t* out_counts) {
    if (!ctx || !images || n == 0 || !d_offsets || (keypoint_capacity && (!d_keypoints || !d_descriptors)))
        return fail(SIFT_ERR_INVALID, "null argument or empty batch");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    for (size_t i = 0; i < n; i++) {
        int rc = check_image(&images[i]);
        if (rc) return rc;
        if (params && (rc = check_params(&params[i]))) return rc;
    }
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    char* base = scratch_of(C->device);
    if (!base) {
        const hipError_t e = C->device >= 0 && C->device < kMaxScratchDevices ? g_scratch_err[C->device]
                                                                             : hipErrorInvalidDevice;
        return fail(SIFT_ERR_DEVICE, std::string("scratch address of this device unknown: ") + hipGetErrorString(e));
    }
    sh::ExternalMem X;
    X.base = base;
    X.cap = SIFT_HIP_SCRATCH_BYTES;
    X.kps = reinterpret_cast<sh::Keypoint*>(d_keypoints);
    X.desc = d_descriptors;
    X.kcap = keypoint_capacity;
    X.offsets = d_offsets;
    if (!C->ext_pinned || (n + 3) * 4 > kExtSummaryBytes)
        return fail(SIFT_ERR_NOMEM, !C->ext_pinned ? "no staging for the no-allocation entry point"
                                                   : "batch too large for the no-allocation entry point");
    sh::Batch B;
    B.ext = &X;
    B.images.assign(images, images + n);
    B.params.assign(n, C->params);
    if (params) B.params.assign(params, params + n);
    B.device_input = true;
    C->prof_blocked = true;  // event timing would create events
    struct Unblock {
        sh::Context* c;
        ~Unblock() { c->prof_blocked = false; }
    } unblock{C};
    // the summary comes back through the end of the staging
    unsigned* sum = reinterpret_cast<unsigned*>(C->ext_pinned + C->ext_pinned_cap - kExtSummaryBytes);
    for (int attempt = 0; attempt < 4; attempt++) {
        sh::RunOutputs R;
        int rc = C->run(B, &R);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(sum, R.d_summary, (n + 3) * 4, hipMemcpyDeviceToHost, C->stream));
        HIP_TRY(hipStreamSynchronize(C->stream));
        if (sum[0] > (unsigned)R.cap) {  // candidates overflowed: grow within the scratch
            X.cand_cap = (size_t)sum[0] + sum[0] / 4 + 1024;
            continue;
        }
        if (out_counts)
            for (size_t i = 0; i < n; i++) out_counts[i] = sum[3 + i] - sum[2 + i];
        if (sum[1] > (unsigned)R.kcap)
            return fail(SIFT_ERR_NOMEM, "keypoint_capacity " + std::to_string(keypoint_capacity) + " < " +
                                            std::to_string(sum[1]) + " keypoints found");
        return SIFT_OK;
    }
    return fail(SIFT_ERR_NOMEM, "candidate capacity did not converge");
}

void sift_result_free(sift_result_t* r) {
    if (!r) return;
    // one pinned pool block holds both arrays, or two malloc'd buffers
    if (!sh::pinned_release(r->keypoints)) {
        free(r->keypoints);
        free(r->descriptors);
    }
    r->keypoints = nullptr;
    r->descriptors = nullptr;
    r->count = 0;
}

int sift_hip_pyramid(sift_hip_ctx* ctx, const sift_image_t* image, uint8_t* gauss, float* dog) {
    if (!ctx || !gauss) return fail(SIFT_ERR_INVALID, "null argument");
    int rc = check_image(image);
    if (rc) return rc;
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    sh::Batch B;
    B.images.assign(image, image + 1);
    B.params.assign(1, C->params);
    B.pyramid_only = true;
    B.want_dog = dog != nullptr;
    sh::RunOutputs R;
    rc = C->run(B, &R);
    if (rc) return rc;
    const int L = (int)C->params.num_intervals + 3;
    int w = (int)image->width, h = (int)image->height;
    size_t go = 0, dgo = 0;
    for (uint32_t o = 0; o < C->params.num_octaves; o++) {
        const int pitch = (int)round_up((size_t)std::max(w, 1), 64);
        if (w > 0 && h > 0) {
            for (int s = 0; s < L; s++) {
                HIP_TRY(hipMemcpy2DAsync(gauss + go + (size_t)s * w * h, w,
                                         R.arena + R.lev_off[0][o] + (size_t)s * pitch * h, pitch, w, h,
                                         hipMemcpyDeviceToHost, C->stream));
            }
            if (dog)
                HIP_TRY(hipMemcpyAsync(dog + dgo, R.arena + R.dog_off[0][o], (size_t)(L - 1) * w * h * 4,
                                       hipMemcpyDeviceToHost, C->stream));
        }
        go += (size_t)L * w * h;
        dgo += (size_t)(L - 1) * w * h;
        w /= 2;
        h /= 2;
    }
    HIP_TRY(hipStreamSynchronize(C->stream));
    return SIFT_OK;
}

int sift_hip_detect_stages(sift_hip_ctx* ctx, const sift_image_t* image, sift_stages_t* out) {
    if (!ctx || !out) return fail(SIFT_ERR_INVALID, "null argument");
    memset(out, 0, sizeof(*out));
    int rc = check_image(image);
    if (rc) return rc;
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    sh::Batch B;
    B.images.assign(image, image + 1);
    B.params.assign(1, C->params);
    B.want_stages = true;
    for (int attempt = 0; attempt < 4; attempt++) {
        sh::RunOutputs R;
        rc = C->run(B, &R);
        if (rc) return rc;
        bool retry = false;
        rc = collect(C, 1, R, &out->result, &retry);
        if (rc) {
            sift_stages_free(out);
            return rc;
        }
        if (retry) continue;
        unsigned nc = 0;
        HIP_TRY(hipMemcpy(&nc, R.d_summary, 4, hipMemcpyDeviceToHost));
        std::vector<sh::Cand> cand(nc);
        std::vector<float> resp(nc);
        std::vector<sh::Keypoint> ref(nc);
        std::vector<int> valid(nc);
        if (nc) {
            HIP_TRY(hipMemcpy(cand.data(), R.d_cand, nc * sizeof(sh::Cand), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(resp.data(), R.d_cand_resp, nc * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(ref.data(), R.d_ref, nc * sizeof(sh::Keypoint), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(valid.data(), R.d_valid, nc * 4, hipMemcpyDeviceToHost));
        }
        out->candidates = (sift_keypoint_t*)malloc(std::max<size_t>(nc, 1) * sizeof(sift_keypoint_t));
        out->n_candidates = nc;
        size_t nr = 0;
        for (unsigned c = 0; c < nc; c++) nr += valid[c] ? 1 : 0;
        out->refined = (sift_keypoint_t*)malloc(std::max<size_t>(nr, 1) * sizeof(sift_keypoint_t));
        out->n_refined = nr;
        size_t k = 0;
        for (unsigned c = 0; c < nc; c++) {
            const sh::OctaveDesc& O = R.octs[cand[c].job];
            const float scale = ldexpf(1.0f, O.octave);
            sift_keypoint_t& kp = out->candidates[c];
            kp.x = ((float)cand[c].x + 0.5f) * scale;
            kp.y = ((float)cand[c].y + 0.5f) * scale;
            kp.size = 0.0f;
            kp.angle = 0.0f;
            kp.response = resp[c];
            kp.octave = O.octave;
            kp.layer = cand[c].layer;
            if (valid[c]) memcpy(&out->refined[k++], &ref[c], sizeof(sift_keypoint_t));
        }
        return SIFT_OK;
    }
    return fail(SIFT_ERR_NOMEM, "keypoint capacity did not converge");
}

int sift_hip_tune_blur(sift_hip_ctx* ctx, uint32_t width, uint32_t height, float sigma, uint32_t jobs,
                       const char* h_form, const char* v_form, uint32_t reps, double* out_ms,
                       int* out_forms) {
    if (!ctx || !out_ms || width == 0 || height == 0 || width > 65536 || height > 65536 || jobs == 0 ||
        jobs > 1024 || reps == 0 || !(sigma > 0.0f) || sigma > 2000.0f)
        return fail(SIFT_ERR_INVALID, "invalid argument");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    const int w = (int)width, h = (int)height;
    std::vector<float> taps;
    const int nt = gaussian_taps(sigma, taps);
    const int npad = (int)round_up((size_t)nt, sh::kTapBlock);
    taps.resize(npad, 0.0f);
    const std::pair<int, int> heur = sh::blur_heuristic(w, h, nt, npad);
    const int hf = h_form && h_form[0] ? sh::blur_form_id(h_form, false) : heur.first;
    const int vf = v_form && v_form[0] ? sh::blur_form_id(v_form, true) : heur.second;
    if (hf < 0 || vf < 0) return fail(SIFT_ERR_INVALID, "unknown blur form");
    if (!sh::blur_h_valid(hf, w, h, nt, npad) || !sh::blur_v_valid(vf, w, h, nt))
        return fail(SIFT_ERR_INVALID, "blur form not valid for this level");
    if (out_forms) {
        out_forms[0] = hf;
        out_forms[1] = vf;
    }
    const size_t toff = C->tables.append(taps.data(), taps.size() * sizeof(float));
    int rc = C->tables.ensure_uploaded(C->stream);
    if (rc) return rc;
    const int pitch = (int)round_up((size_t)w, 64);
    const size_t level = (size_t)pitch * h + 64;
    sh::Carver L;
    const size_t src_off = L.take(level);
    std::vector<size_t> tmp_off(jobs), dst_off(jobs);
    for (uint32_t j = 0; j < jobs; j++) {
        tmp_off[j] = L.take(level);
        dst_off[j] = L.take(level);
    }
    const size_t hj_off = L.take(jobs * sizeof(sh::BlurJob)), vj_off = L.take(jobs * sizeof(sh::BlurJob));
    if ((rc = C->ensure_arena(L.off))) return rc;
    if ((rc = C->ensure_pinned(2 * jobs * sizeof(sh::BlurJob) + 256))) return rc;
    char* A = C->arena;
    std::vector<sh::BlurJob> hjobs(jobs), vjobs(jobs);
    int htiles = 0, vtiles = 0;
    for (uint32_t j = 0; j < jobs; j++) {
        sh::BlurJob J{};
        J.w = w;
        J.h = h;
        J.taps = (const float*)(C->tables.dev + toff);
        J.ntaps = nt;
        J.ntaps_pad = npad;
        J.src = (const uint8_t*)A + src_off;
        J.src_pitch = pitch;
        J.dst = (uint8_t*)A + tmp_off[j];
        J.dst_pitch = pitch;
        sh::blur_h_tiles(J, hf);
        J.tile_begin = htiles;
        htiles += J.tiles;
        hjobs[j] = J;
        J.src = (const uint8_t*)A + tmp_off[j];
        J.dst = (uint8_t*)A + dst_off[j];
        sh::blur_v_tiles(J, vf);
        J.tile_begin = vtiles;
        vtiles += J.tiles;
        vjobs[j] = J;
    }
    HIP_TRY(hipStreamSynchronize(C->stream));  // pinned staging is reused across calls
    memcpy(C->pinned, hjobs.data(), jobs * sizeof(sh::BlurJob));
    memcpy(C->pinned + jobs * sizeof(sh::BlurJob), vjobs.data(), jobs * sizeof(sh::BlurJob));
    HIP_TRY(hipMemcpyAsync(A + hj_off, C->pinned, jobs * sizeof(sh::BlurJob), hipMemcpyHostToDevice, C->stream));
    HIP_TRY(hipMemcpyAsync(A + vj_off, C->pinned + jobs * sizeof(sh::BlurJob), jobs * sizeof(sh::BlurJob),
                           hipMemcpyHostToDevice, C->stream));
    HIP_TRY(hipMemsetAsync(A + src_off, 0x6b, level, C->stream));  // any content: the work is data-independent
    const int n4 = (nt + 3) & ~3, r = nt / 2;
    const sh::BlurJob* dh = (const sh::BlurJob*)(A + hj_off);
    const sh::BlurJob* dv = (const sh::BlurJob*)(A + vj_off);
    auto launch_h = [&]() {
        switch (hf) {
            case sh::BH_W8: sh::launch_blur_h(dh, (int)jobs, htiles, npad, 8, C->stream); break;
            case sh::BH_W16: sh::launch_blur_h(dh, (int)jobs, htiles, npad, 16, C->stream); break;
            case sh::BH_W2: sh::launch_blur_h(dh, (int)jobs, htiles, npad, 2, C->stream); break;
            case sh::BH_SMALL: sh::launch_blur_h_small(dh, (int)jobs, htiles, n4, C->stream); break;
            default: sh::launch_blur_h_generic(dh, (int)jobs, htiles, C->stream); break;
        }
    };
    auto launch_v = [&]() {
        switch (vf) {
            case sh::BV_WIDE: sh::launch_blur_v(dv, (int)jobs, vtiles, 0, C->stream); break;
            case sh::BV_WIDE8: sh::launch_blur_v8(dv, (int)jobs, vtiles, C->stream); break;
            case sh::BV_NARROW: sh::launch_blur_v(dv, (int)jobs, vtiles, std::min(h, 16 + 2 * r), C->stream); break;
            case sh::BV_SMALL:
                sh::launch_blur_v_small(dv, (int)jobs, vtiles, n4, std::min(h, 4 + 2 * r), C->stream);
                break;
            case sh::BV_SMALL8:
                sh::launch_blur_v_small8(dv, (int)jobs, vtiles, n4, std::min(h, 4 + 2 * r), C->stream);
                break;
            default: sh::launch_blur_v_generic(dv, (int)jobs, vtiles, C->stream); break;
        }
    };
    std::vector<hipEvent_t> ev(3 * (reps + 2), nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    for (uint32_t k = 0; k < reps + 2; k++) {  // the first two are warm-up
        HIP_TRY(hipEventRecord(ev[3 * k], C->stream));
        launch_h();
        HIP_TRY(hipEventRecord(ev[3 * k + 1], C->stream));
        launch_v();
        HIP_TRY(hipEventRecord(ev[3 * k + 2], C->stream));
    }
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(C->stream);
    std::vector<float> th, tv;
    for (uint32_t k = 2; k < reps + 2 && le == hipSuccess && se == hipSuccess; k++) {
        float a = 0.0f, b = 0.0f;
        (void)hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
        (void)hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
        th.push_back(a);
        tv.push_back(b);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (le != hipSuccess) return fail(SIFT_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(le));
    if (se != hipSuccess) return fail(SIFT_ERR_DEVICE, std::string("synchronize failed: ") + hipGetErrorString(se));
    std::sort(th.begin(), th.end());
    std::sort(tv.begin(), tv.end());
    out_ms[0] = th[th.size() / 2];
    out_ms[1] = tv[tv.size() / 2];
    return SIFT_OK;
}

int sift_hip_profile(sift_hip_ctx* ctx, int enable) {
    if (!ctx) return fail(SIFT_ERR_INVALID, "null context");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    HIP_TRY(hipStreamSynchronize(C->stream));
    C->prof_resolve();
    C->prof_stats.clear();
    C->profiling = enable != 0;
    return SIFT_OK;
}

int sift_hip_profile_filter(sift_hip_ctx* ctx, const char* kernel) {
    if (!ctx) return fail(SIFT_ERR_INVALID, "null context");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    int k = -1;
    if (kernel && kernel[0]) {
        for (int i = 0; i < K_COUNT; i++)
            if (strcmp(kKernelNames[i], kernel) == 0) k = i;
        if (k < 0) return fail(SIFT_ERR_INVALID, std::string("unknown kernel name: ") + kernel);
    }
    C->prof_only = k;
    return SIFT_OK;
}

int sift_hip_profile_read(sift_hip_ctx* ctx, sift_kernel_stat_t* out, size_t cap, size_t* n) {
    if (!ctx || !n) return fail(SIFT_ERR_INVALID, "null argument");
    sh::Context* C = reinterpret_cast<sh::Context*>(ctx);
    std::lock_guard<std::mutex> lk(C->mu);
    (void)hipSetDevice(C->device);
    HIP_TRY(hipStreamSynchronize(C->stream));
    C->prof_resolve();
    size_t k = 0;
    for (const auto& kv : C->prof_stats) {
        if (out && k < cap) {
            memset(&out[k], 0, sizeof(out[k]));
            strncpy(out[k].name, kKernelNames[kv.first], sizeof(out[k].name) - 1);
            out[k].launches = kv.second.launches;
            out[k].total_ms = kv.second.total_ms;
            out[k].bytes = kv.second.bytes;
            out[k].ops = kv.second.ops;
        }
        k++;
    }
    *n = k;
    return SIFT_OK;
}

void sift_stages_free(sift_stages_t* s) {
    if (!s) return;
    free(s->candidates);
    free(s->refined);
    sift_result_free(&s->result);
    s->candidates = nullptr;
    s->refined = nullptr;
    s->n_candidates = s->n_refined = 0;
}

}  // extern "C"
