// sift_context.h -- host-side context of libsift_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "sift_hip.h"
#include "sift_internal.h"

namespace sh {

// Bytes per pixel of a SIFT_PIXEL_* format (0 = unknown format).
inline int pixel_bytes(uint32_t format) {
    switch (format) {
        case SIFT_PIXEL_GRAY8: return 1;
        case SIFT_PIXEL_RGB8: return 3;
        case SIFT_PIXEL_RGBA8: return 4;
        case SIFT_PIXEL_LUMAA8: return 2;
        case SIFT_PIXEL_RGBA8_CANVAS: return 4;
        default: return 0;
    }
}

struct TapsRef {
    int ntaps = 0, npad = 0;
    size_t off = 0;      // byte offset in the table store
    bool copy = false;   // blur sigma < 1e-4: the level is a copy of the previous one
};

struct ParamTables {
    TapsRef initial;
    std::vector<TapsRef> level;  // [o * (S+3) + s], s >= 1
};

struct LanczosRef {
    int maxtaps = 0;
    size_t left_off = 0, count_off = 0, w_off = 0;
    std::vector<int> left, count;  // host copies (tile spans of the fused resize)
};

struct ParamKey {
    float sigma, assumed_blur;
    uint32_t S, O;
    bool operator<(const ParamKey& o) const {
        return std::tie(sigma, assumed_blur, S, O) < std::tie(o.sigma, o.assumed_blur, o.S, o.O);
    }
};

// Host mirror + device copy of all weight tables, append-only. New tables are
// uploaded asynchronously on the owning stream through a pinned bounce buffer
// (reused only after the stream has been synchronised, which every public call
// does before returning).
struct TableStore {
    std::vector<char> host;
    char* dev = nullptr;
    size_t dev_cap = 0, uploaded = 0;
    char* bounce = nullptr;
    size_t bounce_cap = 0;
    size_t append(const void* p, size_t bytes);
    int ensure_uploaded(hipStream_t stream);
};

// Pinned staging of the no-allocation entry point (per context, allocated at
// creation) and the part of it that receives the summary (up to 16K frames).
constexpr size_t kExtStagingBytes = (size_t)4 << 20;
constexpr size_t kExtSummaryBytes = (size_t)64 << 10;

// Device memory a call works in instead of the context's own (the
// no-allocation entry point): the weight tables at base, then the call's
// layout, all within [base, base + cap); keypoints and descriptors go to the
// caller's buffers (kcap entries). cand_cap: candidate capacity of the call (0 =
// the initial rule).
struct ExternalMem {
    char* base = nullptr;
    size_t cap = 0;
    Keypoint* kps = nullptr;
    float* desc = nullptr;
    size_t kcap = 0;
    size_t cand_cap = 0;
    unsigned* offsets = nullptr;  // the caller's n + 1 per-image keypoint offsets
};

struct Batch {
    const ExternalMem* ext = nullptr;  // set: no allocation, no graphs, no timing
    std::vector<sift_image_t> images;
    std::vector<uint32_t> formats;  // SIFT_PIXEL_* per image; empty = all grey
    std::vector<sift_params_t> params;
    bool desc_u8 = false;           // also produce the u8 descriptors (RunOutputs::d_desc_u8)
    bool device_input = false;
    bool host_pinned = false;  // host inputs are pinned (pinned_alloc): copied without staging
    bool pyramid_only = false;
    bool want_dog = false;
    bool want_stages = false;
    uint32_t format(size_t i) const { return formats.empty() ? SIFT_PIXEL_GRAY8 : formats[i]; }
};

struct RunOutputs {
    Keypoint* d_kps = nullptr;
    float* d_desc = nullptr;
    uint8_t* d_desc_u8 = nullptr;
    unsigned* d_summary = nullptr;
    Cand* d_cand = nullptr;
    float* d_cand_resp = nullptr;
    Keypoint* d_ref = nullptr;
    int* d_valid = nullptr;
    const OctaveDesc* d_octs = nullptr;
    std::vector<OctaveDesc> octs;
    int cap = 0, kcap = 0;
    std::vector<std::vector<size_t>> lev_off, dog_off;
    char* arena = nullptr;
};

struct Server;

// Blur pass forms (context.hip) and the tuning table's key: a level's width,
// height and tap count, and whether its launch carries many jobs (a batch of
// at least kTuneManyJobs frames) or few.
enum BlurHForm { BH_W8 = 0, BH_W2, BH_SMALL, BH_GENERIC, BH_W16, BH_COUNT };
enum BlurVForm { BV_WIDE = 0, BV_NARROW, BV_SMALL, BV_GENERIC, BV_SMALL8, BV_WIDE8, BV_COUNT };
constexpr int kTuneManyJobs = 8;
struct TuneKey {
    int w, h, ntaps, many;
    bool operator<(const TuneKey& o) const {
        return std::tie(w, h, ntaps, many) < std::tie(o.w, o.h, o.ntaps, o.many);
    }
};

// Per-kernel event timing (sift_hip_profile).
struct ProfPending {
    int kernel;
    double bytes, ops;
    hipEvent_t a, b;
};
struct ProfStat {
    uint64_t launches = 0;
    double total_ms = 0.0, bytes = 0.0, ops = 0.0;
};

// A captured launch sequence of one batch layout (see Context::run).
struct GraphEntry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
};

struct Context {
    int device = 0;
    sift_params_t params{};
    hipStream_t stream = nullptr;
    std::mutex mu;

    TableStore tables;
    std::map<ParamKey, ParamTables> param_tables;
    std::map<std::pair<int, int>, LanczosRef> lanczos;

    char* arena = nullptr;
    size_t arena_cap = 0;
    // Growth rule of the arena: at least arena_floor bytes per allocation, and
    // never more than arena_limit (0 = no limit; a server lane's share of the HBM
    // budget). A call whose layout needs more than the limit fails with
    // SIFT_ERR_NOMEM.
    size_t arena_floor = (size_t)64 << 20;
    size_t arena_limit = 0;
    char* pinned = nullptr;
    size_t pinned_cap = 0;
    char* pinned_out = nullptr;  // device-to-host staging for results
    size_t pinned_out_cap = 0;
    size_t cand_cap = 0, kp_cap = 0;

    // The server of this context (owned; created by sift_hip_server_start).
    // server_mu guards the pointer and the server's count of callers inside
    // submit / poll, so that sift_hip_server_stop deletes it only after they left.
    std::mutex server_mu;
    std::condition_variable server_cv;
    Server* server = nullptr;

    // Launch sequences replayed as HIP graphs, keyed by everything the kernel
    // arguments depend on (arena base, layout offsets, launch shapes, capacities).
    bool use_graphs = true;  // SIFT_HIP_GRAPHS=0 in the environment turns them off
    std::vector<GraphEntry> graphs;
    std::vector<uint64_t> seen_once;  // hashes of keys run once, not captured yet
    uint64_t graph_clock = 0;
    void clear_graphs();

    // Pinned staging of the external-memory calls (tables and job blob in, the
    // summary out), allocated with the context so that those calls allocate nothing.
    char* ext_pinned = nullptr;
    size_t ext_pinned_cap = 0;

    bool profiling = false;
    bool prof_blocked = false;  // no event timing (it creates events) during a no-allocation call
    int prof_only = -1;  // kernel id whose launches are timed, -1 = all
    std::vector<ProfPending> prof_pending;
    std::vector<hipEvent_t> prof_events;  // free pool
    std::map<int, ProfStat> prof_stats;
    hipEvent_t prof_event();
    void prof_begin(int kernel, double bytes, hipEvent_t* a);
    void prof_end(int kernel, double bytes, hipEvent_t a, double ops = 0.0);
    void prof_resolve();  // call after the stream is idle

    // Blur forms chosen by the tuning table (w, h, taps, many) -> (H form, V form),
    // loaded at creation; shapes it does not list take the built-in choice.
    std::map<TuneKey, std::pair<int, int>> tune;
    std::pair<int, int> blur_choice(int w, int h, int ntaps, int npad, bool many,
                                    std::pair<int, int> forced) const;

    const ParamTables& tables_for(const sift_params_t& p);
    const LanczosRef& lanczos_for(int in, int out);
    int ensure_arena(size_t bytes);
    int ensure_pinned(size_t bytes);
    int ensure_pinned_out(size_t bytes);
    void release();  // frees every device and pinned allocation of the context
    // Enqueues the whole pipeline for a batch on `stream` (no host sync except
    // for table uploads); outputs stay in the arena until the next run.
    int run(const Batch& B, RunOutputs* outs);
};

// Sets the calling thread's sift_hip_last_error message; returns code.
int set_error(int code, const char* msg);

// Pinned host blocks from a process-wide pool (host_pool.hip): nullptr when
// the pool is at its limit or the allocation fails. pinned_release returns a
// block to the pool, or false for a pointer the pool did not hand out.
void* pinned_alloc(size_t bytes);
bool pinned_release(void* p);

// Device bytes one frame takes in a batch at the initial capacities: an upper
// bound of what Context::run carves for it (server admission).
size_t frame_device_bytes(uint32_t w, uint32_t h, const sift_params_t& p, int bpp);

// Runs a host-input batch and copies the results out (f32 descriptors into
// `out`, or u8 descriptors into `out8` when the batch has desc_u8 set), growing
// the capacities and re-running when a count overflowed them.
int run_and_collect(Context* C, const Batch& B, sift_result_t* out, sift_result_u8_t* out8 = nullptr,
                    uint32_t kp_conv = SIFT_KEYPOINTS_CPU);

}  // namespace sh
