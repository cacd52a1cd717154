// server.hip -- continuous-batching request loop on one device.
//
// sift_hip_submit copies a frame into host memory and queues it. Worker lanes
// (four by default, fewer when the HBM budget is small), each with its own HIP
// stream and device arena, take queued frames (up to max_batch, while the device
// bytes of the lane's batch stay within its share of the HBM budget, and up to
// about two 1080p frames of pixels, or an equal share of a longer queue per
// lane), run them as one batch through the same kernels as
// sift_hip_detect_batch, and post completions. Frames that arrive while batches
// are on the device go into the next batch of whichever lane frees first, so the
// lanes' batches overlap on the device and queueing stays short under load; a
// backlog is taken in fewer, larger batches, which raise the served rate.
//
// Device memory: each lane's arena is allocated once, at start, as its share
// (budget / lanes) and never grows past it, so the lanes together hold at most
// the budget; the weight tables (a few hundred KB per lane) are outside it.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "sift_context.h"

namespace sh {

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t round_up_sz(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr int kDefaultLanes = 4;
// A lane per this many bytes of budget, up to kDefaultLanes.
constexpr size_t kLaneBytes = (size_t)2 << 30;
// Pixels of one device batch while the queue is short (about two 1080p
// frames); a larger first frame still runs alone.
constexpr size_t kMaxBatchPixels = (size_t)4200000;

}  // namespace

size_t frame_device_bytes(uint32_t w, uint32_t h, const sift_params_t& p, int bpp) {
    constexpr size_t kAlign = 256;  // every carved buffer starts on a 256-byte boundary
    const size_t L = p.num_intervals + 3, S = p.num_intervals;
    size_t bytes = round_up_sz(w, 64) * h + kAlign;                          // input copy
    if (bpp > 1) bytes += round_up_sz((size_t)w * bpp, 64) * h + kAlign;     // colour pixels
    size_t px_layers = 0, slots = 0, words = 0;
    uint32_t ww = w, hh = h;
    for (uint32_t o = 0; o < p.num_octaves; o++) {
        bytes += L * round_up_sz(ww ? ww : 1, 64) * hh + 64 + kAlign;  // levels
        bytes += round_up_sz(ww ? ww : 1, 64) * hh + 64 + kAlign;      // its H-pass temporary
        px_layers += S * ww * hh;
        slots += S * hh;
        words += S * hh * ((ww + 63) / 64);
        ww /= 2;
        hh /= 2;
    }
    const size_t pitch0 = round_up_sz(w, 64);
    bytes += pitch0 * std::max<size_t>(h / 2, 1) * 4 + kAlign;           // Lanczos temporary
    bytes += words * 8 + 2 * (slots + kCountExtra) * 4 + (slots / 4096 + 4) * 4 + 4 * kAlign;  // bitmask, counts, scan
    const size_t cand = std::max<size_t>(4096, px_layers / 128);
    // per candidate: Cand, response, refined keypoint, valid, peak count, angles,
    // keypoint offset, two descriptor lists; per keypoint (1.5 per candidate):
    // keypoint and 128 f32
    bytes += cand * (sizeof(Cand) + 4 + sizeof(Keypoint) + 4 + 4 + kMaxPeaks * 4 + 4 + 4 + 4) + 12 * kAlign;
    bytes += (cand + cand / 2) * (sizeof(Keypoint) + 128 * 4) + 2 * kAlign;
    // job blob: this frame's jobs and octave records, and (counted per frame, an
    // over-estimate for batches) the launch records of a whole batch
    const size_t O = p.num_octaves;
    bytes += ((1 + O * L) * 2 + O * 2) * std::max(sizeof(BlurJob), sizeof(ResizeJob)) +
             O * sizeof(OctaveDesc) + sizeof(ImageParams) + 16;
    bytes += ((1 + O * L) * 15 + O * 3 + 8) * 256 + 4096 + 64 + kAlign;
    return bytes;
}

struct Request {
    uint64_t tag;
    uint32_t w, h;
    uint32_t format;  // SIFT_PIXEL_*
    sift_params_t params;
    // the frame's rows, packed: a pinned pool block (copied to the device without
    // staging), or, when the pool has none to give, plain host memory
    struct Pixels {
        uint8_t* p = nullptr;
        bool pinned = false;
        Pixels() = default;
        Pixels(const Pixels&) = delete;
        Pixels& operator=(const Pixels&) = delete;
        Pixels(Pixels&& o) noexcept : p(o.p), pinned(o.pinned) { o.p = nullptr; }
        Pixels& operator=(Pixels&& o) noexcept {
            std::swap(p, o.p);
            std::swap(pinned, o.pinned);
            return *this;
        }
        ~Pixels() {
            if (pinned) pinned_release(p);
            else delete[] p;
        }
        bool alloc(size_t bytes) {
            p = (uint8_t*)pinned_alloc(bytes);
            pinned = p != nullptr;
            if (!p) p = new (std::nothrow) uint8_t[bytes];
            return p != nullptr;
        }
    } pixels;
    double submit_ms;
    size_t bytes;
};

struct Server {
    bool trace = false;  // SIFT_HIP_TRACE=1: one stderr line per batch
    // Pixels of one batch: base_px (SIFT_HIP_BATCH_PIXELS overrides it), or,
    // when more is queued, an equal share of the queue per lane
    // (SIFT_HIP_BATCH_ADAPT=0 turns that off).
    size_t base_px = kMaxBatchPixels;
    bool adapt = true;
    size_t pending_px = 0;  // pixels of the queued frames
    Context* owner = nullptr;
    sift_server_config_t cfg{};
    size_t lane_share = 0;  // device bytes of one lane's arena
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::deque<sift_completion_t> done;
    size_t bytes_in_flight = 0;
    bool stopping = false;
    int users = 0;  // callers inside submit / poll (guarded by the owner's server_mu)
    std::vector<Context*> lanes;
    std::vector<std::thread> workers;

    void worker(Context* lane);
};

namespace {

// Runs a batch on a lane; when the batch as a whole does not fit the lane's
// share (a capacity grown past it by keypoint-rich frames), its frames are run
// one at a time, so that only a frame that cannot fit alone fails.
void run_batch(Context* lane, const std::vector<Request>& batch, std::vector<sift_result_t>& res,
               std::vector<int>& rcs) {
    auto make = [&](size_t i0, size_t i1) {
        Batch B;
        B.host_pinned = true;
        for (size_t i = i0; i < i1; i++) {
            const uint32_t row = batch[i].w * (uint32_t)pixel_bytes(batch[i].format);
            B.host_pinned = B.host_pinned && batch[i].pixels.pinned;
            B.images.push_back(sift_image_t{batch[i].pixels.p, batch[i].w, batch[i].h, row});
            B.formats.push_back(batch[i].format);
            B.params.push_back(batch[i].params);
        }
        return B;
    };
    const int rc = run_and_collect(lane, make(0, batch.size()), res.data());
    std::fill(rcs.begin(), rcs.end(), rc);
    if (rc != SIFT_ERR_NOMEM || batch.size() == 1) return;
    for (size_t i = 0; i < batch.size(); i++) rcs[i] = run_and_collect(lane, make(i, i + 1), &res[i]);
}

}  // namespace

void Server::worker(Context* lane) {
    (void)hipSetDevice(lane->device);
    for (;;) {
        std::vector<Request> batch;
        size_t batch_bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stopping && pending.empty()) return;
                size_t batch_px = 0;
                // small batches while the queue is short (latency); under a
                // backlog each lane takes its share of it (fewer, larger batches),
                // so the backlog still spreads over the lanes
                const size_t cap_px = adapt ? std::max(base_px, pending_px / lanes.size()) : base_px;
                while (!pending.empty() && batch.size() < cfg.max_batch) {
                    Request& r = pending.front();
                    if (bytes_in_flight + r.bytes > cfg.hbm_budget_bytes) break;
                    if (batch_bytes + r.bytes > lane_share) break;
                    const size_t px = (size_t)r.w * r.h;
                    if (!batch.empty() && batch_px + px > cap_px) break;
                    batch_px += px;
                    pending_px -= px;
                    bytes_in_flight += r.bytes;
                    batch_bytes += r.bytes;
                    batch.push_back(std::move(r));
                    pending.pop_front();
                }
                if (!batch.empty()) break;
                cv_work.wait(lk);  // woken by submit, by memory release, or by stop
            }
        }
        const double start = now_ms();
        std::vector<sift_result_t> res(batch.size(), sift_result_t{nullptr, nullptr, 0});
        std::vector<int> rcs(batch.size(), SIFT_OK);
        run_batch(lane, batch, res, rcs);
        const double end = now_ms();
        if (trace) {
            double oldest = start;
            for (const Request& r : batch) oldest = std::min(oldest, r.submit_ms);
            fprintf(stderr, "[sift-hip server] lane %p batch %zu frames %.1f MB queued %.3f ms run %.3f ms rc %d\n",
                    (void*)lane, batch.size(), batch_bytes / 1e6, start - oldest, end - start, rcs[0]);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < batch.size(); i++) {
                sift_completion_t c{};
                c.tag = batch[i].tag;
                c.status = rcs[i];
                if (rcs[i] == SIFT_OK) c.result = res[i];
                else sift_result_free(&res[i]);
                c.submit_ms = batch[i].submit_ms;
                c.start_ms = start;
                c.done_ms = end;
                done.push_back(c);
            }
            bytes_in_flight -= batch_bytes;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

namespace {

// The context's server with its caller count raised (nullptr if none is
// running); release_server lowers it again.
Server* acquire_server(Context* C) {
    std::lock_guard<std::mutex> lk(C->server_mu);
    Server* S = C->server;
    if (S) S->users++;
    return S;
}

void release_server(Context* C, Server* S) {
    {
        std::lock_guard<std::mutex> lk(C->server_mu);
        S->users--;
    }
    C->server_cv.notify_all();
}

}  // namespace

// Copies the frame into host memory and queues it (sift_hip_submit*).
int submit_frame(sift_hip_ctx* ctx, const sift_color_image_t* image, const sift_params_t* params,
                 uint64_t tag) {
    const double t_submit = now_ms();
    const int bpp = image ? pixel_bytes(image->format) : 0;
    if (!ctx || !image || !image->data || bpp == 0 || image->width == 0 || image->height == 0 ||
        image->width > 65536 || image->height > 65536 ||
        (image->stride && image->stride < (size_t)image->width * bpp))
        return set_error(SIFT_ERR_INVALID, "invalid image (null, zero or oversized, unknown format, or stride too small)");
    Context* C = reinterpret_cast<Context*>(ctx);
    const sift_params_t& P = params ? *params : C->params;
    if (P.num_intervals < 1 || P.num_intervals > 13 || P.num_octaves < 1 || P.num_octaves > 16 ||
        !(P.sigma > 0.0f))
        return set_error(SIFT_ERR_INVALID, "invalid params");
    Server* S = acquire_server(C);
    if (!S) return set_error(SIFT_ERR_INVALID, "no server running on this context");
    int rc = SIFT_OK;
    Request r;
    r.tag = tag;
    r.w = image->width;
    r.h = image->height;
    r.format = image->format;
    r.params = P;
    r.bytes = frame_device_bytes(r.w, r.h, P, bpp);
    if (r.bytes > S->lane_share) {
        rc = set_error(SIFT_ERR_NOMEM, "frame needs more device memory than a lane's share of the HBM budget");
    } else {
        const size_t row = (size_t)r.w * bpp;
        const bool ok = r.pixels.alloc(row * r.h);
        const size_t stride = image->stride ? image->stride : row;
        if (ok && stride == row) {
            memcpy(r.pixels.p, image->data, row * r.h);
        } else if (ok) {
            for (uint32_t y = 0; y < r.h; y++)
                memcpy(r.pixels.p + (size_t)y * row, image->data + (size_t)y * stride, row);
        }
        r.submit_ms = t_submit;
        std::lock_guard<std::mutex> lk(S->mu);
        if (!ok) rc = set_error(SIFT_ERR_NOMEM, "host allocation failed");
        else if (S->stopping) rc = set_error(SIFT_ERR_INVALID, "server is stopping");
        else if (S->pending.size() >= S->cfg.max_queue) rc = set_error(SIFT_ERR_BUSY, "server queue is full");
        else {
            S->pending_px += (size_t)r.w * r.h;
            S->pending.push_back(std::move(r));
        }
    }
    if (rc == SIFT_OK) S->cv_work.notify_one();
    release_server(C, S);
    return rc;
}

}  // namespace sh

using namespace sh;

extern "C" {

int sift_hip_server_start(sift_hip_ctx* ctx, const sift_server_config_t* cfg) {
    if (!ctx || !cfg) return set_error(SIFT_ERR_INVALID, "null argument");
    Context* C = reinterpret_cast<Context*>(ctx);
    {
        std::lock_guard<std::mutex> lk(C->server_mu);
        if (C->server) return set_error(SIFT_ERR_INVALID, "a server is already running on this context");
    }
    Server* S = new (std::nothrow) Server();
    if (!S) return set_error(SIFT_ERR_NOMEM, "host allocation failed");
    S->owner = C;
    S->cfg = *cfg;
    const char* tr = getenv("SIFT_HIP_TRACE");
    S->trace = tr && tr[0] == '1';
    if (const char* e = getenv("SIFT_HIP_BATCH_PIXELS")) S->base_px = (size_t)std::max(1L, atol(e));
    if (const char* e = getenv("SIFT_HIP_BATCH_ADAPT")) S->adapt = e[0] != '0';
    if (S->cfg.max_batch == 0) S->cfg.max_batch = 64;
    if (S->cfg.max_queue == 0) S->cfg.max_queue = 4096;
    if (S->cfg.hbm_budget_bytes == 0) S->cfg.hbm_budget_bytes = (size_t)32 << 30;
    // Lanes: independent streams whose batches overlap on the device; small
    // frames leave most of the GPU idle, so several lanes raise the served rate.
    // One lane per kLaneBytes of budget, up to kDefaultLanes.
    size_t nlanes = std::min<size_t>(kDefaultLanes, std::max<size_t>(1, S->cfg.hbm_budget_bytes / kLaneBytes));
    if (const char* e = getenv("SIFT_HIP_LANES")) nlanes = (size_t)std::min(8, std::max(1, atoi(e)));
    S->lane_share = S->cfg.hbm_budget_bytes / nlanes / 256 * 256;
    int err = SIFT_ERR_DEVICE;
    for (size_t i = 0; i < nlanes; i++) {
        Context* L = new (std::nothrow) Context();
        if (!L) break;
        L->device = C->device;
        L->params = C->params;
        if (hipSetDevice(C->device) != hipSuccess ||
            hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess) {
            delete L;
            break;
        }
        // the lane's device pool: its share of the budget, allocated now so that
        // the arena never grows (and synchronises) mid-stream; pinned staging
        // for a batch of a few 4K frames each way
        L->arena_floor = 0;
        L->arena_limit = S->lane_share;
        int rc = L->ensure_arena(S->lane_share);
        if (rc == SIFT_OK) rc = L->ensure_pinned((size_t)64 << 20);
        if (rc == SIFT_OK) rc = L->ensure_pinned_out((size_t)64 << 20);
        if (rc != SIFT_OK) {
            err = rc;
            L->release();
            delete L;
            break;
        }
        S->lanes.push_back(L);
    }
    if (S->lanes.size() < nlanes) {  // all lanes or none: the shares assume nlanes
        for (Context* L : S->lanes) {
            L->release();
            delete L;
        }
        delete S;
        return set_error(err, "server lanes could not be created (device stream or memory)");
    }
    for (Context* L : S->lanes) S->workers.emplace_back([S, L] { S->worker(L); });
    std::lock_guard<std::mutex> lk(C->server_mu);
    C->server = S;
    return SIFT_OK;
}

int sift_hip_submit(sift_hip_ctx* ctx, const sift_image_t* image, const sift_params_t* params,
                    uint64_t tag) {
    if (!image) return set_error(SIFT_ERR_INVALID, "image is null");
    const sift_color_image_t c{image->data, image->width, image->height, image->stride, SIFT_PIXEL_GRAY8};
    return sh::submit_frame(ctx, &c, params, tag);
}

int sift_hip_submit_color(sift_hip_ctx* ctx, const sift_color_image_t* image,
                          const sift_params_t* params, uint64_t tag) {
    return sh::submit_frame(ctx, image, params, tag);
}

int sift_hip_poll(sift_hip_ctx* ctx, sift_completion_t* out, size_t cap, size_t* n, int timeout_ms) {
    if (n) *n = 0;
    if (!ctx || !out || !n) return set_error(SIFT_ERR_INVALID, "null argument");
    Context* C = reinterpret_cast<Context*>(ctx);
    Server* S = acquire_server(C);
    if (!S) return set_error(SIFT_ERR_INVALID, "no server running on this context");
    size_t k = 0;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        // a stop wakes the waiters, so that they leave before the server is deleted
        auto ready = [&] { return !S->done.empty() || S->stopping; };
        if (!ready() && timeout_ms != 0) {
            if (timeout_ms < 0) S->cv_done.wait(lk, ready);
            else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
        }
        while (k < cap && !S->done.empty()) {
            out[k++] = S->done.front();
            S->done.pop_front();
        }
    }
    release_server(C, S);
    *n = k;
    return k ? SIFT_OK : set_error(SIFT_ERR_TIMEOUT, "no completion within the timeout");
}

int sift_hip_server_stop(sift_hip_ctx* ctx) {
    if (!ctx) return set_error(SIFT_ERR_INVALID, "null context");
    Context* C = reinterpret_cast<Context*>(ctx);
    Server* S = nullptr;
    {
        // detach it: later submit / poll calls see no server
        std::lock_guard<std::mutex> lk(C->server_mu);
        S = C->server;
        C->server = nullptr;
    }
    if (!S) return SIFT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    for (auto& t : S->workers) t.join();
    {
        // callers still inside submit / poll leave once they see `stopping`
        std::unique_lock<std::mutex> lk(C->server_mu);
        C->server_cv.wait(lk, [&] { return S->users == 0; });
    }
    for (auto& c : S->done) sift_result_free(&c.result);
    for (Context* L : S->lanes) {
        L->release();
        delete L;
    }
    delete S;
    return SIFT_OK;
}

}  // extern "C"
