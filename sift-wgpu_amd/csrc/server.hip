// server.hip -- continuous-batching request loop on one device.
//
// sift_hip_submit copies a frame into host memory and queues it. Two worker
// lanes, each with its own HIP stream and device arena, take every queued frame
// that fits (up to max_batch, and while the bytes of admitted frames stay within
// the HBM budget), run them as one batch through the same kernels as
// sift_hip_detect_batch, and post completions. Frames that arrive while a batch
// is on the device go into the next batch of whichever lane frees first, so two
// batches overlap on the device and queueing stays short under load.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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
// Pixels of one device batch (about two 1080p frames); a larger first frame
// still runs alone.
constexpr size_t kMaxBatchPixels = (size_t)4200000;

}  // namespace

// Device bytes one frame needs in a batch (pyramid, temporaries, detection
// buffers at the default capacity), used for admission against the budget.
size_t frame_device_bytes(uint32_t w, uint32_t h, const sift_params_t& p, int bpp = 1) {
    const size_t L = p.num_intervals + 3;
    size_t bytes = round_up_sz(w, 64) * h;  // input copy
    if (bpp > 1) bytes += round_up_sz((size_t)w * bpp, 64) * h;  // colour pixels
    size_t px_layers = 0;
    uint32_t ww = w, hh = h;
    for (uint32_t o = 0; o < p.num_octaves; o++) {
        bytes += L * round_up_sz(ww ? ww : 1, 64) * hh + 256;
        px_layers += (size_t)p.num_intervals * ww * hh;
        ww /= 2;
        hh /= 2;
    }
    bytes += round_up_sz(w, 64) * h * 3;  // blur and resize temporaries
    const size_t cand = std::max<size_t>(4096, px_layers / 128);
    bytes += cand * (sizeof(Cand) + 4 + sizeof(Keypoint) + 4 + 4 + kMaxPeaks * 4 + 4);
    bytes += cand * 3 / 2 * (sizeof(Keypoint) + 128 * 4);
    return bytes;
}

struct Request {
    uint64_t tag;
    uint32_t w, h;
    uint32_t format;  // SIFT_PIXEL_*
    sift_params_t params;
    std::vector<uint8_t> pixels;
    double submit_ms;
    size_t bytes;
};

struct Server {
    bool trace = false;  // SIFT_HIP_TRACE=1: one stderr line per batch
    Context* owner = nullptr;
    sift_server_config_t cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::deque<sift_completion_t> done;
    size_t bytes_in_flight = 0;
    bool stopping = false;
    std::vector<Context*> lanes;
    std::vector<std::thread> workers;

    void worker(Context* lane);
};

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
                while (!pending.empty() && batch.size() < cfg.max_batch) {
                    Request& r = pending.front();
                    if (bytes_in_flight + r.bytes > cfg.hbm_budget_bytes) break;
                    // keep batches small enough that a backlog spreads over the lanes
                    const size_t px = (size_t)r.w * r.h;
                    if (!batch.empty() && batch_px + px > kMaxBatchPixels) break;
                    batch_px += px;
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
        Batch B;
        B.images.resize(batch.size());
        B.formats.resize(batch.size());
        B.params.resize(batch.size());
        for (size_t i = 0; i < batch.size(); i++) {
            const uint32_t row = batch[i].w * (uint32_t)pixel_bytes(batch[i].format);
            B.images[i] = sift_image_t{batch[i].pixels.data(), batch[i].w, batch[i].h, row};
            B.formats[i] = batch[i].format;
            B.params[i] = batch[i].params;
        }
        std::vector<sift_result_t> res(batch.size());
        int rc = run_and_collect(lane, B, res.data());
        const double end = now_ms();
        if (trace) {
            double oldest = start;
            for (const Request& r : batch) oldest = std::min(oldest, r.submit_ms);
            fprintf(stderr, "[sift-hip server] lane %p batch %zu frames %.1f MB queued %.3f ms run %.3f ms rc %d\n",
                    (void*)lane, batch.size(), batch_bytes / 1e6, start - oldest, end - start, rc);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < batch.size(); i++) {
                sift_completion_t c{};
                c.tag = batch[i].tag;
                c.status = rc;
                if (rc == SIFT_OK) c.result = res[i];
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

// Copies the frame into host memory and queues it (sift_hip_submit*).
int submit_frame(sift_hip_ctx* ctx, const sift_color_image_t* image, const sift_params_t* params,
                 uint64_t tag) {
    const double t_submit = now_ms();
    const int bpp = image ? pixel_bytes(image->format) : 0;
    if (!ctx || !image || !image->data || bpp == 0 || image->width == 0 || image->height == 0 ||
        image->width > 65536 || image->height > 65536 ||
        (image->stride && image->stride < (size_t)image->width * bpp))
        return SIFT_ERR_INVALID;
    Context* C = reinterpret_cast<Context*>(ctx);
    Server* S = C->server;
    if (!S) return SIFT_ERR_INVALID;
    const sift_params_t& P = params ? *params : C->params;
    if (P.num_intervals < 1 || P.num_intervals > 13 || P.num_octaves < 1 || P.num_octaves > 16 ||
        !(P.sigma > 0.0f))
        return SIFT_ERR_INVALID;
    Request r;
    r.tag = tag;
    r.w = image->width;
    r.h = image->height;
    r.format = image->format;
    r.params = P;
    r.bytes = frame_device_bytes(r.w, r.h, P, bpp);
    if (r.bytes > S->cfg.hbm_budget_bytes) return SIFT_ERR_NOMEM;
    const size_t row = (size_t)r.w * bpp;
    r.pixels.resize(row * r.h);
    const size_t stride = image->stride ? image->stride : row;
    for (uint32_t y = 0; y < r.h; y++)
        memcpy(r.pixels.data() + (size_t)y * row, image->data + (size_t)y * stride, row);
    r.submit_ms = t_submit;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) return SIFT_ERR_INVALID;
        if (S->pending.size() >= S->cfg.max_queue) return SIFT_ERR_BUSY;
        S->pending.push_back(std::move(r));
    }
    S->cv_work.notify_one();
    return SIFT_OK;
}

}  // namespace sh

using namespace sh;

extern "C" {

int sift_hip_server_start(sift_hip_ctx* ctx, const sift_server_config_t* cfg) {
    if (!ctx || !cfg) return SIFT_ERR_INVALID;
    Context* C = reinterpret_cast<Context*>(ctx);
    if (C->server) return SIFT_ERR_INVALID;
    Server* S = new (std::nothrow) Server();
    if (!S) return SIFT_ERR_NOMEM;
    S->owner = C;
    S->cfg = *cfg;
    const char* tr = getenv("SIFT_HIP_TRACE");
    S->trace = tr && tr[0] == '1';
    if (S->cfg.max_batch == 0) S->cfg.max_batch = 64;
    if (S->cfg.max_queue == 0) S->cfg.max_queue = 4096;
    if (S->cfg.hbm_budget_bytes == 0) S->cfg.hbm_budget_bytes = (size_t)32 << 30;
    // Lanes: independent streams whose batches overlap on the device; small
    // frames leave most of the GPU idle, so several lanes raise the served rate.
    int nlanes = kDefaultLanes;
    if (const char* e = getenv("SIFT_HIP_LANES")) nlanes = std::min(8, std::max(1, atoi(e)));
    for (int i = 0; i < nlanes; i++) {
        Context* L = new (std::nothrow) Context();
        if (!L) break;
        L->device = C->device;
        L->params = C->params;
        if (hipSetDevice(C->device) != hipSuccess ||
            hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess) {
            delete L;
            break;
        }
        // preallocated device pool of the lane: room for a full batch of large
        // frames, so the arena does not grow (and synchronise) mid-stream
        // and pinned staging for a batch of a few 4K frames each way
        const size_t pool = std::min(S->cfg.hbm_budget_bytes, (size_t)6 << 30);
        if (L->ensure_arena(pool) != SIFT_OK || L->ensure_pinned((size_t)64 << 20) != SIFT_OK ||
            L->ensure_pinned_out((size_t)64 << 20) != SIFT_OK) {
            L->release();
            delete L;
            break;
        }
        S->lanes.push_back(L);
    }
    if (S->lanes.empty()) {
        delete S;
        return SIFT_ERR_DEVICE;
    }
    for (Context* L : S->lanes) S->workers.emplace_back([S, L] { S->worker(L); });
    C->server = S;
    return SIFT_OK;
}

int sift_hip_submit(sift_hip_ctx* ctx, const sift_image_t* image, const sift_params_t* params,
                    uint64_t tag) {
    if (!image) return SIFT_ERR_INVALID;
    const sift_color_image_t c{image->data, image->width, image->height, image->stride, SIFT_PIXEL_GRAY8};
    return sh::submit_frame(ctx, &c, params, tag);
}

int sift_hip_submit_color(sift_hip_ctx* ctx, const sift_color_image_t* image,
                          const sift_params_t* params, uint64_t tag) {
    return sh::submit_frame(ctx, image, params, tag);
}

int sift_hip_poll(sift_hip_ctx* ctx, sift_completion_t* out, size_t cap, size_t* n, int timeout_ms) {
    if (n) *n = 0;
    if (!ctx || !out || !n) return SIFT_ERR_INVALID;
    Context* C = reinterpret_cast<Context*>(ctx);
    Server* S = C->server;
    if (!S) return SIFT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && timeout_ms != 0) {
        if (timeout_ms < 0) S->cv_done.wait(lk, [&] { return !S->done.empty(); });
        else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                 [&] { return !S->done.empty(); });
    }
    size_t k = 0;
    while (k < cap && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    *n = k;
    return k ? SIFT_OK : SIFT_ERR_TIMEOUT;
}

int sift_hip_server_stop(sift_hip_ctx* ctx) {
    if (!ctx) return SIFT_ERR_INVALID;
    Context* C = reinterpret_cast<Context*>(ctx);
    Server* S = C->server;
    if (!S) return SIFT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& c : S->done) sift_result_free(&c.result);
    for (Context* L : S->lanes) {
        L->release();
        delete L;
    }
    delete S;
    C->server = nullptr;
    return SIFT_OK;
}

}  // extern "C"
