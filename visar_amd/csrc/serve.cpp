// serve.cpp -- the served front end's host runtime (include/visar_serve.h).
//
// Threads: the callers (submit / poll), one dispatcher, one completer and a
// pool of connected-components workers.
//   submit    copies the frame into the stream's pinned staging slot and
//             queues it on the stream (blocking while the stream has
//             queue_depth frames unfinished).
//   dispatcher while fewer than max_inflight batches are on the GPU, takes
//             the oldest queued frame of every stream that has one (round
//             robin, up to max_batch), enqueues their uploads, one fused
//             launch and the mask downloads on the server's stream, and
//             records an event. Frames keep arriving meanwhile; the next
//             batch takes whatever is queued when the GPU has room
//             (continuous batching). A stream's frames run in order: frame
//             f reads the model written by frame f - 1, and batches execute
//             in stream order.
//   completer waits for each batch's event and hands its masks to the
//             workers.
//   workers   label the mask (8-connectivity), hash mask and statistics,
//             copy the mask out if asked, post the result and free the
//             stream's slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/visar_serve.h"
#include "components.hpp"
#include "runtime.hpp"
#include "serve.hpp"
#include "tables.hpp"

using namespace visar;

namespace {

struct Req {
    int stream = 0;
    int slot = 0;       // staging / host mask slot
    uint64_t tag = 0;
    uint64_t index = 0;
    double t_submit = 0;
    uint8_t *mask_out = nullptr;
    size_t mask_stride = 0;
    uint32_t batch_size = 0;
};

// A stream slot's life: kFree -> kAdmitting (admit makes its buffers
// outside the lock) -> kActive -> kReleasing (release waits for its frames,
// then frees the buffers outside the lock) -> kFree. Only kFree slots are
// admitted and only kActive ones take frames, so an admit never reuses a
// slot whose buffers are still being made or freed.
enum SlotState { kFree, kAdmitting, kActive, kReleasing };

struct Stream {
    SlotState state = kFree;
    bool active = false;  // the dispatcher takes its queued frames
    void *dev = nullptr;  // frame[2] | bg[2] | mask[2]
    ServeStream d{};
    std::vector<uint8_t *> stage, hmask;  // pinned, queue_depth each
    std::vector<int> free_slots;
    std::deque<Req> queued;
    // submitted counts a frame from the moment it takes a staging slot (its
    // copy into the slot may still be running), so release and drain wait
    // for it.
    uint64_t submitted = 0, dispatched = 0, finished = 0;
};

struct Batch {
    hipEvent_t ev = nullptr;
    std::vector<Req> reqs;
};

uint64_t mix64(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xff51afd7ed558ccdull;
    return h ^ (h >> 33);
}

}  // namespace

struct visar_serve {
    visar_serve_config cfg{};
    int device = 0;
    hipStream_t st = nullptr;
    long p8 = 0, p16 = 0;
    size_t per_stream = 0;
    int kx[5] = {0, 0, 0, 0, 0};
    ServeStream *dtable = nullptr;  // device table of every stream slot
    std::vector<Stream> streams;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_dispatch, cv_complete, cv_slots, cv_results, cv_work;
    std::deque<Batch> inflight;
    std::deque<visar_serve_result> results;
    std::deque<std::function<void()>> work;
    uint64_t used_bytes = 0, frames_done = 0, batches = 0, admitted = 0, rejected = 0, submitted = 0;
    uint32_t max_batch_seen = 0;
    size_t rr = 0;
    bool stop = false;
    int sticky = VISAR_OK;
    std::thread dispatcher, completer;
    std::vector<std::thread> workers;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int check_config(const visar_serve_config &c) {
    if (c.width == 0 || c.height == 0 || c.width > (1u << 15) || c.height > (1u << 15)) return VISAR_INVALID_ARGUMENT;
    if (c.shift < 0 || c.shift > 15) return VISAR_INVALID_ARGUMENT;
    if (c.mask_threshold < 0 || c.mask_threshold > 254) return VISAR_INVALID_ARGUMENT;
    if (c.max_streams < 1 || c.max_streams > 4096) return VISAR_INVALID_ARGUMENT;
    if (c.max_batch < 1 || c.max_batch > (uint32_t)kServeMaxBatch) return VISAR_INVALID_ARGUMENT;
    if (c.queue_depth < 1 || c.queue_depth > 64) return VISAR_INVALID_ARGUMENT;
    if (c.host_workers < 1 || c.host_workers > 64) return VISAR_INVALID_ARGUMENT;
    if (c.max_inflight < 1 || c.max_inflight > 8) return VISAR_INVALID_ARGUMENT;
    if (c.batch_window_us > 1000000) return VISAR_INVALID_ARGUMENT;
    return VISAR_OK;
}

void set_sticky(visar_serve *s, int rc) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->sticky == VISAR_OK) s->sticky = rc;
    s->stop = true;
    s->cv_dispatch.notify_all();
    s->cv_complete.notify_all();
    s->cv_slots.notify_all();
    s->cv_results.notify_all();
    s->cv_work.notify_all();
}

// Labels one returned mask and posts its result (worker thread).
void finish_frame(visar_serve *s, const Req &r) {
    Stream &sm = s->streams[(size_t)r.stream];
    const uint32_t w = s->cfg.width, h = s->cfg.height;
    const uint8_t *m = sm.hmask[(size_t)r.slot];
    std::vector<CompStats> stats;
    const size_t n = label_components(m, w, nullptr, 0, w, h, true, stats);
    uint64_t hsh = 0x243f6a8885a308d3ull;
    const size_t bytes = (size_t)w * h;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t v;
        std::memcpy(&v, m + i, 8);
        hsh = mix64(hsh, v);
    }
    for (; i < bytes; i++) hsh = mix64(hsh, m[i]);
    uint32_t objects = 0;
    for (size_t k = 1; k < n; k++) {
        const CompStats &c = stats[k];
        hsh = mix64(hsh, c.x0 | (c.y0 << 16) | (c.x1 << 32) | (c.y1 << 48));
        hsh = mix64(hsh, c.area);
        if (c.area >= s->cfg.min_area) objects++;
    }
    if (r.mask_out)
        for (uint32_t y = 0; y < h; y++) std::memcpy(r.mask_out + y * r.mask_stride, m + (size_t)y * w, w);
    visar_serve_result res{};
    res.stream_id = r.stream;
    res.status = VISAR_OK;
    res.tag = r.tag;
    res.frame_index = r.index;
    res.components = (uint32_t)n;
    res.objects = objects;
    res.checksum = hsh;
    res.t_submit = r.t_submit;
    res.batch_size = r.batch_size;
    std::lock_guard<std::mutex> lk(s->mu);
    res.t_done = s->now();
    s->results.push_back(res);
    sm.free_slots.push_back(r.slot);
    sm.finished++;
    s->frames_done++;
    s->cv_results.notify_all();
    s->cv_slots.notify_all();
}

void worker_loop(visar_serve *s) {
    for (;;) {
        std::function<void()> job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->work.empty(); });
            if (s->work.empty()) return;
            job = std::move(s->work.front());
            s->work.pop_front();
        }
        job();
    }
}

void dispatch_loop(visar_serve *s) {
    if (hipSetDevice(s->device) != hipSuccess) return set_sticky(s, VISAR_INTERNAL);
    const uint32_t w = s->cfg.width, h = s->cfg.height;
    for (;;) {
        Batch b;
        ServeLaunch L{};
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_dispatch.wait(lk, [&] {
                if (s->stop) return true;
                if (s->inflight.size() >= s->cfg.max_inflight) return false;
                for (const Stream &sm : s->streams)
                    if (sm.active && !sm.queued.empty()) return true;
                return false;
            });
            if (s->stop) return;
            // Batching window: with fewer than max_batch streams ready, wait
            // until max_batch are or the oldest ready frame has waited
            // batch_window_us.
            if (s->cfg.batch_window_us > 0) {
                auto ready = [&](double *oldest) {
                    size_t n = 0;
                    for (const Stream &sm : s->streams)
                        if (sm.active && !sm.queued.empty()) {
                            n++;
                            if (oldest) *oldest = std::min(*oldest, sm.queued.front().t_submit);
                        }
                    return n;
                };
                double oldest = 1e300;
                if (ready(&oldest) < s->cfg.max_batch) {
                    const double left = oldest + 1e-6 * s->cfg.batch_window_us - s->now();
                    if (left > 0)
                        s->cv_dispatch.wait_for(lk, std::chrono::duration<double>(left),
                                                [&] { return s->stop || ready(nullptr) >= s->cfg.max_batch; });
                }
                if (s->stop) return;
            }
            const size_t ns = s->streams.size();
            for (size_t k = 0; k < ns && b.reqs.size() < s->cfg.max_batch; k++) {
                const size_t i = (s->rr + k) % ns;
                Stream &sm = s->streams[i];
                if (!sm.active || sm.queued.empty()) continue;
                Req r = sm.queued.front();
                sm.queued.pop_front();
                r.index = sm.dispatched++;
                b.reqs.push_back(r);
            }
            s->rr = (s->rr + 1) % std::max<size_t>(1, ns);
        }
        L.streams = s->dtable;
        L.w = (int)w;
        L.h = (int)h;
        L.p8 = s->p8;
        L.p16 = s->p16;
        L.shift = s->cfg.shift;
        L.thr = s->cfg.threshold < 0 ? 0 : s->cfg.threshold > 255 ? 255 : s->cfg.threshold;
        L.mthr = s->cfg.mask_threshold;
        for (int i = 0; i < 5; i++) L.kx[i] = s->kx[i];
        L.n = (int)b.reqs.size();
        bool ok = true;
        for (size_t k = 0; k < b.reqs.size(); k++) {
            Req &r = b.reqs[k];
            r.batch_size = (uint32_t)b.reqs.size();
            Stream &sm = s->streams[(size_t)r.stream];
            const int par = (int)(r.index & 1);
            L.job[k] = (uint32_t)r.stream | ((uint32_t)par << 16) | ((r.index > 0 ? 1u : 0u) << 17);
            ok = ok && hipMemcpy2DAsync(sm.d.frame[par], (size_t)s->p8, sm.stage[(size_t)r.slot], w, w, h,
                                        hipMemcpyHostToDevice, s->st) == hipSuccess;
        }
        ok = ok && launch_serve(L, s->st) == OK;
        for (const Req &r : b.reqs) {
            Stream &sm = s->streams[(size_t)r.stream];
            const int par = (int)(r.index & 1);
            ok = ok && hipMemcpy2DAsync(sm.hmask[(size_t)r.slot], w, sm.d.mask[par], (size_t)s->p8, w, h,
                                        hipMemcpyDeviceToHost, s->st) == hipSuccess;
        }
        ok = ok && hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) == hipSuccess;
        ok = ok && hipEventRecord(b.ev, s->st) == hipSuccess;
        if (!ok) return set_sticky(s, VISAR_INTERNAL);
        std::lock_guard<std::mutex> lk(s->mu);
        s->batches++;
        s->max_batch_seen = std::max<uint32_t>(s->max_batch_seen, (uint32_t)b.reqs.size());
        s->inflight.push_back(std::move(b));
        s->cv_complete.notify_all();
    }
}

void complete_loop(visar_serve *s) {
    if (hipSetDevice(s->device) != hipSuccess) return set_sticky(s, VISAR_INTERNAL);
    for (;;) {
        hipEvent_t ev;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_complete.wait(lk, [&] { return s->stop || !s->inflight.empty(); });
            if (s->inflight.empty()) return;
            ev = s->inflight.front().ev;
        }
        if (hipEventSynchronize(ev) != hipSuccess) return set_sticky(s, VISAR_INTERNAL);
        std::vector<Req> reqs;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            reqs = std::move(s->inflight.front().reqs);
            s->inflight.pop_front();
            for (const Req &r : reqs) s->work.push_back([s, r] { finish_frame(s, r); });
            s->cv_work.notify_all();
            s->cv_dispatch.notify_all();
        }
        (void)hipEventDestroy(ev);
    }
}

void free_stream_memory(Stream &sm) {
    if (sm.dev) (void)hipFree(sm.dev);
    sm.dev = nullptr;
    for (uint8_t *p : sm.stage) (void)hipHostFree(p);
    for (uint8_t *p : sm.hmask) (void)hipHostFree(p);
    sm.stage.clear();
    sm.hmask.clear();
    sm.free_slots.clear();
}

}  // namespace

extern "C" {

void visar_serve_config_default(visar_serve_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->width = 1280;
    c->height = 720;
    c->shift = 4;
    c->threshold = 28;
    c->mask_threshold = 127;
    c->min_area = 250;
    c->hbm_budget = 0;
    c->max_streams = 64;
    c->max_batch = 64;
    c->queue_depth = 4;
    c->host_workers = 4;
    c->max_inflight = 2;
    c->batch_window_us = 0;
}

int visar_serve_create(const visar_serve_config *cfg, visar_serve **out) {
    if (!cfg || !out) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    *out = nullptr;
    if (check_config(*cfg) != VISAR_OK) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    int rc = ensure_runtime();
    if (rc) return rc;
    visar_serve *s = new (std::nothrow) visar_serve();
    if (!s) return fail(VISAR_OUT_OF_MEMORY, "OutOfMemory");
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    const uint32_t w = cfg->width, h = cfg->height;
    s->p8 = (long)((w + 255) / 256 * 256);
    s->p16 = (long)((2 * (size_t)w + 255) / 256 * 256);
    s->per_stream = 2 * (size_t)s->p8 * h * 2 + 2 * (size_t)s->p16 * h;
    uint32_t taps[kMaxTaps];
    gaussian_kernel_fixed(taps, 5, 0.0, 8);
    for (int i = 0; i < 5; i++) s->kx[i] = (int)taps[i];
    s->streams.resize(cfg->max_streams);
    bool ok = hipGetDevice(&s->device) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&s->dtable, sizeof(ServeStream) * cfg->max_streams) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        if (s->dtable) (void)hipFree(s->dtable);
        if (s->st) (void)hipStreamDestroy(s->st);
        delete s;
        return fail(VISAR_INTERNAL, "DeviceSetupFailed");
    }
    s->dispatcher = std::thread(dispatch_loop, s);
    s->completer = std::thread(complete_loop, s);
    for (uint32_t i = 0; i < cfg->host_workers; i++) s->workers.emplace_back(worker_loop, s);
    *out = s;
    return VISAR_OK;
}

int visar_serve_admit(visar_serve *s, int32_t *stream_id) {
    if (!s || !stream_id) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    size_t slot;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->sticky) return fail(s->sticky, "ServerFailed");
        for (slot = 0; slot < s->streams.size(); slot++)
            if (s->streams[slot].state == kFree) break;
        if (slot == s->streams.size() || (s->cfg.hbm_budget && s->used_bytes + s->per_stream > s->cfg.hbm_budget)) {
            s->rejected++;
            return fail(VISAR_OUT_OF_MEMORY, "HbmBudget");
        }
        s->used_bytes += s->per_stream;  // reserved while the buffers are made
        s->streams[slot].state = kAdmitting;
    }
    Stream &sm = s->streams[slot];
    const uint32_t w = s->cfg.width, h = s->cfg.height;
    bool ok = hipMalloc(&sm.dev, s->per_stream) == hipSuccess;
    if (ok) {
        uint8_t *p = static_cast<uint8_t *>(sm.dev);
        const size_t f = (size_t)s->p8 * h, b = (size_t)s->p16 * h;
        sm.d.frame[0] = p;
        sm.d.frame[1] = p + f;
        sm.d.bg[0] = reinterpret_cast<uint16_t *>(p + 2 * f);
        sm.d.bg[1] = reinterpret_cast<uint16_t *>(p + 2 * f + b);
        sm.d.mask[0] = p + 2 * f + 2 * b;
        sm.d.mask[1] = p + 3 * f + 2 * b;
        ok = hipMemcpy(s->dtable + slot, &sm.d, sizeof(ServeStream), hipMemcpyHostToDevice) == hipSuccess;
    }
    for (uint32_t q = 0; ok && q < s->cfg.queue_depth; q++) {
        void *a = nullptr, *c = nullptr;
        ok = hipHostMalloc(&a, (size_t)w * h, hipHostMallocDefault) == hipSuccess;
        if (ok) sm.stage.push_back(static_cast<uint8_t *>(a));
        ok = ok && hipHostMalloc(&c, (size_t)w * h, hipHostMallocDefault) == hipSuccess;
        if (ok) {
            sm.hmask.push_back(static_cast<uint8_t *>(c));
            sm.free_slots.push_back((int)q);
        }
    }
    if (!ok) {
        (void)hipGetLastError();
        free_stream_memory(sm);
        std::lock_guard<std::mutex> lk(s->mu);
        s->used_bytes -= s->per_stream;
        s->rejected++;
        sm.state = kFree;
        return fail(VISAR_OUT_OF_MEMORY, "OutOfMemory");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    sm.state = kActive;
    sm.active = true;
    sm.submitted = sm.dispatched = sm.finished = 0;
    sm.queued.clear();
    s->admitted++;
    *stream_id = (int32_t)slot;
    return VISAR_OK;
}

int visar_serve_submit(visar_serve *s, int32_t id, const uint8_t *frame, size_t stride, uint64_t tag,
                       uint8_t *mask_out, size_t mask_stride) {
    if (!s || !frame) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    const uint32_t w = s->cfg.width, h = s->cfg.height;
    if (stride < w || (mask_out && mask_stride < w)) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    const double t_arrive = s->now();
    int slot;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (id < 0 || (size_t)id >= s->streams.size() || s->streams[(size_t)id].state != kActive)
            return fail(VISAR_INVALID_ARGUMENT, "UnknownStream");
        Stream &sm = s->streams[(size_t)id];
        s->cv_slots.wait(lk, [&] { return s->stop || sm.state != kActive || !sm.free_slots.empty(); });
        if (s->sticky) return fail(s->sticky, "ServerFailed");
        if (s->stop) return fail(VISAR_INVALID_ARGUMENT, "ServerStopped");
        if (sm.state != kActive) return fail(VISAR_INVALID_ARGUMENT, "UnknownStream");
        slot = sm.free_slots.back();
        sm.free_slots.pop_back();
        // Pending from here: release and drain wait for this frame.
        sm.submitted++;
        s->submitted++;
    }
    Stream &sm = s->streams[(size_t)id];
    uint8_t *dst = sm.stage[(size_t)slot];
    for (uint32_t y = 0; y < h; y++) std::memcpy(dst + (size_t)y * w, frame + y * stride, w);
    std::lock_guard<std::mutex> lk(s->mu);
    Req r;
    r.stream = id;
    r.slot = slot;
    r.tag = tag;
    r.t_submit = t_arrive;
    r.mask_out = mask_out;
    r.mask_stride = mask_stride;
    sm.queued.push_back(r);
    s->cv_dispatch.notify_all();
    return VISAR_OK;
}

int visar_serve_poll(visar_serve *s, visar_serve_result *out, size_t max, size_t *n, int wait_ms) {
    if (!s || !n || (!out && max)) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait_ms > 0 && s->results.empty())
        s->cv_results.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return s->stop || !s->results.empty(); });
    size_t k = 0;
    while (k < max && !s->results.empty()) {
        out[k++] = s->results.front();
        s->results.pop_front();
    }
    *n = k;
    if (s->sticky) return fail(s->sticky, "ServerFailed");
    return VISAR_OK;
}

int visar_serve_drain(visar_serve *s) {
    if (!s) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_results.wait(lk, [&] { return s->stop || s->frames_done == s->submitted; });
    if (s->sticky) return fail(s->sticky, "ServerFailed");
    return VISAR_OK;
}

int visar_serve_release(visar_serve *s, int32_t id) {
    if (!s) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    Stream gone;  // the slot's buffers, detached under the lock and freed outside it
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (id < 0 || (size_t)id >= s->streams.size() || s->streams[(size_t)id].state != kActive)
            return fail(VISAR_INVALID_ARGUMENT, "UnknownStream");
        Stream &sm = s->streams[(size_t)id];
        sm.state = kReleasing;   // no new frames; the queued ones still run
        s->cv_slots.notify_all();
        s->cv_results.wait(lk, [&] { return s->stop || s->sticky != VISAR_OK || sm.finished == sm.submitted; });
        sm.active = false;
        std::swap(gone.dev, sm.dev);
        gone.stage.swap(sm.stage);
        gone.hmask.swap(sm.hmask);
        gone.free_slots.swap(sm.free_slots);
        sm.queued.clear();
    }
    free_stream_memory(gone);
    std::lock_guard<std::mutex> lk(s->mu);
    s->streams[(size_t)id].state = kFree;
    s->used_bytes -= s->per_stream;
    if (s->sticky) return fail(s->sticky, "ServerFailed");
    return VISAR_OK;
}

int visar_serve_get_stats(visar_serve *s, visar_serve_stats *st) {
    if (!s || !st) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof *st);
    st->frames_done = s->frames_done;
    st->batches = s->batches;
    st->admitted = s->admitted;
    st->rejected = s->rejected;
    st->device_bytes = s->used_bytes;
    st->bytes_per_stream = s->per_stream;
    st->max_batch_seen = s->max_batch_seen;
    for (const Stream &sm : s->streams) st->active_streams += sm.active ? 1u : 0u;
    return VISAR_OK;
}

int visar_serve_time_kernel(visar_serve *s, uint32_t reps, double *ms_per_launch, uint32_t *frames_per_launch) {
    if (!s || !ms_per_launch || reps == 0) return fail(VISAR_INVALID_ARGUMENT, "InvalidArgument");
    ServeLaunch L{};
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->sticky) return fail(s->sticky, "ServerFailed");
        if (s->frames_done != s->submitted || !s->inflight.empty()) return fail(VISAR_INVALID_ARGUMENT, "ServerBusy");
        for (size_t i = 0; i < s->streams.size() && L.n < kServeMaxBatch; i++)
            if (s->streams[i].active) L.job[L.n++] = (uint32_t)i | (1u << 17);
    }
    if (L.n == 0) return fail(VISAR_INVALID_ARGUMENT, "NoStreams");
    L.streams = s->dtable;
    L.w = (int)s->cfg.width;
    L.h = (int)s->cfg.height;
    L.p8 = s->p8;
    L.p16 = s->p16;
    L.shift = s->cfg.shift;
    L.thr = s->cfg.threshold < 0 ? 0 : s->cfg.threshold > 255 ? 255 : s->cfg.threshold;
    L.mthr = s->cfg.mask_threshold;
    for (int i = 0; i < 5; i++) L.kx[i] = s->kx[i];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    ok = ok && launch_serve(L, s->st) == OK && hipEventRecord(e0, s->st) == hipSuccess;
    for (uint32_t r = 0; ok && r < reps; r++) ok = launch_serve(L, s->st) == OK;
    ok = ok && hipEventRecord(e1, s->st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!ok) return fail(VISAR_INTERNAL, "KernelFailed");
    *ms_per_launch = ms / reps;
    if (frames_per_launch) *frames_per_launch = (uint32_t)L.n;
    return VISAR_OK;
}

double visar_serve_now(const visar_serve *s) { return s ? s->now() : 0.0; }

void visar_serve_free(visar_serve *s) {
    if (!s) return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_results.wait(lk, [&] { return s->sticky != VISAR_OK || s->frames_done == s->submitted; });
        s->stop = true;
        s->cv_dispatch.notify_all();
        s->cv_complete.notify_all();
        s->cv_work.notify_all();
        s->cv_slots.notify_all();
        s->cv_results.notify_all();
    }
    if (s->dispatcher.joinable()) s->dispatcher.join();
    if (s->completer.joinable()) s->completer.join();
    for (std::thread &t : s->workers)
        if (t.joinable()) t.join();
    if (s->st) (void)hipStreamSynchronize(s->st);
    for (Batch &b : s->inflight)
        if (b.ev) (void)hipEventDestroy(b.ev);
    for (Stream &sm : s->streams) free_stream_memory(sm);
    if (s->dtable) (void)hipFree(s->dtable);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

}  // extern "C"
