// capi.cpp -- the C ABI of include/visar.h and include/visar_gpu.h.
//
// Mirrors the reference's C-ABI layer (src/capi.zig, src/capi_calib3d.zig):
// descriptor validation (capi.zig:123-132), error names copied into a
// thread-local buffer (capi.zig:59-75), depth dispatch (capi.zig:136-144)
// and per-call argument checks. Underneath, every call resolves its images
// to device views: device buffers are used in place, host memory is staged
// through pooled device buffers with 2-D copies on the call's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/visar.h"
#include "../../include/visar_gpu.h"
#include "kernels.hpp"
#include "runtime.hpp"
#include "undistort.hpp"
#include "tables.hpp"

namespace visar {

// ------------------------------------------------------------- errors

static thread_local char g_last_error[256] = "";

int fail(int status, const char *name) {
    std::snprintf(g_last_error, sizeof g_last_error, "%s", name);
    return status;
}

// ------------------------------------------------------------ runtime

static std::once_flag g_init_once;
static int g_init_status = INTERNAL;
static hipStream_t g_own_stream = nullptr;
static thread_local hipStream_t t_stream = nullptr;
static thread_local bool t_stream_set = false;
static thread_local int t_sync = 1;

int ensure_runtime() {
    std::call_once(g_init_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            g_init_status = INTERNAL;
            return;
        }
        if (hipStreamCreateWithFlags(&g_own_stream, hipStreamNonBlocking) != hipSuccess) {
            g_init_status = INTERNAL;
            return;
        }
        g_init_status = OK;
    });
    if (g_init_status != OK) return fail(INTERNAL, "NoDevice");
    return OK;
}

hipStream_t current_stream() { return t_stream_set ? t_stream : g_own_stream; }
bool sync_mode() { return t_sync != 0; }

// Buffers are reused by size class: a call takes the smallest cached buffer
// that fits. A call that waited for its stream returns its buffers at once;
// one that only queued work returns them behind an event recorded on its
// stream, and they become free when that event has passed, so queued calls
// never wait for the GPU to reuse memory. The same pool type serves device
// buffers and pinned host staging buffers.
struct DeviceAlloc {
    static hipError_t alloc(void **p, size_t n) { return hipMalloc(p, n); }
    static void release(void *p) { (void)hipFree(p); }
};
struct PinnedAlloc {
    static hipError_t alloc(void **p, size_t n) { return hipHostMalloc(p, n, hipHostMallocDefault); }
    static void release(void *p) { (void)hipHostFree(p); }
};

template <typename A>
class BufferPool {
  public:
    void *get(size_t bytes) {
        {
            std::lock_guard<std::mutex> g(mu_);
            reap_locked();
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= bytes * 2 + (1 << 20)) {
                void *p = it->second;
                sizes_[p] = it->first;
                free_.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        if (A::alloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (A::alloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        std::lock_guard<std::mutex> g(mu_);
        sizes_[p] = bytes;
        return p;
    }
    void put(void *p) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        put_locked(p);
    }
    // Returns `bufs` once the work queued on `st` so far has finished.
    void put_after(const std::vector<void *> &bufs, hipStream_t st) {
        if (bufs.empty()) return;
        std::lock_guard<std::mutex> g(mu_);
        hipEvent_t ev = nullptr;
        if (!events_.empty()) {
            ev = events_.back();
            events_.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
        }
        if (!ev || hipEventRecord(ev, st) != hipSuccess) {
            // No event: wait for the stream instead.
            (void)hipGetLastError();
            if (ev) events_.push_back(ev);
            (void)hipStreamSynchronize(st);
            for (void *p : bufs) put_locked(p);
            return;
        }
        deferred_.push_back(Deferred{ev, bufs});
    }
    void trim() {
        std::lock_guard<std::mutex> g(mu_);
        reap_locked();
        for (auto &kv : free_) A::release(kv.second);
        free_.clear();
    }

  private:
    struct Deferred {
        hipEvent_t ev;
        std::vector<void *> bufs;
    };
    void put_locked(void *p) {
        auto it = sizes_.find(p);
        if (it == sizes_.end()) return;
        free_.emplace(it->second, p);
        sizes_.erase(it);
    }
    void reap_locked() {
        while (!deferred_.empty()) {
            const hipError_t q = hipEventQuery(deferred_.front().ev);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                break;
            }
            for (void *p : deferred_.front().bufs) put_locked(p);
            events_.push_back(deferred_.front().ev);
            deferred_.pop_front();
        }
    }
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> sizes_;
    std::deque<Deferred> deferred_;
    std::vector<hipEvent_t> events_;
};

static BufferPool<DeviceAlloc> &pool() {
    static auto *p = new BufferPool<DeviceAlloc>();  // never destroyed: outlives static HIP teardown
    return *p;
}
static BufferPool<PinnedAlloc> &pinned_pool() {
    static auto *p = new BufferPool<PinnedAlloc>();
    return *p;
}

void *pool_get(size_t bytes) { return pool().get(bytes); }
void pool_put(void *p) { pool().put(p); }

bool is_device_ptr(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// ------------------------------------------------------------- images

static inline size_t row_bytes(const visar_image *c) { return (size_t)c->width * c->channels * depth_size(c->depth); }
static inline size_t span_bytes(const visar_image *c) { return (c->height - 1) * c->stride + row_bytes(c); }

// capi.zig:123-132 typed() + image.zig:71-77, 172-176.
int check_image(const visar_image *c, uint32_t depth) {
    if (!c || !c->data) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (c->depth != depth) return fail(UNSUPPORTED, "Unsupported");
    if (c->channels == 0 || c->channels > 512) return fail(INVALID_ARGUMENT, "InvalidArgument");
    size_t es = depth_size(depth);
    if (es == 0 || c->stride % es != 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (c->width == 0 || c->height == 0) return fail(INVALID_ARGUMENT, "InvalidSize");
    if (c->stride < row_bytes(c)) return fail(INVALID_ARGUMENT, "InvalidSize");
    if ((uint64_t)c->width * c->channels > (1u << 30) || c->height > (1u << 30)) return fail(INVALID_ARGUMENT, "InvalidSize");
    return OK;
}

Call::~Call() {
    // Buffers of a call that queued work on a stream (the default stream
    // included) return behind that work.
    if (waited_ || !begun_) {
        for (void *p : owned_) pool_put(p);
        for (void *p : pinned_) pinned_pool().put(p);
    } else {
        pool().put_after(owned_, st_);
        pinned_pool().put_after(pinned_, st_);
    }
}

int Call::begin() {
    int rc = ensure_runtime();
    if (rc) return rc;
    st_ = current_stream();
    begun_ = true;
    return OK;
}

DImg Call::view(const visar_image *c, void *p, size_t stride) {
    DImg d;
    d.p = static_cast<uint8_t *>(p);
    d.w = (int)c->width;
    d.h = (int)c->height;
    d.cn = (int)c->channels;
    d.stride = (long)stride;
    return d;
}

void *Call::scratch(size_t bytes) {
    void *p = pool_get(bytes ? bytes : 1);
    if (p) owned_.push_back(p);
    return p;
}

void *Call::upload(const void *src, size_t bytes) {
    void *dev = scratch(bytes);
    void *pin = pinned_pool().get(bytes ? bytes : 1);
    if (!dev || !pin) {
        if (pin) pinned_pool().put(pin);
        return nullptr;
    }
    pinned_.push_back(pin);
    std::memcpy(pin, src, bytes);
    if (hipMemcpyAsync(dev, pin, bytes, hipMemcpyHostToDevice, st_) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return dev;
}

void *Call::pinned(size_t bytes) {
    void *p = pinned_pool().get(bytes ? bytes : 1);
    if (p) pinned_.push_back(p);
    return p;
}

// Rows of `rb` bytes from src (stride ss) to dst (stride ds) on the CPU.
static void copy_rows(void *dst, size_t ds, const void *src, size_t ss, size_t rb, size_t h) {
    if (ds == rb && ss == rb) {
        std::memcpy(dst, src, rb * h);
        return;
    }
    for (size_t y = 0; y < h; y++)
        std::memcpy(static_cast<uint8_t *>(dst) + y * ds, static_cast<const uint8_t *>(src) + y * ss, rb);
}

// A host image's rows into device buffer `buf` (pitch), through pinned
// staging rows on the call's stream.
static int stage_up(Call &call, void *buf, size_t pitch, const visar_image *c, size_t rb) {
    void *pin = call.pinned(rb * c->height);
    if (!pin) return fail(OUT_OF_MEMORY, "OutOfMemory");
    copy_rows(pin, rb, c->data, c->stride, rb, c->height);
    if (hipMemcpy2DAsync(buf, pitch, pin, rb, rb, c->height, hipMemcpyHostToDevice, call.stream()) != hipSuccess)
        return fail(INTERNAL, "CopyFailed");
    return OK;
}

int Call::input(const visar_image *c, DImg &out) {
    if (is_device_ptr(c->data)) {
        out = view(c, c->data, c->stride);
        dev_ranges_.push_back({(const uint8_t *)c->data, span_bytes(c)});
        return OK;
    }
    size_t rb = row_bytes(c);
    size_t pitch = (rb + 255) & ~(size_t)255;
    void *buf = scratch(pitch * c->height);
    if (!buf) return fail(OUT_OF_MEMORY, "OutOfMemory");
    const int rc = stage_up(*this, buf, pitch, c, rb);
    if (rc) return rc;
    staged_ = true;
    out = view(c, buf, pitch);
    return OK;
}

int Call::output(const visar_image *c, DImg &out) {
    if (is_device_ptr(c->data)) {
        out = view(c, c->data, c->stride);
        // An output that overlaps a device input is computed into a
        // temporary and copied over afterwards (kernels read neighbours).
        const uint8_t *p = (const uint8_t *)c->data;
        size_t n = span_bytes(c);
        bool overlap = false;
        for (auto &r : dev_ranges_)
            if (p < r.first + r.second && r.first < p + n) overlap = true;
        if (!overlap) return OK;
        size_t rb = row_bytes(c);
        size_t pitch = (rb + 255) & ~(size_t)255;
        void *buf = scratch(pitch * c->height);
        if (!buf) return fail(OUT_OF_MEMORY, "OutOfMemory");
        pending_.push_back({c, buf, pitch, true, nullptr});
        out = view(c, buf, pitch);
        return OK;
    }
    size_t rb = row_bytes(c);
    size_t pitch = (rb + 255) & ~(size_t)255;
    void *buf = scratch(pitch * c->height);
    if (!buf) return fail(OUT_OF_MEMORY, "OutOfMemory");
    staged_ = true;
    pending_.push_back({c, buf, pitch, false, nullptr});
    out = view(c, buf, pitch);
    return OK;
}

int Call::output_keep(const visar_image *c, DImg &out) {
    int rc = output(c, out);
    if (rc || out.p == c->data) return rc;
    if (!is_device_ptr(c->data)) return stage_up(*this, out.p, (size_t)out.stride, c, row_bytes(c));
    if (hipMemcpy2DAsync(out.p, (size_t)out.stride, c->data, c->stride, row_bytes(c), c->height, hipMemcpyDeviceToDevice,
                         st_) != hipSuccess)
        return fail(INTERNAL, "CopyFailed");
    return OK;
}

int Call::temp_like(const DImg &like, size_t elem, DImg &out) {
    size_t rb = (size_t)like.w * like.cn * elem;
    size_t pitch = (rb + 255) & ~(size_t)255;
    void *buf = scratch(pitch * like.h);
    if (!buf) return fail(OUT_OF_MEMORY, "OutOfMemory");
    out = like;
    out.p = (uint8_t *)buf;
    out.stride = (long)pitch;
    return OK;
}

int Call::finish(int launch_status) {
    if (launch_status != OK) {
        (void)hipStreamSynchronize(st_);
        waited_ = true;
        return launch_status == UNSUPPORTED ? fail(UNSUPPORTED, "Unsupported") : fail(INTERNAL, "LaunchFailed");
    }
    for (auto &pd : pending_) {
        const visar_image *c = pd.desc;
        const size_t rb = row_bytes(c);
        if (pd.device_target) {
            if (hipMemcpy2DAsync(c->data, c->stride, pd.buf, pd.pitch, rb, c->height, hipMemcpyDeviceToDevice, st_) !=
                hipSuccess)
                return fail(INTERNAL, "CopyFailed");
            continue;
        }
        // Host targets: device rows to pinned rows here, to the caller's
        // memory on the CPU once the stream has drained.
        pd.pin = pinned(rb * c->height);
        if (!pd.pin) return fail(OUT_OF_MEMORY, "OutOfMemory");
        if (hipMemcpy2DAsync(pd.pin, rb, pd.buf, pd.pitch, rb, c->height, hipMemcpyDeviceToHost, st_) != hipSuccess)
            return fail(INTERNAL, "CopyFailed");
    }
    if (staged_ || sync_mode()) {
        if (hipStreamSynchronize(st_) != hipSuccess) return fail(INTERNAL, "KernelFailed");
        waited_ = true;
    }
    for (auto &pd : pending_)
        if (pd.pin) copy_rows(pd.desc->data, pd.desc->stride, pd.pin, row_bytes(pd.desc), row_bytes(pd.desc), pd.desc->height);
    return OK;
}

// threshold.zig:184-207 getThreshVal_Otsu_8u
double threshold_otsu(const uint32_t *hist, double total) {
    const double scale = 1.0 / total;
    const double eps = 2.220446049250313e-16;
    double mu = 0;
    for (int i = 0; i < 256; i++) mu += (double)i * (double)hist[i];
    mu *= scale;
    double mu1 = 0, q1 = 0, max_sigma = 0, max_val = 0;
    for (int i = 0; i < 256; i++) {
        double p_i = (double)hist[i] * scale;
        mu1 *= q1;
        q1 += p_i;
        double q2 = 1.0 - q1;
        if (std::min(q1, q2) < eps || std::max(q1, q2) > 1.0 - eps) continue;
        mu1 = (mu1 + (double)i * p_i) / q1;
        double mu2 = (mu - q1 * mu1) / q2;
        double sigma = q1 * q2 * (mu1 - mu2) * (mu1 - mu2);
        if (sigma > max_sigma) max_sigma = sigma, max_val = (double)i;
    }
    return max_val;
}

// threshold.zig:210-253 getThreshVal_Triangle_8u
double threshold_triangle(const uint32_t *hist_in) {
    int left = 0, right = 0, max_ind = 0;
    uint32_t mx = 0;
    for (int i = 0; i < 256; i++)
        if (hist_in[i] > 0) {
            left = i;
            break;
        }
    if (left > 0) left -= 1;
    for (int i = 255; i > 0; i--)
        if (hist_in[i] > 0) {
            right = i;
            break;
        }
    if (right < 255) right += 1;
    for (int i = 0; i < 256; i++)
        if (hist_in[i] > mx) mx = hist_in[i], max_ind = i;
    uint32_t h[256];
    std::memcpy(h, hist_in, sizeof h);
    bool flip = false;
    if (max_ind - left < right - max_ind) {
        flip = true;
        std::reverse(h, h + 256);
        left = 255 - right;
        max_ind = 255 - max_ind;
    }
    int thresh = left;
    double a = (double)mx, b = (double)(left - max_ind), dist = 0;
    for (int j = left + 1; j <= max_ind; j++) {
        double t = a * (double)j + b * (double)h[j];
        if (t > dist) dist = t, thresh = j;
    }
    thresh -= 1;
    if (flip) thresh = 255 - thresh;
    return (double)thresh;
}

// threshold.zig:25-82: the per-depth form of (thresh, maxval) the kernel
// applies; for u8 a 256-entry table of every type's result.
void threshold_setup(ThreshParams &p, uint32_t T, int kind, double thresh, double maxval) {
    p.depth = (int)T;
    p.kind = kind;
    if (T == 0) {
        // threshold.zig:40-60: floored threshold; the LUT covers every case.
        double fl = std::floor(thresh);
        long ti = fl < -2147483648.0 ? -2147483648L : fl > 2147483647.0 ? 2147483647L : (long)fl;
        int m = sat_cast<uint8_t>(maxval);
        int tc = ti < 0 ? 0 : ti > 255 ? 255 : (int)ti;
        p.ti = (int)(ti < -1 ? -1 : ti > 256 ? 256 : ti);
        p.mi = m;
        for (int i = 0; i < 256; i++) {
            bool gt = i > ti;
            int v;
            switch (kind) {
            case 0: v = gt ? m : 0; break;
            case 1: v = gt ? 0 : m; break;
            case 2: v = gt ? tc : i; break;
            case 3: v = gt ? i : 0; break;
            default: v = gt ? 0 : i; break;
            }
            p.lut[i] = (uint8_t)v;
        }
    } else if (T == 5) {
        p.t = (double)(float)thresh;
        p.m = (double)(float)maxval;
    } else {
        // threshold.zig:62-81: thresholds below the range make every value
        // "above" (tozero copies, binary fills maxval, the rest give 0).
        double lo = T == 2 ? 0 : -32768.0;
        p.below = std::floor(thresh) < lo;
        p.t = T == 2 ? (double)sat_cast<uint16_t>(std::floor(thresh)) : (double)sat_cast<int16_t>(std::floor(thresh));
        p.m = T == 2 ? (double)sat_cast<uint16_t>(maxval) : (double)sat_cast<int16_t>(maxval);
    }
}

}  // namespace visar

using namespace visar;

// ================================================================ C ABI

extern "C" {

const char *visar_version(void) { return "0.1.0"; }

const char *visar_build_info(void) { return "visar_amd gfx950 (HIP)"; }

const char *visar_last_error(void) { return g_last_error; }

static std::atomic<int> g_threads{0};

void visar_set_num_threads(int n) { g_threads.store(n <= 0 ? 0 : n); }

int visar_get_num_threads(void) {
    int n = g_threads.load();
    if (n > 0) return n;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

int visar_gpu_set_stream(void *s) {
    t_stream = (hipStream_t)s;
    t_stream_set = true;
    return VISAR_OK;
}

int visar_gpu_reset_stream(void) {
    t_stream = nullptr;
    t_stream_set = false;
    return VISAR_OK;
}

void *visar_gpu_get_stream(void) {
    if (ensure_runtime()) return nullptr;
    return (void *)current_stream();
}

int visar_gpu_set_sync(int sync) {
    t_sync = sync != 0;
    return VISAR_OK;
}

int visar_gpu_synchronize(void) {
    int rc = ensure_runtime();
    if (rc) return rc;
    return hipStreamSynchronize(current_stream()) == hipSuccess ? VISAR_OK : fail(INTERNAL, "KernelFailed");
}

int visar_gpu_malloc(size_t bytes, void **out) {
    int rc = ensure_runtime();
    if (rc) return rc;
    if (!out) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(OUT_OF_MEMORY, "OutOfMemory");
    }
    return VISAR_OK;
}

int visar_gpu_free(void *p) {
    if (!p) return VISAR_OK;
    return hipFree(p) == hipSuccess ? VISAR_OK : fail(INVALID_ARGUMENT, "InvalidArgument");
}

int visar_gpu_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    int rc = ensure_runtime();
    if (rc) return rc;
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (kind < 1 || kind > 3 || !dst || !src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (hipMemcpyAsync(dst, src, bytes, k, current_stream()) != hipSuccess) return fail(INTERNAL, "CopyFailed");
    return hipStreamSynchronize(current_stream()) == hipSuccess ? VISAR_OK : fail(INTERNAL, "CopyFailed");
}

int visar_gpu_set_tuning(const char *key, int value) {
    if (!key) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!set_tuning(key, value)) return fail(INVALID_ARGUMENT, "InvalidArgument");
    return VISAR_OK;
}

int visar_gpu_get_tuning(const char *key, int *value) {
    if (!key || !value) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!get_tuning(key, value)) return fail(INVALID_ARGUMENT, "InvalidArgument");
    return VISAR_OK;
}

int visar_gpu_device_info(char *buf, size_t n) {
    int rc = ensure_runtime();
    if (rc) return rc;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(INTERNAL, "NoDevice");
    if (buf && n) std::snprintf(buf, n, "%s %d CUs", prop.gcnArchName, prop.multiProcessorCount);
    return VISAR_OK;
}

// ------------------------------------------------------------ filters

static bool valid_depth(uint32_t d) { return d == 0 || d == 2 || d == 3 || d == 5; }

static bool same_shape(const visar_image *a, const visar_image *b) {
    return a->width == b->width && a->height == b->height && a->channels == b->channels;
}

// capi.zig:197-206 + filter.zig:498-607
// filter.zig:498-607 gaussianBlur's taps for images of depth `T` (8.8 / 16.16
// fixed point for u8 / u16, f32 otherwise), sizes from sigma where 0, and a
// 1-pixel side reduced to one tap for non-constant borders. The caller sets
// the source and destination depths.
static int gaussian_params(SepParams &p, uint32_t T, int kw, int kh, double sx, double sy, int border,
                           const visar_image *src) {
    double sy_in = sy > 0 ? sy : sx;
    if (kw == 0) kw = gaussian_ksize(sx, T == 0);
    if (kh == 0) kh = gaussian_ksize(sy_in, T == 0);
    if (kw <= 0 || kh <= 0 || kw % 2 == 0 || kh % 2 == 0 || kw > 127 || kh > 127)
        return fail(INVALID_ARGUMENT, "InvalidKernel");
    if (border != BORDER_CONSTANT) {
        if (src->width == 1) kw = 1;
        if (src->height == 1) kh = 1;
    }
    std::memset(&p, 0, sizeof p);
    p.sdepth = p.ddepth = (int)T;
    p.nkx = kw;
    p.nky = kh;
    p.border = border;
    p.scale = 1.0f;
    if (T == 0 || T == 2) {
        uint32_t fx[kMaxTaps], fy[kMaxTaps];
        gaussian_kernel_fixed(fx, kw, sx, T == 0 ? 8 : 16);
        gaussian_kernel_fixed(fy, kh, sy_in, T == 0 ? 8 : 16);
        for (int i = 0; i < kw; i++) p.kxi[i] = (int32_t)fx[i];
        for (int i = 0; i < kh; i++) p.kyi[i] = (int32_t)fy[i];
        p.fin = T == 0 ? FIN_FIX8 : FIN_FIX16;
    } else {
        double kxd[kMaxTaps], kyd[kMaxTaps];
        gaussian_kernel(kxd, kw, sx);
        gaussian_kernel(kyd, kh, sy_in);
        for (int i = 0; i < kw; i++) p.kxf[i] = (float)kxd[i];
        for (int i = 0; i < kh; i++) p.kyf[i] = (float)kyd[i];
        p.f32mode = 1;
        p.fin = FIN_CAST;
        p.vsym = symmetry(p.kyf, kh);
        p.hsym = symmetry(p.kxf, kw);
        // filter.zig:586-600: f32 images with OpenCV's 3x3 / 5x5 tables take
        // the compile-time kernel path (symmetric body, tail summed from 0).
        static const float t3[3] = {0.25f, 0.5f, 0.25f};
        static const float t5[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
        if (T == 5 && kw == kh && std::equal(p.kxf, p.kxf + kw, p.kyf) &&
            ((kw == 3 && std::equal(p.kxf, p.kxf + 3, t3)) || (kw == 5 && std::equal(p.kxf, p.kxf + 5, t5)))) {
            p.const_kernel = 1;
            p.vsym = p.hsym = SYM_EVEN;
        }
        p.body = ((long)src->width * src->channels / 16) * 16;
    }
    return OK;
}

int visar_gaussian_blur(const visar_image *src, const visar_image *dst, int kw, int kh, double sx, double sy,
                        int border) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (kw < 0 || kh < 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    // Kernel checks before the shape check, as in filter.zig:498-520.
    double sy_in = sy > 0 ? sy : sx;
    const int kw0 = kw == 0 ? gaussian_ksize(sx, T == 0) : kw, kh0 = kh == 0 ? gaussian_ksize(sy_in, T == 0) : kh;
    if (kw0 <= 0 || kh0 <= 0 || kw0 % 2 == 0 || kh0 % 2 == 0 || kw0 > 127 || kh0 > 127)
        return fail(INVALID_ARGUMENT, "InvalidKernel");
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    SepParams p;
    if ((rc = gaussian_params(p, T, kw, kh, sx, sy, border, src))) return rc;
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.src)) || (rc = call.output(dst, p.dst))) return rc;
    return call.finish(launch_sep(p, call.stream()));
}

// capi.zig:208-217 visar_box_filter -> filter.zig:617-633 boxFilter: unit
// taps, exact integer sums (u8, u16) or the f32 engine's order, then
// f32(sum) * f32(scale) saturated (FinishScaled). i16 is rejected: the
// reference widens i16 samples into u32 accumulators with @intCast, which is
// undefined for negative samples. Kernels wider or taller than 127 (the
// reference allows 255) are not supported on the GPU.
int visar_box_filter(const visar_image *src, const visar_image *dst, int kw, int kh, int normalize, int border) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (kw <= 0 || kh <= 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (kw > 255 || kh > 255 || kw % 2 == 0 || kh % 2 == 0) return fail(INVALID_ARGUMENT, "InvalidKernel");
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    const bool wide = T == 3 || kw > 127 || kh > 127;
    SepParams p;
    std::memset(&p, 0, sizeof p);
    p.sdepth = p.ddepth = (int)T;
    p.nkx = kw;
    p.nky = kh;
    p.border = border;
    p.fin = FIN_SCALED;
    p.scale = (float)(normalize ? 1.0 / (double)(kw * kh) : 1.0);
    // The engine's tap arrays hold kMaxTaps; wider windows take the
    // two-pass kernels, which need no taps.
    if (!wide) {
        for (int i = 0; i < kw; i++) p.kxi[i] = 1, p.kxf[i] = 1.0f;
        for (int i = 0; i < kh; i++) p.kyi[i] = 1, p.kyf[i] = 1.0f;
    }
    if (T == 5) {
        p.f32mode = 1;
        p.vsym = p.hsym = SYM_EVEN;
        p.body = ((long)src->width * src->channels / 16) * 16;
    }
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.src)) || (rc = call.output(dst, p.dst))) return rc;
    if (wide) {
        void *tmp = call.scratch(box_wide_tmp_bytes(p.src));
        if (!tmp) return fail(OUT_OF_MEMORY, "OutOfMemory");
        return call.finish(launch_box_wide(p.src, p.dst, (int)T, kw, kh, border, p.scale, tmp, call.stream()));
    }
    if (T == 0 && src->channels <= 4)
        return call.finish(launch_box_u8(p.src, p.dst, kw, kh, border, p.scale, call.stream()));
    return call.finish(launch_sep(p, call.stream()));
}

// capi.zig:234-250 visar_pyr_down / visar_pyr_up -> imgproc/pyramid.zig:65-122
// pyrDown (dst ((w + 1) / 2, (h + 1) / 2), any border but constant) and
// :403-427 pyrUp (dst (2w, 2h), reflect-101 only); u8, u16, i16, f32.
static int pyramid_call(const visar_image *src, const visar_image *dst, int border, bool down) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (down ? border == BORDER_CONSTANT : border != BORDER_REFLECT101) return fail(UNSUPPORTED, "Unsupported");
    const uint64_t w = down ? (src->width + 1) / 2 : 2 * (uint64_t)src->width;
    const uint64_t h = down ? (src->height + 1) / 2 : 2 * (uint64_t)src->height;
    if (dst->channels != src->channels || dst->width != w || dst->height != h) return fail(SIZE_MISMATCH, "SizeMismatch");
    Call call;
    DImg s, d;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    return call.finish(launch_pyramid(s, d, (int)T, border, down, call.stream()));
}

int visar_pyr_down(const visar_image *src, const visar_image *dst, int border) { return pyramid_call(src, dst, border, true); }
int visar_pyr_up(const visar_image *src, const visar_image *dst, int border) { return pyramid_call(src, dst, border, false); }

// capi.zig:326-337 visar_adaptive_threshold -> threshold.zig:306-360: the
// block x block mean with a replicated border (boxFilter, or the f32
// Gaussian blur of the u8 source saturated back to u8: one pass of the f32
// engine, whose u8 loads convert exactly), then dst = s + idelta > m ? on :
// off with idelta = ceil(delta) (BINARY) or floor(delta) (BINARY_INV), the
// compare done by the kernel that computes m.
// Blocks above 127 are not supported on the GPU.
int visar_adaptive_threshold(const visar_image *src, const visar_image *dst, double maxval, int method, int type,
                             int block, double delta) {
    if (method != 0 && method != 1) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (type != 0 && type != 1) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (block < 3) return fail(INVALID_ARGUMENT, "InvalidArgument");
    int rc;
    if ((rc = check_image(src, 0)) || (rc = check_image(dst, 0))) return rc;
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    if (src->channels != 1) return fail(UNSUPPORTED, "UnsupportedChannels");
    if (block % 2 != 1) return fail(INVALID_ARGUMENT, "InvalidKernel");
    // The mean is boxFilter (windows up to 255) or GaussianBlur (up to 127),
    // whose InvalidKernel the reference returns beyond those.
    if (block > (method == 0 ? 255 : 127)) return fail(INVALID_ARGUMENT, "InvalidKernel");
    SepParams p;
    if (method == 0) {
        std::memset(&p, 0, sizeof p);
        p.nkx = p.nky = block;
        p.border = BORDER_REPLICATE;
        p.fin = FIN_SCALED;
        p.scale = (float)(1.0 / (double)(block * block));
        // Blocks above kMaxTaps take the two-pass kernels, which need no taps.
        for (int i = 0; i < block && i < kMaxTaps; i++) p.kxi[i] = p.kyi[i] = 1;
    } else if ((rc = gaussian_params(p, 5, block, block, 0, 0, BORDER_REPLICATE, src))) {
        return rc;
    }
    p.sdepth = p.ddepth = 0;
    const double dl = type == 0 ? std::ceil(delta) : std::floor(delta);
    const int idelta = (int)std::max(-1024.0, std::min(1024.0, dl));
    const int imax = (int)sat_cast<uint8_t>(maxval);
    // One kernel: the mean's producer compares each mean with its source
    // sample before the store, so the call needs no temporary and can be
    // captured in a graph.
    p.adapt.enabled = 1;
    p.adapt.idelta = idelta;
    p.adapt.on = type == 0 ? imax : 0;
    p.adapt.off = type == 0 ? 0 : imax;
    Call call;
    DImg s, d;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    p.src = s;
    p.dst = d;
    if (method == 0 && block > 127) {
        void *tmp = call.scratch(box_wide_tmp_bytes(s));
        if (!tmp) return fail(OUT_OF_MEMORY, "OutOfMemory");
        return call.finish(launch_box_wide(s, d, 0, block, block, BORDER_REPLICATE, p.scale, tmp, call.stream(), p.adapt));
    }
    return call.finish(method == 0 ? launch_box_u8(s, d, block, block, BORDER_REPLICATE, p.scale, call.stream(), p.adapt)
                                   : launch_sep(p, call.stream()));
}

// capi.zig:218-232 + filter.zig:856-906
int visar_sobel(const visar_image *src, const visar_image *dst, int dx, int dy, int ksize, double scale, int border) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (dx < 0 || dy < 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!src || !dst) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t S = src->depth;
    if (!valid_depth(S)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, S))) return rc;
    if (dst->depth == 3) {
        if (S != 0) return fail(UNSUPPORTED, "Unsupported");
    } else if (dst->depth != 5) {
        return fail(UNSUPPORTED, "Unsupported");
    }
    if ((rc = check_image(dst, dst->depth))) return rc;
    // OpenCV's apertures; larger Pascal kernels overflow the reference's i32 sums.
    if (ksize != -1 && ksize != 1 && ksize != 3 && ksize != 5 && ksize != 7) return fail(INVALID_ARGUMENT, "InvalidKernel");
    int n = ksize <= 1 ? 3 : ksize;
    int nx = (ksize == 1 && dx > 0) ? 3 : n, ny = (ksize == 1 && dy > 0) ? 3 : n;
    int32_t kx[32], ky[32];
    if (!deriv_kernel(kx, nx, dx, (ksize == 1 && dx == 0) ? 1 : (ksize == 1 ? 3 : ksize)) ||
        !deriv_kernel(ky, ny, dy, (ksize == 1 && dy == 0) ? 1 : (ksize == 1 ? 3 : ksize)))
        return fail(INVALID_ARGUMENT, "InvalidKernel");
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    SepParams p;
    std::memset(&p, 0, sizeof p);
    p.sdepth = (int)S;
    p.ddepth = (int)dst->depth;
    p.nkx = nx;
    p.nky = ny;
    p.border = border;
    p.scale = (float)scale;
    if (dst->depth == 5 || S == 5) {
        for (int i = 0; i < nx; i++) p.kxf[i] = (float)kx[i];
        for (int i = 0; i < ny; i++) p.kyf[i] = (float)ky[i];
        p.f32mode = 1;
        p.fin = FIN_SCALED;
        p.vsym = symmetry(p.kyf, ny);
        p.hsym = symmetry(p.kxf, nx);
        p.body = ((long)src->width * src->channels / 16) * 16;
    } else {
        for (int i = 0; i < nx; i++) p.kxi[i] = kx[i];
        for (int i = 0; i < ny; i++) p.kyi[i] = ky[i];
        p.fin = scale == 1 ? FIN_CAST : FIN_SCALED;
    }
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.src)) || (rc = call.output(dst, p.dst))) return rc;
    return call.finish(launch_sep(p, call.stream()));
}

// filter.zig:909-914 (sepFilter2D: f32 taps and accumulation, FinishCast)
int visar_sep_filter_2d(const visar_image *src, const visar_image *dst, const float *kx, size_t nkx, const float *ky,
                        size_t nky, int border) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (!src || !dst) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!valid_depth(src->depth) || !valid_depth(dst->depth)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, src->depth)) || (rc = check_image(dst, dst->depth))) return rc;
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    if (!kx || !ky || nkx == 0 || nky == 0 || nkx % 2 == 0 || nky % 2 == 0 || nkx > 127 || nky > 127)
        return fail(INVALID_ARGUMENT, "InvalidKernel");
    SepParams p;
    std::memset(&p, 0, sizeof p);
    p.sdepth = (int)src->depth;
    p.ddepth = (int)dst->depth;
    p.nkx = (int)nkx;
    p.nky = (int)nky;
    p.border = border;
    p.scale = 1.0f;
    std::memcpy(p.kxf, kx, sizeof(float) * nkx);
    std::memcpy(p.kyf, ky, sizeof(float) * nky);
    p.f32mode = 1;
    p.fin = FIN_CAST;
    p.vsym = symmetry(p.kyf, (int)nky);
    p.hsym = symmetry(p.kxf, (int)nkx);
    p.body = ((long)src->width * src->channels / 16) * 16;
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.src)) || (rc = call.output(dst, p.dst))) return rc;
    return call.finish(launch_sep(p, call.stream()));
}

// ----------------------------------------------------- color, threshold

// capi.zig:149-161 + color.zig:88-105 (gray codes)
int visar_cvt_color(const visar_image *src, const visar_image *dst, int code) {
    int scn, sb;
    switch (code) {
    case 6: scn = 3, sb = 0; break;
    case 7: scn = 3, sb = 2; break;
    case 10: scn = 4, sb = 0; break;
    case 11: scn = 4, sb = 2; break;
    default: return fail(UNSUPPORTED, "Unsupported");
    }
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (src->width != dst->width || src->height != dst->height) return fail(SIZE_MISMATCH, "SizeMismatch");
    if ((int)src->channels != scn || dst->channels != 1) return fail(UNSUPPORTED, "UnsupportedChannels");
    if (T == 3) return fail(UNSUPPORTED, "Unsupported");
    Call call;
    DImg s, d;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    return call.finish(launch_gray(s, d, (int)T, scn, sb, call.stream()));
}

// capi.zig:176-182 visar_gray_downsample -> videoio/frame.zig:170-173
// grayDownsample: BGR / BGRA u8 -> gray shrunk by k (1..16) with box
// averaging; dst is (src.w / k) x (src.h / k), 1 channel.
static int gray_down_check(const visar_image *src, const visar_image *dst, int k) {
    if (k <= 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    int rc;
    if ((rc = check_image(src, 0)) || (rc = check_image(dst, 0))) return rc;
    if (k > 16 || dst->channels != 1 || (src->channels != 3 && src->channels != 4))
        return fail(SIZE_MISMATCH, "SizeMismatch");
    if (dst->width != src->width / (uint32_t)k || dst->height != src->height / (uint32_t)k)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    return OK;
}

static int gray_down_run(const visar_image *src, const visar_image *dst, int k) {
    Call call;
    DImg s, d;
    int rc;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    return call.finish(launch_gray_down(s, d, k, call.stream()));
}

int visar_gray_downsample(const visar_image *src, const visar_image *dst, int k) {
    int rc = gray_down_check(src, dst, k);
    return rc ? rc : gray_down_run(src, dst, k);
}

// videoio/frame.zig:179-189 grayDownsampleRect: the (x, y, w, h) window of
// dst from the k x k source blocks under it; the rest of dst is untouched.
int visar_gray_downsample_rect(const visar_image *src, const visar_image *dst, int k, uint32_t x, uint32_t y,
                               uint32_t w, uint32_t h) {
    int rc = gray_down_check(src, dst, k);
    if (rc) return rc;
    if (w == 0 || h == 0) return OK;
    if ((uint64_t)x + w > dst->width || (uint64_t)y + h > dst->height) return fail(INVALID_ARGUMENT, "OutOfBounds");
    visar_image sv = *src, dv = *dst;
    const uint32_t ku = (uint32_t)k;
    sv.data = (uint8_t *)src->data + (size_t)y * ku * src->stride + (size_t)x * ku * src->channels;
    sv.width = w * ku;
    sv.height = h * ku;
    dv.data = (uint8_t *)dst->data + (size_t)y * dst->stride + x;
    dv.width = w;
    dv.height = h;
    return gray_down_run(&sv, &dv, k);
}

// calib3d/camera.zig:17-22 expandDist with capi_calib3d.zig:93-95
// distSlice (a null `dist` is no coefficients): 0, 4, 5 or 8 of them.
static int undistort_lens(const double k[9], const double *dist, size_t nd, UndistortLens &L) {
    if (!dist) nd = 0;
    if (nd != 0 && nd != 4 && nd != 5 && nd != 8) return fail(INVALID_ARGUMENT, "InvalidArgument");
    for (size_t i = 0; i < 8; i++) L.k[i] = i < nd ? dist[i] : 0.0;
    L.rational = L.k[5] != 0 || L.k[6] != 0 || L.k[7] != 0;
    L.fx = k[0];
    L.fy = k[4];
    L.u0 = k[2];
    L.v0 = k[5];
    return OK;
}

// capi_calib3d.zig:186-191 visar_undistort -> calib3d/undistort.zig:205-272:
// 8U, 1-4 channels, dst the shape of src, new_k defaulting to K; stripes of
// max(1, 4096 / width) rows, each with its own inverse of new_k.
int visar_undistort(const visar_image *src, const visar_image *dst, const double k[9], const double *dist, size_t nd,
                    const double *new_k) {
    int rc;
    if ((rc = check_image(src, 0)) || (rc = check_image(dst, 0))) return rc;
    if (!k) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (src->width != dst->width || src->height != dst->height || src->channels != dst->channels)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    if (src->channels < 1 || src->channels > 4) return fail(UNSUPPORTED, "Unsupported");
    UndistortArgs p;
    if ((rc = undistort_lens(k, dist, nd, p.lens))) return rc;
    for (int i = 0; i < 9; i++) p.ar[i] = new_k ? new_k[i] : k[i];
    p.stripe = (int)std::min(std::max<uint32_t>(1, 4096 / src->width), src->height);
    // Rows of a stripe whose inverse is singular stay unwritten; then the
    // output starts as the destination's contents (the same inv3 as the
    // kernel's, so host and device agree on which stripes those are). With
    // a bottom row (0, 0, x) the moved element only meets exact zeros in the
    // determinant, so the first stripe decides for all.
    bool partial = false;
    const uint32_t last = p.ar[6] == 0 && p.ar[7] == 0 ? 1 : src->height;
    for (uint32_t y0 = 0; y0 < last && !partial; y0 += (uint32_t)p.stripe) {
        double ar[9], ir[9];
        for (int i = 0; i < 9; i++) ar[i] = p.ar[i];
        ar[5] = p.ar[5] - (double)y0;
        partial = !inv3(ar, ir);
    }
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.s))) return rc;
    if ((rc = partial ? call.output_keep(dst, p.d) : call.output(dst, p.d))) return rc;
    return call.finish(launch_undistort(p, call.stream()));
}

// capi_calib3d.zig:195-201 visar_init_undistort_rectify_map ->
// undistort.zig:36-41 MapParams.init (inverse of new_k * R; singular:
// InvalidArgument) and :82-103 initUndistortRectifyMap (32F, one channel).
int visar_init_undistort_rectify_map(const double k[9], const double *dist, size_t nd, const double *r,
                                     const double *new_k, const visar_image *mapx, const visar_image *mapy) {
    int rc;
    if ((rc = check_image(mapx, 5)) || (rc = check_image(mapy, 5))) return rc;
    if (!k) return fail(INVALID_ARGUMENT, "InvalidArgument");
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double *a = new_k ? new_k : k, *b = r ? r : I;
    double m[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
    UndistortMapArgs p;
    if (!inv3(m, p.ir)) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if ((rc = undistort_lens(k, dist, nd, p.lens))) return rc;
    if (mapx->width != mapy->width || mapx->height != mapy->height || mapx->channels != mapy->channels ||
        mapx->channels != 1)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.output(mapx, p.mx)) || (rc = call.output(mapy, p.my))) return rc;
    return call.finish(launch_undistort_map(p, call.stream()));
}

// capi.zig:310-322 + threshold.zig:25-82
int visar_threshold(const visar_image *src, const visar_image *dst, double thresh, double maxval, int type,
                    double *used) {
    int kind = type & 7;
    if (kind > 4) return fail(UNSUPPORTED, "Unsupported");
    int autom = (type & 8) ? 1 : (type & 16) ? 2 : 0;
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    Call call;
    ThreshParams p;
    std::memset(&p, 0, sizeof p);
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.s))) return rc;
    if (autom) {
        if (T != 0 || src->channels != 1) return fail(UNSUPPORTED, "Unsupported");
        uint32_t *hd = (uint32_t *)call.scratch(256 * sizeof(uint32_t));
        uint32_t *hist = (uint32_t *)call.pinned(256 * sizeof(uint32_t));
        if (!hd || !hist) return fail(OUT_OF_MEMORY, "OutOfMemory");
        if (hipMemsetAsync(hd, 0, 256 * sizeof(uint32_t), call.stream()) != hipSuccess ||
            launch_hist_u8(p.s, hd, call.stream()) != OK ||
            hipMemcpyAsync(hist, hd, 256 * sizeof(uint32_t), hipMemcpyDeviceToHost, call.stream()) != hipSuccess ||
            hipStreamSynchronize(call.stream()) != hipSuccess)
            return fail(INTERNAL, "KernelFailed");
        thresh = autom == 1 ? threshold_otsu(hist, (double)src->width * src->height) : threshold_triangle(hist);
    }
    if (used) *used = thresh;
    threshold_setup(p, T, kind, thresh, maxval);
    DImg d;
    if ((rc = call.output(dst, d))) return rc;
    p.d = d;
    return call.finish(launch_threshold(p, call.stream()));
}

// ------------------------------------------------------------- resize

// capi.zig:184-191 + resize.zig:28-62
int visar_resize(const visar_image *src, const visar_image *dst, int interp) {
    if (interp < 0 || interp > 4) return fail(UNSUPPORTED, "Unsupported");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (src->channels != dst->channels) return fail(SIZE_MISMATCH, "SizeMismatch");
    const bool same = src->width == dst->width && src->height == dst->height;
    const bool u8fast = T == 0 && src->channels <= 4;
    // INTER_CUBIC and LANCZOS4 are outside this path.
    if (!same && interp != 0 && interp != 1 && interp != 3) return fail(UNSUPPORTED, "Unsupported");
    Call call;
    DImg s, d;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    hipStream_t st = call.stream();
    if (same) return call.finish(launch_copy(s, d, depth_size(T), st));
    if (interp == 0) return call.finish(launch_nearest(s, d, depth_size(T), st));
    if (!u8fast) {
        // resize.zig:47-60: the generic kernels (AREA enlarging is LINEAR
        // with area-style taps; u8 keeps linearTapU8's fixed point).
        const bool up = dst->width > src->width || dst->height > src->height;
        const int cn = (int)src->channels;
        if (interp == 1 || up) {
            const bool area = interp == 3;
            void *dev = nullptr;
            if (T == 0) {
                std::vector<LinTap> tabs(dst->width + dst->height);
                for (uint32_t x = 0; x < dst->width; x++) {
                    LinTap t = linear_tap_u8(x, src->width, dst->width, area, true);
                    t.o0 *= cn;
                    t.o1 *= cn;
                    tabs[x] = t;
                }
                for (uint32_t y = 0; y < dst->height; y++)
                    tabs[dst->width + y] = linear_tap_u8(y, src->height, dst->height, area, false);
                dev = call.upload(tabs.data(), tabs.size() * sizeof(LinTap));
                if (!dev) return fail(OUT_OF_MEMORY, "OutOfMemory");
                const LinTap *xt = (const LinTap *)dev;
                return call.finish(launch_linear_any(s, d, (int)T, xt, xt + dst->width, st));
            }
            std::vector<FLinTap> tabs(dst->width + dst->height);
            const double sx = (double)src->width / (double)dst->width, sy = (double)src->height / (double)dst->height;
            for (uint32_t x = 0; x < dst->width; x++) {
                long i;
                double f;
                linear_coord(x, sx, src->width, area, &i, &f);
                const long i1 = i + 1 < (long)src->width ? i + 1 : i;
                tabs[x] = FLinTap{(int32_t)(i * cn), (int32_t)(i1 * cn), (float)(1.0 - f), (float)f};
            }
            for (uint32_t y = 0; y < dst->height; y++) {
                long i;
                double f;
                linear_coord(y, sy, src->height, area, &i, &f);
                const long i1 = i + 1 < (long)src->height ? i + 1 : i;
                tabs[dst->width + y] = FLinTap{(int32_t)i, (int32_t)i1, (float)(1.0 - f), (float)f};
            }
            dev = call.upload(tabs.data(), tabs.size() * sizeof(FLinTap));
            if (!dev) return fail(OUT_OF_MEMORY, "OutOfMemory");
            const FLinTap *xt = (const FLinTap *)dev;
            return call.finish(launch_linear_any(s, d, (int)T, xt, xt + dst->width, st));
        }
        const uint32_t ix = src->width / dst->width, iy = src->height / dst->height;
        if (ix * dst->width == src->width && iy * dst->height == src->height)
            return call.finish(launch_area_int_any(s, d, (int)T, (int)ix, (int)iy, st));
    }
    const bool exact_half = src->width == 2 * dst->width && src->height == 2 * dst->height;
    bool area_up = interp == 3 && (dst->width > src->width || dst->height > src->height);
    if (interp == 1 || area_up) {
        if (exact_half && interp == 1) return call.finish(launch_half_u8(s, d, st));
        // resize.zig:216-242: x taps pinned to the edges, y taps not.
        std::vector<LinTap> tabs(dst->width + dst->height);
        for (uint32_t x = 0; x < dst->width; x++) {
            LinTap t = linear_tap_u8(x, src->width, dst->width, area_up, true);
            t.o0 *= (int32_t)src->channels;
            t.o1 *= (int32_t)src->channels;
            tabs[x] = t;
        }
        for (uint32_t y = 0; y < dst->height; y++) tabs[dst->width + y] = linear_tap_u8(y, src->height, dst->height, area_up, false);
        LinTap *dev = (LinTap *)call.upload(tabs.data(), tabs.size() * sizeof(LinTap));
        if (!dev) return fail(OUT_OF_MEMORY, "OutOfMemory");
        const bool box = linear_box_fits(tabs.data(), (int)dst->width, (int)src->channels, tabs.data() + dst->width, (int)dst->height);
        const bool stream = linear_stream_fits(tabs.data(), (int)dst->width, (int)src->channels, tabs.data() + dst->width, (int)dst->height);
        return call.finish(launch_linear_u8(s, d, dev, dev + dst->width, st, box, stream));
    }
    // INTER_AREA shrinking
    if (exact_half) return call.finish(launch_half_u8(s, d, st));
    uint32_t ix = src->width / dst->width, iy = src->height / dst->height;
    if (ix * dst->width == src->width && iy * dst->height == src->height)
        return call.finish(launch_area_int_u8(s, d, (int)ix, (int)iy, st));
    std::vector<AreaEntry> xe, ye;
    std::vector<int32_t> xs, ys;
    area_tab(src->width, dst->width, (double)src->width / (double)dst->width, xe, xs);
    area_tab(src->height, dst->height, (double)src->height / (double)dst->height, ye, ys);
    size_t bytes = (xe.size() + ye.size()) * sizeof(AreaEntry) + (xs.size() + ys.size()) * sizeof(int32_t);
    std::vector<uint8_t> blob(bytes);
    size_t o = 0;
    auto put = [&](const void *p, size_t n) {
        std::memcpy(blob.data() + o, p, n);
        o += n;
    };
    put(xe.data(), xe.size() * sizeof(AreaEntry));
    put(ye.data(), ye.size() * sizeof(AreaEntry));
    put(xs.data(), xs.size() * sizeof(int32_t));
    put(ys.data(), ys.size() * sizeof(int32_t));
    uint8_t *dev = (uint8_t *)call.upload(blob.data(), bytes);
    if (!dev) return fail(OUT_OF_MEMORY, "OutOfMemory");
    const AreaEntry *dxe = (const AreaEntry *)dev;
    const AreaEntry *dye = dxe + xe.size();
    const int32_t *dxs = (const int32_t *)(dye + ye.size());
    const int32_t *dys = dxs + xs.size();
    if (!u8fast) return call.finish(launch_area_gen_any(s, d, (int)T, dxe, dxs, dye, dys, st));
    const bool box = area_box_fits(xe.data(), xs.data(), (int)dst->width, (int)src->channels, ye.data(), ys.data(), (int)dst->height);
    return call.finish(launch_area_frac_u8(s, d, dxe, dxs, dye, dys, st, box));
}

// ------------------------------------------------------------------ warps

// capi.zig:372-380 + geometry.zig:410-426
int visar_warp_affine(const visar_image *src, const visar_image *dst, const double m[6], int interp, int border,
                      double value) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (!m) return fail(INVALID_ARGUMENT, "InvalidArgument");
    double inv[6];
    if (!invert_affine(m, inv)) return fail(UNSUPPORTED, "Unsupported");
    if (src->channels != dst->channels) return fail(SIZE_MISMATCH, "SizeMismatch");
    AffineParams p;
    std::memset(&p, 0, sizeof p);
    p.depth = (int)T;
    p.interp = interp == 0 ? 0 : 1;
    p.border = border;
    for (int i = 0; i < 6; i++) p.m[i] = (float)inv[i];
    switch (T) {
    case 0: p.cvd = sat_cast<uint8_t>(value); break;
    case 2: p.cvd = sat_cast<uint16_t>(value); break;
    case 3: p.cvd = sat_cast<int16_t>(value); break;
    default: p.cvd = (double)(float)value; break;
    }
    p.cv = (float)p.cvd;
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.s)) || (rc = call.output(dst, p.d))) return rc;
    return call.finish(launch_warp_affine(p, call.stream()));
}

// geometry.zig:586-589, 615-672
int visar_warp_perspective(const visar_image *src, const visar_image *dst, const double m[9], int interp, int border,
                           double value) {
    if (border < 0 || border > 4) return fail(UNSUPPORTED, "Unsupported");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (!m) return fail(INVALID_ARGUMENT, "InvalidArgument");
    PerspParams p;
    std::memset(&p, 0, sizeof p);
    if (!invert3(m, p.m)) return fail(UNSUPPORTED, "Unsupported");
    if (src->channels != dst->channels) return fail(SIZE_MISMATCH, "SizeMismatch");
    p.depth = (int)T;
    p.interp = interp == 0 ? 0 : 1;
    p.border = border;
    switch (T) {
    case 0: p.cvd = sat_cast<uint8_t>(value); break;
    case 2: p.cvd = sat_cast<uint16_t>(value); break;
    case 3: p.cvd = sat_cast<int16_t>(value); break;
    default: p.cvd = (double)(float)value; break;
    }
    Call call;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, p.s)) || (rc = call.output(dst, p.d))) return rc;
    return call.finish(launch_warp_perspective(p, call.stream()));
}

// capi_calib3d.zig:205-213 + undistort.zig:280-323
int visar_remap(const visar_image *src, const visar_image *dst, const visar_image *mapx, const visar_image *mapy,
                double value) {
    int rc;
    if ((rc = check_image(src, 0)) || (rc = check_image(dst, 0)) || (rc = check_image(mapx, 5)) ||
        (rc = check_image(mapy, 5)))
        return rc;
    if (dst->width != mapx->width || dst->height != mapx->height || !same_shape(mapx, mapy) || mapx->channels != 1)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    if (src->channels != dst->channels || src->channels < 1 || src->channels > 4) return fail(UNSUPPORTED, "Unsupported");
    double v = std::round(max_num(0.0, min_num(value, 255.0)));
    Call call;
    DImg s, d, mx, my;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.input(mapx, mx)) || (rc = call.input(mapy, my)) ||
        (rc = call.output(dst, d)))
        return rc;
    return call.finish(launch_remap(s, d, mx, my, (int)v, call.stream()));
}

// ------------------------------------------------------------- morphology

// morph.zig:111-154: the element as horizontal runs of "on" cells.
static int element_runs(const std::vector<uint8_t> &mask, int w, int h, MorphParams &p) {
    p.nruns = 0;
    for (int ky = 0; ky < h; ky++) {
        int x = 0;
        while (x < w) {
            if (!mask[(size_t)ky * w + x]) {
                x++;
                continue;
            }
            int x0 = x;
            while (x < w && mask[(size_t)ky * w + x]) x++;
            if (p.nruns == kMaxRuns) return UNSUPPORTED;
            p.runs[p.nruns++] = MorphRun{(int16_t)ky, (int16_t)x0, (int16_t)(x - x0), 0};
        }
    }
    return OK;
}

// erode^n or dilate^n from `s` into `d` (morph.zig:63-96); n = 0 copies.
static int morph_iter(Call &call, const DImg &s, const DImg &d, MorphParams p, int op_max, int iterations,
                      size_t elem) {
    hipStream_t st = call.stream();
    if (iterations == 0) return launch_copy(s, d, elem, st);
    DImg t1, t2;
    if (iterations > 1) {
        if (call.temp_like(s, elem, t1) || call.temp_like(s, elem, t2)) return OUT_OF_MEMORY;
    }
    p.op_max = op_max;
    DImg cur = s;
    for (int it = 0; it < iterations; it++) {
        DImg out = it == iterations - 1 ? d : (it % 2 == 0 ? t1 : t2);
        p.s = cur;
        p.d = out;
        int rc = launch_morph_pass(p, st);
        if (rc) return rc;
        cur = out;
    }
    return OK;
}

// Passes of one fused launch: at most this many per chain, so that the
// summed halo stays within 6 rows and 64 bytes.
static int morph_fuse_limit(int cn, const MorphRects &r) {
    int f = 6;
    if (r.ev > 0) f = std::min(f, 6 / r.ev);
    if (r.eh > 0) f = std::min(f, 64 / (r.eh * cn));
    return std::max(f, 1);
}

enum { COMB_NONE = 0, COMB_SRC = 1, COMB_CHAIN = 2, COMB_IMAGE = 3 };

// One launch of the fused kernel: `nchains` chains of passes (op_max per
// pass) from `in`, the last pass of the last chain storing to `out`. The
// chains all start from the staged input; with two, the first one's result
// is kept for the combine (comb COMB_CHAIN). COMB_SRC combines with the
// staged input itself, COMB_IMAGE with image `cimg`. UNSUPPORTED when the
// tile does not fit the device's LDS.
static int morph_chain_launch(const DImg &in, const DImg &out, const MorphRects &r, const std::vector<int> *chains,
                              int nchains, int post, int comb, const DImg *cimg, hipStream_t st) {
    MorphChainArgs a;
    std::memset(&a, 0, sizeof a);
    a.s = in;
    a.d = out;
    a.c = cimg ? *cimg : in;
    int F = 0;
    for (int c = 0; c < nchains; c++) F = std::max(F, (int)chains[c].size());
    const int cn = in.cn;
    const int K = std::max(1, std::max(r.eh, r.ev));
    a.ev = r.ev;
    a.vb = F * r.ev;
    a.hb = (F * r.eh * cn + 7) & ~7;
    a.tw = 512 - 2 * a.hb;
    a.nr = r.nr;
    for (int k = 0; k < r.nr; k++) {
        a.rw[k] = r.rw[k];
        a.rv[k] = r.rv[k];
        for (int e = -r.rv[k]; e <= r.rv[k]; e++) a.cmask[r.rw[k]] |= 1u << (e + K);
    }
    // Buffers: 0 holds the staged input while a later chain or the combine
    // still needs it; a chain's result stays while the combine needs it.
    int ra = -1, nbuf = 1, np = 0;
    for (int c = 0; c < nchains; c++) {
        int cur = 0;
        const int len = (int)chains[c].size();
        for (int i = 0; i < len; i++) {
            MorphChainPass ps{(int8_t)cur, (int8_t)-1, (int8_t)chains[c][i], (int8_t)(i + 1)};
            if (!(c == nchains - 1 && i == len - 1)) {
                int dst = -1;
                for (int b = 0; b < 3 && dst < 0; b++) {
                    const bool busy = b == cur || b == ra || (b == 0 && (comb == COMB_SRC || c < nchains - 1));
                    if (!busy) dst = b;
                }
                if (dst < 0) return UNSUPPORTED;
                ps.dst = (int8_t)dst;
                nbuf = std::max(nbuf, dst + 1);
                cur = dst;
            }
            if (np == kChainMaxPasses) return UNSUPPORTED;
            a.pass[np++] = ps;
        }
        if (c < nchains - 1) ra = cur;
    }
    a.npass = np;
    a.post = post;
    a.comb = comb == COMB_SRC ? 0 : comb == COMB_CHAIN ? ra : comb == COMB_IMAGE ? 3 : 0;
    a.nbuf = nbuf;
    const size_t pitch = (size_t)morph_chain_pitch_dw(cn, K) * 4;
    const size_t limit = std::min<size_t>(morph_chain_lds_limit(), 160 * 1024);
    int th = tuning().morph_rows >= 8 ? tuning().morph_rows : F == 1 ? 32 : 16;
    auto lds_of = [&](int t) { return nbuf * (t + 2 * a.vb) * pitch + (comb == COMB_IMAGE ? t * pitch : 0); };
    while (th > 8 && lds_of(th) > limit) th /= 2;
    if (lds_of(th) > limit) return UNSUPPORTED;
    a.th = th;
    a.rows = th + 2 * a.vb;
    const int n = in.w * cn;
    const int tiles_y = (in.h + th - 1) / th;
    a.tiles_x = (n + a.tw - 1) / a.tw;
    const int k = tuning().tile_order;
    a.ntiles = a.tiles_x * tiles_y;
    a.xcd = k != 1;
    return launch_morph_chain(a, K, lds_of(th), st);
}

// One or two erode / dilate passes of morph_direct.hpp from `in` into `out`
// (ops bit p: pass p dilates); post 1: out = saturate(c - v), 2: out =
// saturate(v - c). UNSUPPORTED (nothing launched) when the element has no
// kernel there.
static int morph_direct_launch(const DImg &in, const DImg &out, uint32_t el, int np, int ops, int post, const DImg *c,
                               hipStream_t st) {
    MorphDirectArgs a;
    a.s = in;
    a.d = out;
    a.c = c ? *c : in;
    auto span = [](const DImg &m) { return (uint64_t)(m.h - 1) * (uint64_t)m.stride + (uint64_t)m.w * m.cn; };
    if (span(a.s) >= (1ull << 31) || span(a.d) >= (1ull << 31) || span(a.c) >= (1ull << 31) || a.s.stride >= (1l << 31) ||
        a.d.stride >= (1l << 31) || a.c.stride >= (1l << 31))
        return UNSUPPORTED;
    a.span_s = (uint32_t)span(a.s);
    a.span_d = (uint32_t)span(a.d);
    a.span_c = (uint32_t)span(a.c);
    const int sw = 512 - 16 * (np - 1);
    a.strips = (in.w * in.cn + sw - 1) / sw;
    a.bands4 = ((in.h + kMorphDirectRows - 1) / kMorphDirectRows + 3) / 4;
    a.ntiles = a.strips * a.bands4;
    a.xcd = tuning().tile_order != 1;
    a.ops = ops;
    const bool ok = np == 2 ? launch_morph_direct2(a, el, post, st) : launch_morph_direct1(a, el, post, st);
    if (!ok) return UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

// One launch of morph_stream.hpp: base element `el` taken n times (n = 2,
// 3; n = 1 for gradient), mode 0 erode, 1 dilate, 2 gradient; post as in
// morph_direct_launch. UNSUPPORTED (nothing launched) when there is no kernel.
static int morph_stream_launch(const DImg &in, const DImg &out, uint32_t el, int n, int mode, int post, const DImg *c,
                               hipStream_t st) {
    MorphStreamArgs a;
    std::memset(&a, 0, sizeof a);
    a.s = in;
    a.d = out;
    a.c = c ? *c : in;
    auto span = [](const DImg &m) { return (uint64_t)(m.h - 1) * (uint64_t)m.stride + (uint64_t)m.w * m.cn; };
    if (in.w < 8 || span(a.s) >= (1ull << 31) || span(a.d) >= (1ull << 31) || span(a.c) >= (1ull << 31))
        return UNSUPPORTED;
    a.span_s = (uint32_t)span(a.s);
    a.span_d = (uint32_t)span(a.d);
    a.span_c = (uint32_t)span(a.c);
    a.tb = tuning().morph_band > 0 ? tuning().morph_band : 16;
    a.xcd = tuning().tile_order != 1;
    if (!launch_morph_stream(a, el, n, mode, post, st)) return UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

// morphologyEx (morph.zig:532-571) with the streamed n-fold passes of
// morph_stream.hpp: n iterations of an erode or dilate as one launch (runs
// of 3 and 2 for n > 3), gradient's dilation and erosion as one launch (n
// up to 3), the subtraction of tophat and blackhat folded into the last
// launch. For one-channel images whose rows are a multiple of 4 bytes (at
// least 8) starting on dwords, base elements with codes of their own where
// morph_stream_pays, and n >= 2 (1 <= n <= 3 for gradient). UNSUPPORTED
// (nothing launched) otherwise.
// Whether one streamed pass of `el` taken n times beats n passes (and so
// has a kernel): mstream::kStreamMaxRows.
static bool morph_stream_pays(uint32_t el, int n) {
    return mstream::acc_rows(mstream::power(el, n)) <= mstream::kStreamMaxRows;
}

static int morph_band_launch(const DImg &in, const DImg &out, uint32_t el, int op_max, int post, const DImg *c,
                             hipStream_t st);

// The n-fold sums morph_pow.hpp has launches for: half-width up to 11,
// half-height up to 9, where the accumulator kernel does not pay.
static bool pow_fits(uint32_t el, int n) { return morph_pow_has(el, n); }

// n passes of a base element as one launch of morph_pow.hpp (post as in
// morph_direct_launch). UNSUPPORTED (nothing launched) when it has none.
static int morph_pow_launch(const DImg &in, const DImg &out, uint32_t el, int n, int op_max, int post, const DImg *c,
                            hipStream_t st) {
    MorphDirectArgs a;
    std::memset(&a, 0, sizeof a);
    a.s = in;
    a.d = out;
    a.c = c ? *c : in;
    auto span = [](const DImg &m) { return (uint64_t)(m.h - 1) * (uint64_t)m.stride + (uint64_t)m.w * m.cn; };
    if (in.w < 8 || span(a.s) >= (1ull << 31) || span(a.d) >= (1ull << 31) || span(a.c) >= (1ull << 31))
        return UNSUPPORTED;
    a.span_s = (uint32_t)span(a.s);
    a.span_d = (uint32_t)span(a.d);
    a.span_c = (uint32_t)span(a.c);
    if (!launch_morph_pow(a, el, n, op_max, post, st)) return UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

static int morph_stream_ex(Call &call, const DImg &s, const DImg &d, int op, int n, uint32_t el) {
    auto aligned = [](const DImg &m) { return ((((uintptr_t)m.p) | (uintptr_t)m.stride) & 3) == 0; };
    auto span = [](const DImg &m) { return (uint64_t)(m.h - 1) * (uint64_t)m.stride + (uint64_t)m.w * m.cn; };
    if (tuning().morph_stream == 1 || s.cn != 1 || (s.w & 3) || s.w < 8 || !aligned(s) || !aligned(d) ||
        span(s) >= (1ull << 31) || span(d) >= (1ull << 31))
        return UNSUPPORTED;
    if ((el & 15u) == 0) return UNSUPPORTED;
    hipStream_t st = call.stream();
    if (op == 4 && n <= 3 && morph_stream_pays(el, n)) return morph_stream_launch(s, d, el, n, 2, 0, nullptr, st);
    if (n < 2) return UNSUPPORTED;
    // Runs of r passes in one launch each: the accumulator kernel where it
    // pays, else the n-fold column-window kernel, else single passes. Runs
    // of 3 and 2 (4 as 2 + 2), single passes last; no run of 2 or more:
    // UNSUPPORTED (pass by pass). The first launch is a multi-pass run, so
    // an element without kernels fails before anything is launched.
    auto run_ok = [&](int r) { return morph_stream_pays(el, r) || pow_fits(el, r); };
    std::vector<int> runs;
    int ones = 0;
    for (int left = n; left > 0;) {
        int r = left == 4 ? 2 : std::min(3, left);
        while (r > 1 && !run_ok(r)) r--;
        if (r == 1) ones++;
        else runs.push_back(r);
        left -= r;
    }
    if (runs.empty()) return UNSUPPORTED;
    runs.insert(runs.end(), ones, 1);
    auto run_launch = [&](const DImg &in, const DImg &out, int r, int mode, int post, const DImg *c) {
        if (r >= 2 && morph_stream_pays(el, r)) return morph_stream_launch(in, out, el, r, mode, post, c, st);
        if (r >= 2) return morph_pow_launch(in, out, el, r, mode, post, c, st);
        return morph_band_launch(in, out, el, mode, post, c, st);
    };
    // Passes of one kind from `in` into `out`, the last run with post
    // against c.
    DImg tmp[2] = {};
    int k = 0;
    bool first = true;
    auto passes = [&](const DImg &in, const DImg &out, int mode, int post, const DImg *c) -> int {
        DImg cur = in;
        for (size_t q = 0; q < runs.size(); q++) {
            const bool last = q + 1 == runs.size();
            DImg o = out;
            if (!last) {
                DImg &t = tmp[k++ & 1];
                if (!t.p && call.temp_like(s, 1, t)) return OUT_OF_MEMORY;
                o = t;
            }
            const int rc = run_launch(cur, o, runs[q], mode, last ? post : 0, c);
            if (rc == UNSUPPORTED && !first) return INTERNAL;
            if (rc) return rc;
            first = false;
            cur = o;
        }
        return OK;
    };
    if (op <= 1) return passes(s, d, op, 0, nullptr);
    DImg t;
    if (call.temp_like(s, 1, t)) return OUT_OF_MEMORY;
    if (op == 4) {
        // gradient: the n-fold dilation, then the erosion subtracted from it
        int rc = passes(s, t, 1, 0, nullptr);
        if (rc) return rc;
        rc = passes(s, d, 0, 1, &t);
        return rc == UNSUPPORTED ? INTERNAL : rc;
    }
    // open / tophat: erode then dilate; close / blackhat: dilate then erode.
    const int m1 = op == 2 || op == 5 ? 0 : 1;
    int rc = passes(s, t, m1, 0, nullptr);
    if (rc) return rc;
    rc = passes(t, d, 1 - m1, op == 5 ? 1 : op == 6 ? 2 : 0, &s);
    return rc == UNSUPPORTED ? INTERNAL : rc;
}

// One erode / dilate pass of morph_band.hpp from `in` into `out` (post as in
// morph_direct_launch). UNSUPPORTED (nothing launched) when it has no kernel.
static int morph_band_launch(const DImg &in, const DImg &out, uint32_t el, int op_max, int post, const DImg *c,
                             hipStream_t st) {
    MorphDirectArgs a;
    std::memset(&a, 0, sizeof a);
    a.s = in;
    a.d = out;
    a.c = c ? *c : in;
    auto span = [](const DImg &m) { return (uint64_t)(m.h - 1) * (uint64_t)m.stride + (uint64_t)m.w * m.cn; };
    if (in.w < 8 || span(a.s) >= (1ull << 31) || span(a.d) >= (1ull << 31) || span(a.c) >= (1ull << 31))
        return UNSUPPORTED;
    a.span_s = (uint32_t)span(a.s);
    a.span_d = (uint32_t)span(a.d);
    a.span_c = (uint32_t)span(a.c);
    a.xcd = tuning().tile_order != 1;
    if (!launch_morph_band(a, el, op_max, post, st)) return UNSUPPORTED;
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

// morphologyEx (morph.zig:532-571) as one-channel launches of
// morph_direct.hpp, one or two passes each, intermediates in temporaries, the
// subtraction folded into the last launch: for one-channel images whose
// rows are a multiple of 4 bytes starting on dwords, with elements that
// have kernels there. UNSUPPORTED (nothing launched) otherwise.
static int morph_direct_ex(Call &call, const DImg &s, const DImg &d, int op, int n, const MorphRects &r) {
    auto aligned = [](const DImg &m) { return ((((uintptr_t)m.p) | (uintptr_t)m.stride) & 3) == 0; };
    if (s.cn != 1 || (s.w & 3) || !aligned(s) || !aligned(d) || tuning().morph_chain == 2) return UNSUPPORTED;
    uint32_t el = (uint32_t)r.nr;
    for (int k = 0; k < r.nr; k++) el |= (uint32_t)(r.rw[k] | r.rv[k] << 2) << (4 + 4 * k);
    {
        const int rc = morph_stream_ex(call, s, d, op, n, el);
        if (rc != UNSUPPORTED) return rc;
    }
    std::vector<int> A, B;
    int post = 0;
    switch (op) {
    case 0: A.assign(n, 0); break;
    case 1: A.assign(n, 1); break;
    case 2: A.assign(n, 0); A.insert(A.end(), n, 1); break;
    case 3: A.assign(n, 1); A.insert(A.end(), n, 0); break;
    case 4: A.assign(n, 1); B.assign(n, 0); post = 1; break;   // dilate - erode
    case 5: A.assign(n, 0); A.insert(A.end(), n, 1); post = 1; break;  // src - open
    default: A.assign(n, 1); A.insert(A.end(), n, 0); post = 2; break; // close - src
    }
    hipStream_t st = call.stream();
    DImg tmp[2] = {};
    bool first = true;
    int k = 0;
    // Two passes per launch where the element is at most 5 x 5 (a 7 x 7
    // element's second pass costs more than the memory round trip it saves).
    const int per_launch = std::max(r.eh, r.ev) <= 2 ? 2 : 1;
    // Runs `ops` from s; the last launch writes `out`, combining with `comb`.
    auto run = [&](const std::vector<int> &ops, const DImg &out, int pst, const DImg *comb) -> int {
        DImg cur = s;
        for (size_t i = 0; i < ops.size();) {
            const int np = per_launch == 2 && i + 1 < ops.size() ? 2 : 1;
            const bool last = i + np == ops.size();
            DImg o = out;
            if (!last) {
                DImg &t = tmp[k++ & 1];
                if (!t.p && call.temp_like(s, 1, t)) return OUT_OF_MEMORY;
                o = t;
            }
            const int bits = ops[i] | (np == 2 ? ops[i + 1] << 1 : 0);
            int rc = UNSUPPORTED;
            if (np == 1 && tuning().morph_pass != 1) rc = morph_band_launch(cur, o, el, ops[i], last ? pst : 0, comb, st);
            if (rc == UNSUPPORTED) rc = morph_direct_launch(cur, o, el, np, bits, last ? pst : 0, comb, st);
            if (rc == UNSUPPORTED && !first) return INTERNAL;
            if (rc) return rc;
            first = false;
            cur = o;
            i += np;
        }
        return OK;
    };
    if (op == 4) {
        DImg ta;
        if (call.temp_like(s, 1, ta)) return OUT_OF_MEMORY;
        int rc = run(A, ta, 0, nullptr);
        return rc ? rc : run(B, d, post, &ta);
    }
    return run(A, d, post, &s);
}

// morphologyEx (morph.zig:532-571) through the fused kernel: the passes of
// each chain in launches of at most morph_fuse_limit passes, intermediates
// in temporaries, the subtraction folded into the last launch.
// UNSUPPORTED (nothing launched) when the kernel does not take the call.
static int morph_chain_ex(Call &call, const DImg &s, const DImg &d, int op, int n, const MorphRects &r) {
    const int F = morph_fuse_limit(s.cn, r);
    std::vector<int> A, B;
    int post = 0;
    switch (op) {
    case 0: A.assign(n, 0); break;
    case 1: A.assign(n, 1); break;
    case 2: A.assign(n, 0); A.insert(A.end(), n, 1); break;
    case 3: A.assign(n, 1); A.insert(A.end(), n, 0); break;
    case 4: A.assign(n, 1); B.assign(n, 0); post = 1; break;   // dilate - erode
    case 5: A.assign(n, 0); A.insert(A.end(), n, 1); post = 1; break;  // src - open
    default: A.assign(n, 1); A.insert(A.end(), n, 0); post = 2; break; // close - src
    }
    hipStream_t st = call.stream();
    if (op == 4 && n <= F) {
        std::vector<int> ch[2] = {A, B};
        return morph_chain_launch(s, d, r, ch, 2, post, COMB_CHAIN, nullptr, st);
    }
    // One chain from s into `out` in segments; the last one combines.
    auto run_chain = [&](const std::vector<int> &ops, const DImg &out, int pst, int comb, const DImg *cimg) -> int {
        const int P = (int)ops.size();
        const int nseg = (P + F - 1) / F;
        DImg cur = s;
        DImg tmp[2] = {};
        for (int k = 0, i0 = 0; k < nseg; k++) {
            const int len = (P - i0) / (nseg - k);
            std::vector<int> seg(ops.begin() + i0, ops.begin() + i0 + len);
            const bool last = k == nseg - 1;
            DImg o = out;
            if (!last) {
                if (!tmp[k & 1].p && call.temp_like(s, 1, tmp[k & 1])) return OUT_OF_MEMORY;
                o = tmp[k & 1];
            }
            int cm = COMB_NONE;
            const DImg *ci = nullptr;
            if (last && pst) {
                cm = comb;
                ci = cimg;
                if (comb == COMB_SRC && k > 0) {  // the source is no longer the staged input
                    cm = COMB_IMAGE;
                    ci = &s;
                }
            }
            int rc = morph_chain_launch(cur, o, r, &seg, 1, last ? pst : 0, cm, ci, st);
            if (rc) return rc;
            cur = o;
            i0 += len;
        }
        return OK;
    };
    if (op == 4) {
        DImg ta;
        if (call.temp_like(s, 1, ta)) return OUT_OF_MEMORY;
        int rc = run_chain(A, ta, 0, COMB_NONE, nullptr);
        return rc ? rc : run_chain(B, d, post, COMB_IMAGE, &ta);
    }
    return run_chain(A, d, post, COMB_SRC, nullptr);
}

// capi.zig:339-350 + morph.zig:532-571
int visar_morphology(const visar_image *src, const visar_image *dst, int op, int shape, int kw, int kh,
                     int iterations) {
    if (op < 0 || op > 6 || shape < 0 || shape > 2) return fail(UNSUPPORTED, "Unsupported");
    if (kw <= 0 || kh <= 0 || kw > 255 || kh > 255 || iterations < 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!src) return fail(INVALID_ARGUMENT, "InvalidArgument");
    uint32_t T = src->depth;
    if (!valid_depth(T)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (!same_shape(src, dst)) return fail(SIZE_MISMATCH, "SizeMismatch");
    std::vector<uint8_t> mask;
    structuring_element(shape, kw, kh, mask);
    MorphParams p;
    std::memset(&p, 0, sizeof p);
    p.depth = (int)T;
    p.kw = kw;
    p.kh = kh;
    p.ax = kw / 2;
    p.ay = kh / 2;
    if (element_runs(mask, kw, kh, p)) return fail(UNSUPPORTED, "Unsupported");
    const size_t elem = depth_size(T);
    Call call;
    DImg s, d;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(dst, d))) return rc;
    int st = OK;
    MorphRects mr;
    if (T == 0 && s.cn <= 4 && iterations > 0 && tuning().morph_chain != 1 && morph_rects(p, mr)) {
        st = morph_direct_ex(call, s, d, op, iterations, mr);
        if (st == UNSUPPORTED) st = morph_chain_ex(call, s, d, op, iterations, mr);
        if (st == OUT_OF_MEMORY) return fail(OUT_OF_MEMORY, "OutOfMemory");
        if (st != UNSUPPORTED) return call.finish(st);
        st = OK;
    }
    DImg t1, t2;
    if (op >= 2 && (call.temp_like(s, elem, t1) || call.temp_like(s, elem, t2))) return fail(OUT_OF_MEMORY, "OutOfMemory");
    switch (op) {
    case 0: st = morph_iter(call, s, d, p, 0, iterations, elem); break;
    case 1: st = morph_iter(call, s, d, p, 1, iterations, elem); break;
    case 2:
        if (!(st = morph_iter(call, s, t1, p, 0, iterations, elem))) st = morph_iter(call, t1, d, p, 1, iterations, elem);
        break;
    case 3:
        if (!(st = morph_iter(call, s, t1, p, 1, iterations, elem))) st = morph_iter(call, t1, d, p, 0, iterations, elem);
        break;
    case 4:
        if (!(st = morph_iter(call, s, t1, p, 0, iterations, elem)) && !(st = morph_iter(call, s, t2, p, 1, iterations, elem)))
            st = launch_sub_sat(t2, t1, d, (int)T, call.stream());
        break;
    case 5:
        if (!(st = morph_iter(call, s, t1, p, 0, iterations, elem)) && !(st = morph_iter(call, t1, t2, p, 1, iterations, elem)))
            st = launch_sub_sat(s, t2, d, (int)T, call.stream());
        break;
    default:
        if (!(st = morph_iter(call, s, t1, p, 1, iterations, elem)) && !(st = morph_iter(call, t1, t2, p, 0, iterations, elem)))
            st = launch_sub_sat(t2, s, d, (int)T, call.stream());
        break;
    }
    if (st == OUT_OF_MEMORY) return fail(OUT_OF_MEMORY, "OutOfMemory");
    return call.finish(st);
}

// ------------------------------------------------------------ background

}  // extern "C"

struct visar_background {
    uint32_t w, h;
    int shift;
    bool primed;
    uint16_t *model;  // device, w*h contiguous
};

extern "C" {

// capi.zig:446-456 + track.zig:22-24
int visar_background_create(uint32_t width, uint32_t height, int shift, visar_background **out) {
    if (shift < 0 || shift > 15) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (!out) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (width == 0 || height == 0) return fail(INVALID_ARGUMENT, "InvalidSize");
    int rc = ensure_runtime();
    if (rc) return rc;
    visar_background *b = new (std::nothrow) visar_background();
    if (!b) return fail(OUT_OF_MEMORY, "OutOfMemory");
    b->w = width;
    b->h = height;
    b->shift = shift;
    b->primed = false;
    if (hipMalloc(&b->model, (size_t)width * height * sizeof(uint16_t)) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return fail(OUT_OF_MEMORY, "OutOfMemory");
    }
    *out = b;
    return VISAR_OK;
}

// capi.zig:458-464 + track.zig:33-81
int visar_background_apply(visar_background *b, const visar_image *frame, const visar_image *mask, int threshold) {
    if (!b) return fail(INVALID_ARGUMENT, "InvalidArgument");
    int rc;
    if ((rc = check_image(frame, 0)) || (rc = check_image(mask, 0))) return rc;
    if (frame->channels != 1 || !same_shape(frame, mask) || frame->width != b->w || frame->height != b->h)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    int t = threshold < 0 ? 0 : threshold > 255 ? 255 : threshold;
    Call call;
    DImg f, m;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(frame, f)) || (rc = call.output(mask, m))) return rc;
    rc = launch_background(f, m, b->model, (long)b->w * 2, b->shift, t, b->primed ? 0 : 1, call.stream());
    if (rc == OK) b->primed = true;
    return call.finish(rc);
}

void visar_background_free(visar_background *b) {
    if (!b) return;
    (void)hipFree(b->model);
    delete b;
}

}  // extern "C"
