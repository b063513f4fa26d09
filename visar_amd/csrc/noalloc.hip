// noalloc.hip -- visar_threshold_noalloc: visar_threshold for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The automatic modes (Otsu, triangle) need one piece of device memory
// beyond the images: the 256-bin histogram of the source. visar_threshold
// takes it from the library's buffer pool; here it lives in the file-scope
// device array `scratch`, which every GPU the library runs on holds until
// the process ends. Its device address differs per GPU and is not the
// array's host-side name: it is looked up once per device with
// hipGetSymbolAddress and only that address is passed to kernels.
//
// Every call on one GPU shares that GPU's array: calls whose automatic
// thresholds overlap in time on one GPU race on the histogram. Each call
// clears the bins with a fill kernel before counting (no hipMemsetAsync, so
// the fixed-threshold form stays capturable).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/visar_gpu.h"
#include "kernels.hpp"
#include "runtime.hpp"

namespace visar {
namespace {

constexpr int kBins = 256;
// The histogram's u32 bins count at most 2^32 - 1 samples.
constexpr uint64_t kMaxSamples = 0xffffffffull;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kBins];

std::mutex g_mu;
uint32_t *g_ptr[kMaxDevices] = {};

__global__ __launch_bounds__(kBins) void clear_bins(uint32_t *bins) { bins[threadIdx.x] = 0; }

// The device address of this GPU's `scratch`, looked up on its first use.
int scratch_ptr(uint32_t **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return fail(INTERNAL, "NoDevice");
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(INTERNAL, "ScratchUnavailable");
        }
        g_ptr[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_ptr[dev];
    return OK;
}

DImg dview(const visar_image *c) {
    DImg d;
    d.p = static_cast<uint8_t *>(c->data);
    d.w = (int)c->width;
    d.h = (int)c->height;
    d.cn = (int)c->channels;
    d.stride = (long)c->stride;
    return d;
}

size_t span(const visar_image *c) {
    return (c->height - 1) * c->stride + (size_t)c->width * c->channels * depth_size(c->depth);
}

}  // namespace
}  // namespace visar

using namespace visar;

extern "C" {

int visar_threshold_noalloc(const visar_image *src, const visar_image *dst, double thresh, double maxval, int type,
                            double *used) {
    const int kind = type & 7;
    if (kind > 4) return fail(UNSUPPORTED, "Unsupported");
    const int autom = (type & 8) ? 1 : (type & 16) ? 2 : 0;
    if (!src || !dst) return fail(INVALID_ARGUMENT, "InvalidArgument");
    const uint32_t T = src->depth;
    if (!(T == 0 || T == 2 || T == 3 || T == 5)) return fail(UNSUPPORTED, "Unsupported");
    int rc;
    if ((rc = check_image(src, T)) || (rc = check_image(dst, T))) return rc;
    if (src->width != dst->width || src->height != dst->height || src->channels != dst->channels)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    if ((rc = ensure_runtime())) return rc;
    // Host images would need staging buffers: device memory only.
    if (!is_device_ptr(src->data) || !is_device_ptr(dst->data)) return fail(INVALID_ARGUMENT, "NotDeviceMemory");
    // The kernel is element-wise: the same view in and out is fine, any
    // other overlap would need a temporary.
    const uint8_t *a = static_cast<const uint8_t *>(src->data), *b = static_cast<const uint8_t *>(dst->data);
    const bool same = a == b && src->stride == dst->stride;
    if (!same && a < b + span(dst) && b < a + span(src)) return fail(INVALID_ARGUMENT, "Overlap");
    const hipStream_t st = current_stream();
    ThreshParams p;
    std::memset(&p, 0, sizeof p);
    p.s = dview(src);
    p.d = dview(dst);
    if (autom) {
        if (T != 0 || src->channels != 1) return fail(UNSUPPORTED, "Unsupported");
        if ((uint64_t)src->width * src->height > kMaxSamples) return fail(INVALID_ARGUMENT, "TooManyPixels");
        uint32_t *bins = nullptr;
        if ((rc = scratch_ptr(&bins))) return rc;
        uint32_t hist[kBins];
        hipLaunchKernelGGL(clear_bins, dim3(1), dim3(kBins), 0, st, bins);
        if (hipGetLastError() != hipSuccess || launch_hist_u8(p.s, bins, st) != OK ||
            hipMemcpyAsync(hist, bins, sizeof hist, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(INTERNAL, "KernelFailed");
        thresh = autom == 1 ? threshold_otsu(hist, (double)src->width * src->height) : threshold_triangle(hist);
    }
    if (used) *used = thresh;
    threshold_setup(p, T, kind, thresh, maxval);
    if (launch_threshold(p, st) != OK) return fail(INTERNAL, "KernelFailed");
    if (sync_mode() && hipStreamSynchronize(st) != hipSuccess) return fail(INTERNAL, "KernelFailed");
    return OK;
}

}  // extern "C"
