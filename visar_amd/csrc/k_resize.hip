// k_resize.hip -- the generic resize kernels: LINEAR and AREA for u16, i16
// and f32 samples and for u8 with more than 4 channels (reference
// resize.zig:47-60: the path that u8 with 1-4 channels does not take).
//
// One thread per destination sample (x, channel) of one destination row:
// nothing here is on a benchmarked workload, so the kernels are the plain
// per-sample form of the reference's arithmetic, in its operation order.
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ __attribute__((always_inline)) inline float tof(T v) {
    return (float)v;
}

// resize.zig:284-370 with Weights(T) of :401-418 for T != u8: f32 weights,
// h = s0 a0 + s1 a1 per source row and v = h0 b0 + h1 b1, unfused, then
// sat.cast(T, v).
template <typename T>
__global__ __launch_bounds__(kThreads) void resize_linear_f(DImg s, DImg d, const FLinTap *xt, const FLinTap *yt) {
    const int cn = d.cn;
    const int e = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (e >= d.w * cn) return;
    const int x = e / cn, ch = e - x * cn;
    const FLinTap tx = xt[x], ty = yt[y];
    const gptr<const T> r0 = grow<const T>(s, ty.i0), r1 = grow<const T>(s, ty.i1);
    const float h0 = tof(r0[tx.i0 + ch]) * tx.w0 + tof(r0[tx.i1 + ch]) * tx.w1;
    const float h1 = tof(r1[tx.i0 + ch]) * tx.w0 + tof(r1[tx.i1 + ch]) * tx.w1;
    grow<T>(d, y)[e] = sat_cast<T>((double)(h0 * ty.w0 + h1 * ty.w1));
}

// The same for u8 with Weights(u8): linearTapU8's 11-bit taps, integer
// horizontal sums, ((h >> 4) * b) >> 16 per row, (t + 2) >> 2 saturated.
__global__ __launch_bounds__(kThreads) void resize_linear_u8_any(DImg s, DImg d, const LinTap *xt, const LinTap *yt) {
    const int cn = d.cn;
    const int e = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (e >= d.w * cn) return;
    const int x = e / cn, ch = e - x * cn;
    const LinTap tx = xt[x], ty = yt[y];
    const gptr<const uint8_t> r0 = grow<const uint8_t>(s, ty.o0), r1 = grow<const uint8_t>(s, ty.o1);
    const int32_t h0 = ((int32_t)r0[tx.o0 + ch] * tx.w0 + (int32_t)r0[tx.o1 + ch] * tx.w1) >> 4;
    const int32_t h1 = ((int32_t)r1[tx.o0 + ch] * tx.w0 + (int32_t)r1[tx.o1 + ch] * tx.w1) >> 4;
    const int32_t t = ((h0 * ty.w0) >> 16) + ((h1 * ty.w1) >> 16);
    grow<uint8_t>(d, y)[e] = (uint8_t)min(max((t + 2) >> 2, 0), 255);
}

// resize.zig:420-457 areaInt: unsigned samples the block mean rounded half
// up in integers (u8 2x2 included: (sum + 2) >> 2), signed and float the
// f64 mean of the samples, sat.cast to T.
template <typename T>
__global__ __launch_bounds__(kThreads) void resize_area_int_any(DImg s, DImg d, int ix, int iy) {
    const int cn = d.cn;
    const int e = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (e >= d.w * cn) return;
    const int x = e / cn, ch = e - x * cn;
    const long base = (long)x * ix * cn + ch;
    const uint32_t area = (uint32_t)(ix * iy);
    if constexpr (std::is_same<T, uint8_t>::value || std::is_same<T, uint16_t>::value) {
        uint64_t sum = 0;
        for (int dy = 0; dy < iy; dy++) {
            const gptr<const T> r = grow<const T>(s, (long)y * iy + dy) + base;
            for (int dx = 0; dx < ix; dx++) sum += r[(long)dx * cn];
        }
        grow<T>(d, y)[e] = (T)((sum + area / 2) / area);
    } else {
        double sum = 0;
        for (int dy = 0; dy < iy; dy++) {
            const gptr<const T> r = grow<const T>(s, (long)y * iy + dy) + base;
            for (int dx = 0; dx < ix; dx++) sum += (double)tof(r[(long)dx * cn]);
        }
        grow<T>(d, y)[e] = sat_cast<T>(sum / (double)area);
    }
}

// resize.zig:527-565 areaGeneric: per destination row, the x table's f32
// sums of every source row of the row's y entries, each starting from 0 in
// table order, weighted by the y entry's alpha and summed from 0 in table
// order, no FMA, then sat.cast to T.
template <typename T>
__global__ __launch_bounds__(kThreads) void resize_area_gen_any(DImg s, DImg d, const AreaEntry *xe, const int32_t *xs,
                                                                const AreaEntry *ye, const int32_t *ys) {
    const int cn = d.cn;
    const int e = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int y = (int)blockIdx.y;
    if (e >= d.w * cn) return;
    const int x = e / cn, ch = e - x * cn;
    const int x0 = xs[x], x1 = xs[x + 1];
    float sum = 0.0f;
    for (int k = ys[y]; k < ys[y + 1]; k++) {
        const AreaEntry yk = ye[k];
        const gptr<const T> r = grow<const T>(s, yk.si);
        float h = 0.0f;
        for (int j = x0; j < x1; j++) {
            const AreaEntry xj = xe[j];
            h = h + tof(r[(long)xj.si * cn + ch]) * xj.alpha;
        }
        sum = sum + h * yk.alpha;
    }
    grow<T>(d, y)[e] = sat_cast<T>((double)sum);
}

inline dim3 rows_grid(const DImg &d) { return dim3((unsigned)(((long)d.w * d.cn + kThreads - 1) / kThreads), (unsigned)d.h); }
inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

}  // namespace

int launch_linear_any(const DImg &s, const DImg &d, int depth, const void *xt, const void *yt, hipStream_t st) {
    const dim3 g = rows_grid(d), b(kThreads);
    switch (depth) {
    case 0: hipLaunchKernelGGL(resize_linear_u8_any, g, b, 0, st, s, d, (const LinTap *)xt, (const LinTap *)yt); break;
    case 2: hipLaunchKernelGGL(resize_linear_f<uint16_t>, g, b, 0, st, s, d, (const FLinTap *)xt, (const FLinTap *)yt); break;
    case 3: hipLaunchKernelGGL(resize_linear_f<int16_t>, g, b, 0, st, s, d, (const FLinTap *)xt, (const FLinTap *)yt); break;
    case 5: hipLaunchKernelGGL(resize_linear_f<float>, g, b, 0, st, s, d, (const FLinTap *)xt, (const FLinTap *)yt); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_area_int_any(const DImg &s, const DImg &d, int depth, int ix, int iy, hipStream_t st) {
    const dim3 g = rows_grid(d), b(kThreads);
    switch (depth) {
    case 0: hipLaunchKernelGGL(resize_area_int_any<uint8_t>, g, b, 0, st, s, d, ix, iy); break;
    case 2: hipLaunchKernelGGL(resize_area_int_any<uint16_t>, g, b, 0, st, s, d, ix, iy); break;
    case 3: hipLaunchKernelGGL(resize_area_int_any<int16_t>, g, b, 0, st, s, d, ix, iy); break;
    case 5: hipLaunchKernelGGL(resize_area_int_any<float>, g, b, 0, st, s, d, ix, iy); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_area_gen_any(const DImg &s, const DImg &d, int depth, const AreaEntry *xe, const int32_t *xs,
                        const AreaEntry *ye, const int32_t *ys, hipStream_t st) {
    const dim3 g = rows_grid(d), b(kThreads);
    switch (depth) {
    case 0: hipLaunchKernelGGL(resize_area_gen_any<uint8_t>, g, b, 0, st, s, d, xe, xs, ye, ys); break;
    case 2: hipLaunchKernelGGL(resize_area_gen_any<uint16_t>, g, b, 0, st, s, d, xe, xs, ye, ys); break;
    case 3: hipLaunchKernelGGL(resize_area_gen_any<int16_t>, g, b, 0, st, s, d, xe, xs, ye, ys); break;
    case 5: hipLaunchKernelGGL(resize_area_gen_any<float>, g, b, 0, st, s, d, xe, xs, ye, ys); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

}  // namespace visar
