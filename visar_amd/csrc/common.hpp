// common.hpp -- shared types and exact-arithmetic helpers of libvisar_amd.
//
// Every translation unit is built with -ffp-contract=off (no implicit FMA):
// float expressions evaluate exactly as written, and fmaf()/fma() appear only
// where VISAR's reference kernels use @mulAdd.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace visar {

// Status codes of visar_status (reference include/visar.h:22-30).
enum Status : int {
    OK = 0,
    INVALID_ARGUMENT = -1,
    SIZE_MISMATCH = -2,
    UNSUPPORTED = -3,
    OUT_OF_MEMORY = -4,
    INVALID_DATA = -5,
    INTERNAL = -6,
};

// Border kinds of reference src/imgproc/border.zig:5-18 (OpenCV numbering).
enum Border : int { BORDER_CONSTANT = 0, BORDER_REPLICATE = 1, BORDER_REFLECT = 2, BORDER_WRAP = 3, BORDER_REFLECT101 = 4 };

// A device view of one image: interleaved HWC samples, stride in bytes.
struct DImg {
    uint8_t *p;
    int w, h, cn;
    long stride;
};

template <typename T>
__host__ __device__ inline T *row_ptr(const DImg &im, long y) {
    return reinterpret_cast<T *>(im.p + y * im.stride);
}

// border.zig:22-39: maps a coordinate into [0, n), or -1 for a constant border.
// 32-bit arithmetic: image sides are at most 2^30 and coordinates within
// +-1e9 (the warps clamp them there first). The reference reflects
// repeatedly; the results are periodic (period 2n for reflect, 2n - 2 for
// reflect-101), which the closed forms below use.
__host__ __device__ inline int border_map_far(int p, int n, int border) {
    switch (border) {
    case BORDER_CONSTANT: return -1;
    case BORDER_REPLICATE: return p < 0 ? 0 : n - 1;
    case BORDER_REFLECT: {
        int q = p % (2 * n);
        if (q < 0) q += 2 * n;
        return q >= n ? 2 * n - 1 - q : q;
    }
    case BORDER_REFLECT101: {
        if (n == 1) return 0;
        int q = p % (2 * n - 2);
        if (q < 0) q += 2 * n - 2;
        return q >= n ? 2 * n - 2 - q : q;
    }
    case BORDER_WRAP: {
        int m = p % n;
        return m < 0 ? m + n : m;
    }
    default: return -1;
    }
}

__host__ __device__ __attribute__((noinline)) inline int border_map_outside(int p, int n, int border) {
    return border_map_far(p, n, border);
}

// One reflection covers everything a stencil halo or a warp footprint
// reaches on images at least as wide as the halo: p in [-(n - 1), 2n - 2].
template <bool FAR_INLINE>
__host__ __device__ inline int border_map_t(int p, int n, int border) {
    if ((unsigned)p < (unsigned)n) return p;
    if (border == BORDER_CONSTANT) return -1;
    if (border == BORDER_REPLICATE) return p < 0 ? 0 : n - 1;
    if (p >= -(n - 1) && p <= 2 * n - 2) {
        if (border == BORDER_REFLECT101) return p < 0 ? -p : 2 * n - 2 - p;
        if (border == BORDER_REFLECT) return p < 0 ? -p - 1 : 2 * n - 1 - p;
        if (border == BORDER_WRAP) return p < 0 ? p + n : p - n;
    }
    if constexpr (FAR_INLINE) return border_map_far(p, n, border);
    else return border_map_outside(p, n, border);
}

// Generic kernels: the far case is an out-of-line call (smaller code).
__host__ __device__ inline int border_map(int p, int n, int border) { return border_map_t<false>(p, n, border); }

// Stencil and morphology halos (p within a few pixels of the image, or far
// out only on images narrower than the halo): no call, no division, small
// code. Reflects repeatedly as border.zig does; the loop runs once per image
// width the coordinate lies outside.
__host__ __device__ inline int border_map_halo(int p, int n, int border) {
    if ((unsigned)p < (unsigned)n) return p;
    if (border == BORDER_CONSTANT) return -1;
    if (border == BORDER_REPLICATE) return p < 0 ? 0 : n - 1;
    if (border == BORDER_WRAP) {
        while (p < 0) p += n;
        while (p >= n) p -= n;
        return p;
    }
    if (n == 1) return 0;
    const int d = border == BORDER_REFLECT101 ? 1 : 0;
    while (p < 0 || p >= n) p = p < 0 ? -p - 1 + d : 2 * n - 1 - d - p;
    return p;
}

// Tile code: no call at all. A call site makes the compiler drain every
// outstanding load of the wave before it, which would serialise a tile's
// loads even when no lane takes the call.
__host__ __device__ inline int border_map_inl(int p, int n, int border) { return border_map_t<true>(p, n, border); }

// saturate.zig:8-18: round half to even with the 1.5*2^k magic constant.
__host__ __device__ inline float rint_magic(float x) { return (x + 12582912.0f) - 12582912.0f; }
__host__ __device__ inline double rint_magic(double x) { return (x + 6755399441055744.0) - 6755399441055744.0; }

// Zig's @min/@max on floats: a NaN operand yields the other operand.
__host__ __device__ inline float max_num(float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
__host__ __device__ inline float min_num(float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }
__host__ __device__ inline double max_num(double a, double b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
__host__ __device__ inline double min_num(double a, double b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }

template <typename T> struct Lim;
template <> struct Lim<uint8_t> { static constexpr double lo = 0, hi = 255; };
template <> struct Lim<uint16_t> { static constexpr double lo = 0, hi = 65535; };
template <> struct Lim<int16_t> { static constexpr double lo = -32768, hi = 32767; };
template <> struct Lim<int32_t> { static constexpr double lo = -2147483648.0, hi = 2147483647.0; };

// saturate.zig:21-49 sat.cast(D, float): NaN -> 0, round half to even in f64, clamp.
template <typename D>
__host__ __device__ inline D sat_cast(double x) {
    if constexpr (std::is_same<D, float>::value) {
        return (float)x;
    } else {
        if (x != x) return 0;
        double r = rint_magic(x);
        if (r <= Lim<D>::lo) return (D)Lim<D>::lo;
        if (r >= Lim<D>::hi) return (D)Lim<D>::hi;
        return (D)r;
    }
}

// filter.zig:44-58 satV(D, f32): round in f32, clamp with NaN -> low bound.
template <typename D>
__host__ __device__ inline D sat_v(float v) {
    if constexpr (std::is_same<D, float>::value) {
        return v;
    } else {
        float r = rint_magic(v);
        float c = min_num(max_num(r, (float)Lim<D>::lo), (float)Lim<D>::hi);
        return (D)c;
    }
}

// simd.zig:90-126 satF32(D, v): clamp (NaN -> low bound, 0 for signed), round.
template <typename D>
__host__ __device__ inline D sat_f32(float v) {
    float c = min_num(max_num(v, (float)Lim<D>::lo), (float)Lim<D>::hi);
    if (Lim<D>::lo != 0 && v != v) c = 0;
    return (D)rint_magic((double)c);
}

__host__ __device__ inline size_t depth_size(uint32_t depth) {
    switch (depth) {
    case 0: return 1;
    case 2: return 2;
    case 3: return 2;
    case 4: return 4;
    case 5: return 4;
    default: return 0;
    }
}

}  // namespace visar
