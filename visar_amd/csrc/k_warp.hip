// k_warp.hip -- warpAffine, warpPerspective and remap for gfx950.
//
// The reference warps are float kernels, not fixed point, so each output
// pixel reproduces their IEEE operation sequence exactly:
//   warpAffine  geometry.zig:410-533 (f32 inverse, fma coordinates, fma blend)
//   warpPerspective geometry.zig:615-672 (f64 coordinates, f32 weights, no fma)
//   remap       undistort.zig:280-394 (f32 maps, fma blend, magic rounding)
// One thread computes one output pixel (all channels): stores are
// coalesced along the row, the four taps are gathers that hit L2/MALL
// because neighbouring threads read neighbouring source pixels.
#include "kernels.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ inline float ld_f(const DImg &s, long y, long i) {
    return (float)row_ptr<const T>(s, y)[i];
}

// geometry.zig:504-506 roundTo
template <typename T>
__device__ inline T round_to(float v) {
    return sat_cast<T>((double)v);
}

// geometry.zig:492-494 toIdx
__device__ inline long to_idx(float f) { return (long)min_num(max_num(f, -1.0e9f), 1.0e9f); }

template <typename T>
__global__ __launch_bounds__(kThreads) void warp_affine_k(AffineParams p) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    const DImg &s = p.s, &d = p.d;
    long y = gid / d.w, x = gid - y * d.w;
    if (y >= d.h) return;
    const int cn = d.cn;
    const long W = s.w, H = s.h;
    T *o = row_ptr<T>(d, y) + x * cn;
    float yf = (float)y;
    float bx = p.m[1] * yf + p.m[2];
    float by = p.m[4] * yf + p.m[5];
    float xf = (float)x;
    float sx = fmaf(p.m[0], xf, bx);
    float sy = fmaf(p.m[3], xf, by);
    if (p.interp == 0) {
        // geometry.zig:496-502 nearestPixel
        long ix = to_idx(rint_magic(min_num(max_num(sx, -4.0e6f), 4.0e6f)));
        long iy = to_idx(rint_magic(min_num(max_num(sy, -4.0e6f), 4.0e6f)));
        long mx = border_map(ix, W, p.border), my = border_map(iy, H, p.border);
        if (mx >= 0 && my >= 0) {
            const T *sp = row_ptr<const T>(s, my) + mx * cn;
            for (int c = 0; c < cn; c++) o[c] = sp[c];
        } else {
            for (int c = 0; c < cn; c++) o[c] = (T)p.cvd;
        }
        return;
    }
    // geometry.zig:508-533 linearPixel (the 16-pixel fast block computes the same values)
    float fx = floorf(sx), fy = floorf(sy);
    float ax = sx - fx, ay = sy - fy;
    long ix = to_idx(fx), iy = to_idx(fy);
    if (p.border == BORDER_CONSTANT && (ix < -1 || iy < -1 || ix >= W || iy >= H)) {
        for (int c = 0; c < cn; c++) o[c] = (T)p.cvd;
        return;
    }
    long x0 = border_map(ix, W, p.border), x1 = border_map(ix + 1, W, p.border);
    long y0 = border_map(iy, H, p.border), y1 = border_map(iy + 1, H, p.border);
    for (int c = 0; c < cn; c++) {
        float p00 = (x0 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x0 * cn + c) : p.cv;
        float p01 = (x1 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x1 * cn + c) : p.cv;
        float p10 = (x0 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x0 * cn + c) : p.cv;
        float p11 = (x1 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x1 * cn + c) : p.cv;
        float v0 = fmaf(ax, p01 - p00, p00);
        float v1 = fmaf(ax, p11 - p10, p10);
        o[c] = round_to<T>(fmaf(ay, v1 - v0, v0));
    }
}

// std.math.clamp(v, lo, hi) = @max(lo, @min(v, hi)), NaN-ignoring.
__device__ inline double clampd(double v, double lo, double hi) { return max_num(lo, min_num(v, hi)); }

template <typename T>
__global__ __launch_bounds__(kThreads) void warp_persp_k(PerspParams p) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    const DImg &s = p.s, &d = p.d;
    long y = gid / d.w, x = gid - y * d.w;
    if (y >= d.h) return;
    const int cn = d.cn;
    const long W = s.w, H = s.h;
    T *o = row_ptr<T>(d, y) + x * cn;
    double fx = (double)x, fy = (double)y;
    double sx = p.m[0] * fx + p.m[1] * fy + p.m[2];
    double sy = p.m[3] * fx + p.m[4] * fy + p.m[5];
    double w = p.m[6] * fx + p.m[7] * fy + p.m[8];
    double iw = w != 0 ? 1.0 / w : 0.0;
    sx *= iw;
    sy *= iw;
    if (p.interp == 0) {
        long ix = (long)round(clampd(sx, -1e9, 1e9));
        long iy = (long)round(clampd(sy, -1e9, 1e9));
        long mx = border_map(ix, W, p.border), my = border_map(iy, H, p.border);
        if (mx >= 0 && my >= 0) {
            const T *sp = row_ptr<const T>(s, my) + mx * cn;
            for (int c = 0; c < cn; c++) o[c] = sp[c];
        } else {
            for (int c = 0; c < cn; c++) o[c] = (T)p.cvd;
        }
        return;
    }
    double x0f = floor(clampd(sx, -1e9, 1e9));
    double y0f = floor(clampd(sy, -1e9, 1e9));
    float ax = (float)(sx - x0f), ay = (float)(sy - y0f);
    long ix = (long)x0f, iy = (long)y0f;
    if (p.border == BORDER_CONSTANT && (ix < -1 || iy < -1 || ix >= W || iy >= H)) {
        for (int c = 0; c < cn; c++) o[c] = (T)p.cvd;
        return;
    }
    long x0 = border_map(ix, W, p.border), x1 = border_map(ix + 1, W, p.border);
    long y0 = border_map(iy, H, p.border), y1 = border_map(iy + 1, H, p.border);
    float w0 = (1 - ax) * (1 - ay), w1 = ax * (1 - ay), w2 = (1 - ax) * ay, w3 = ax * ay;
    float cv = (float)p.cvd;
    for (int c = 0; c < cn; c++) {
        float acc = 0;
        acc += ((x0 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x0 * cn + c) : cv) * w0;
        acc += ((x1 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x1 * cn + c) : cv) * w1;
        acc += ((x0 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x0 * cn + c) : cv) * w2;
        acc += ((x1 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x1 * cn + c) : cv) * w3;
        o[c] = round_to<T>(acc);
    }
}

// undistort.zig:365-394 linearF32 (the 16-lane all-inside block is identical).
__global__ __launch_bounds__(kThreads) void remap_k(DImg s, DImg d, DImg mxi, DImg myi, int bv) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / d.w, x = gid - y * d.w;
    if (y >= d.h) return;
    const int cn = d.cn;
    const long W = s.w, H = s.h;
    uint8_t *o = row_ptr<uint8_t>(d, y) + x * cn;
    float sx = row_ptr<const float>(mxi, y)[x], sy = row_ptr<const float>(myi, y)[x];
    float fx = floorf(sx), fy = floorf(sy);
    float ax = sx - fx, ay = sy - fy;
    if (!(fx >= -1 && fy >= -1 && fx < (float)W && fy < (float)H)) {
        for (int c = 0; c < cn; c++) o[c] = (uint8_t)bv;
        return;
    }
    long ix = (long)fx, iy = (long)fy;
    bool in00 = ix >= 0 && ix < W && iy >= 0 && iy < H;
    bool in01 = ix + 1 >= 0 && ix + 1 < W && iy >= 0 && iy < H;
    bool in10 = ix >= 0 && ix < W && iy + 1 >= 0 && iy + 1 < H;
    bool in11 = ix + 1 >= 0 && ix + 1 < W && iy + 1 >= 0 && iy + 1 < H;
    for (int c = 0; c < cn; c++) {
        float p00 = (float)(in00 ? row_ptr<const uint8_t>(s, iy)[ix * cn + c] : (uint8_t)bv);
        float p01 = (float)(in01 ? row_ptr<const uint8_t>(s, iy)[(ix + 1) * cn + c] : (uint8_t)bv);
        float p10 = (float)(in10 ? row_ptr<const uint8_t>(s, iy + 1)[ix * cn + c] : (uint8_t)bv);
        float p11 = (float)(in11 ? row_ptr<const uint8_t>(s, iy + 1)[(ix + 1) * cn + c] : (uint8_t)bv);
        float v0 = fmaf(ax, p01 - p00, p00);
        float v1 = fmaf(ax, p11 - p10, p10);
        float v = fmaf(ay, v1 - v0, v0);
        float r = rint_magic(v);
        o[c] = (uint8_t)min_num(max_num(r, 0.0f), 255.0f);
    }
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

}  // namespace

int launch_warp_affine(const AffineParams &p, hipStream_t st) {
    long n = (long)p.d.w * p.d.h;
    switch (p.depth) {
    case 0: hipLaunchKernelGGL(warp_affine_k<uint8_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 2: hipLaunchKernelGGL(warp_affine_k<uint16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 3: hipLaunchKernelGGL(warp_affine_k<int16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 5: hipLaunchKernelGGL(warp_affine_k<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_warp_perspective(const PerspParams &p, hipStream_t st) {
    long n = (long)p.d.w * p.d.h;
    switch (p.depth) {
    case 0: hipLaunchKernelGGL(warp_persp_k<uint8_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 2: hipLaunchKernelGGL(warp_persp_k<uint16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 3: hipLaunchKernelGGL(warp_persp_k<int16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    case 5: hipLaunchKernelGGL(warp_persp_k<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_remap(const DImg &s, const DImg &d, const DImg &mx, const DImg &my, int bv, hipStream_t st) {
    long n = (long)d.w * d.h;
    hipLaunchKernelGGL(remap_k, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, d, mx, my, bv);
    return launched();
}

}  // namespace visar
