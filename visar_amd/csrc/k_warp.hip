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
#include <cmath>

#include "warp_blk.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;

// One output pixel of warpPerspective (geometry.zig:615-672).
template <typename T>
__device__ inline void persp_pixel(const PerspParams &p, int x, int y, T *o) {
    const DImg &s = p.s;
    const int cn = s.cn;
    const long W = s.w, H = s.h;
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
        long mx = border_map_inl(ix, W, p.border), my = border_map_inl(iy, H, p.border);
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
    long x0, x1, y0, y1;
    if (ix >= 0 && ix + 1 < W && iy >= 0 && iy + 1 < H) {
        x0 = ix, x1 = ix + 1, y0 = iy, y1 = iy + 1;
    } else {
        x0 = border_map_inl(ix, W, p.border), x1 = border_map_inl(ix + 1, W, p.border);
        y0 = border_map_inl(iy, H, p.border), y1 = border_map_inl(iy + 1, H, p.border);
    }
    float w0 = (1 - ax) * (1 - ay), w1 = ax * (1 - ay), w2 = (1 - ax) * ay, w3 = ax * ay;
    float cv = (float)p.cvd;
    for (int c = 0; c < cn; c++) {
        float acc = 0;
        acc += ((x0 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x0 * cn + c) : cv) * w0;
        acc += ((x1 >= 0 && y0 >= 0) ? ld_f<T>(s, y0, x1 * cn + c) : cv) * w1;
        acc += ((x0 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x0 * cn + c) : cv) * w2;
        acc += ((x1 >= 0 && y1 >= 0) ? ld_f<T>(s, y1, x1 * cn + c) : cv) * w3;
        o[c] = sat_cast<T>((double)acc);
    }
}

// Four consecutive bilinear u8 pixels of warpPerspective (geometry.zig:615-672).
__device__ inline uint32_t persp4_u8(const PerspParams &p, int x, int dx, int y) {
    int ix[4], iy[4];
    float ax[4], ay[4];
    const double fy = (double)y;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const double fx = (double)(x + j * dx);
        double sx = p.m[0] * fx + p.m[1] * fy + p.m[2];
        double sy = p.m[3] * fx + p.m[4] * fy + p.m[5];
        const double w = p.m[6] * fx + p.m[7] * fy + p.m[8];
        const double iw = w != 0 ? 1.0 / w : 0.0;
        sx *= iw;
        sy *= iw;
        const double x0f = floor(clampd(sx, -1e9, 1e9)), y0f = floor(clampd(sy, -1e9, 1e9));
        ax[j] = (float)(sx - x0f);
        ay[j] = (float)(sy - y0f);
        ix[j] = (int)x0f;
        iy[j] = (int)y0f;
    }
    Quad q[4];
    bool out[4];
    quads_u8(p.s, p.border, (float)p.cvd, ix, iy, q, out);
    uint32_t res = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float a = ax[j], b = ay[j];
        const float w0 = (1 - a) * (1 - b), w1 = a * (1 - b), w2 = (1 - a) * b, w3 = a * b;
        float acc = 0;
        acc += q[j].p00 * w0;
        acc += q[j].p01 * w1;
        acc += q[j].p10 * w2;
        acc += q[j].p11 * w3;
        res = pack_u8(out[j] ? (float)p.cvd : acc, (uint32_t)j, res);
    }
    return res;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void warp_persp_k(PerspParams p) {
    const DImg &d = p.d;
    const int y = blockIdx.y * 4 + tile_ty();
    const int x = blockIdx.x * 64 + tile_tx();
    if (y >= d.h || x >= d.w) return;
    persp_pixel<T>(p, x, y, row_ptr<T>(d, y) + (long)x * d.cn);
}

__global__ __launch_bounds__(kThreads) void warp_persp_lin4_u8(PerspParams p, TileOrder o) {
    const DImg &d = p.d;
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int y = by * 4 + tile_ty();
    const int x0 = bx * 256;
    if (y >= d.h || x0 + tile_tx() >= d.w) return;
    store4_strided_u8(d, x0, y, persp4_u8(p, x0 + tile_tx(), 64, y));
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

__global__ __launch_bounds__(kThreads) void warp_affine_lin4_u8(AffineParams p, TileOrder o) {
    int bx, by;
    if (tile_of(o, bx, by)) affine_lin4_tile(p, bx, by);
}

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void warp_affine_box_u8(AffineParams p, TileOrder o) {
    __shared__ uint32_t lds[kWarpLdsWords];
    int bx, by;
    if (tile_of(o, bx, by)) affine_box_tile(p, bx, by, lds);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void warp_affine_k(AffineParams p) {
    affine_px_tile<T>(p, blockIdx.x, blockIdx.y);
}

// The block-layout device functions (warp_blk.hpp) are shared with the
// batch kernels.


__global__ __launch_bounds__(kTileThreads) void warp_affine_c3_u8(AffineParams p, TileOrder o, uint32_t lim) {
    int bx, by;
    if (tile_of(o, bx, by)) affine_c3_u8_tile(p, lim, bx, by);
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void warp_affine_blk_u8(AffineParams p, TileOrder o, uint32_t lim) {
    __shared__ uint32_t lds[NW][64];
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int bx0 = bx * 4 * kBlkW, by0 = (by * NW + wv) * kBlkH;
    const int y = by0 + l / kBlkW;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2], byv = p.m[4] * yf + p.m[5];
    BlkCoord c[4];
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) c[j] = blk_affine_coord(p.m, bx0 + kBlkW * j + l % kBlkW, bxv, byv);
    blk_sample<false>(p.s, p.cv, lim, c, v);
    blk_store(p.d, bx0, by0, v, lds[wv]);
}

// The eager form: each wave T block rows of one 64-pixel column.
template <int NW, int T>
__global__ __launch_bounds__(64 * NW) void warp_affine_blk_pipe_u8(AffineParams p, TileOrder o, uint32_t lim) {
    __shared__ uint32_t lds[NW][64];
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int bx0 = bx * 4 * kBlkW, by0 = (by * NW + wv) * T * kBlkH;
    if (by0 >= p.d.h) return;
    const int rows_left = (p.d.h - by0 + kBlkH - 1) / kBlkH;
    auto make_coord = [&](int t) {
        const float yf = (float)(by0 + t * kBlkH + l / kBlkW);
        const float bxv = p.m[1] * yf + p.m[2], byv = p.m[4] * yf + p.m[5];
        return [&p, bx0, l, bxv, byv](int j) { return blk_affine_coord(p.m, bx0 + kBlkW * j + l % kBlkW, bxv, byv); };
    };
    blk_pipeline<false, T>(p.s, p.cv, lim, rows_left, make_coord,
                           [&](int t, const float (&v)[4]) { blk_store_wave(p.d, bx0, by0 + t * kBlkH, v, lds[wv]); });
}

// 1 / w in f64, bit for bit as the compiler's correctly rounded division
// computes it where v_div_scale leaves its operands unscaled and
// v_div_fixup passes the quotient through, which holds for 1 / w with
// |w| in [2^-60, 2^60]: the reciprocal estimate, two Newton-Raphson steps
// and one correction of the quotient, fused.
__device__ __attribute__((always_inline)) inline double recip_f64(double w) {
    double r = __builtin_amdgcn_rcp(w);
    double e = fma(-w, r, 1.0);
    r = fma(r, e, r);
    e = fma(-w, r, 1.0);
    r = fma(r, e, r);
    const double q = fma(-w, r, 1.0);
    return fma(q, r, r);
}

// Coordinates of pixel (x, y) of warpPerspective (geometry.zig:625-645):
// f64 map, x0 = floor(clamp(sx, +-1e9)), ax = f32(sx - x0). FAST: the host
// has found every |w| of the launch's pixels in [2^-60, 2^60] with one sign
// (w is affine in x and y, so its corners bound it), where recip_f64 equals
// the division and a coordinate beyond 1e9 only ever lands outside the
// image, so the clamp is left out; otherwise 1 / w (0 for w = 0) and the
// clamp. c1, c4, c7 are m[1] y, m[4] y, m[7] y.
template <bool FAST>
__device__ __attribute__((always_inline)) inline BlkCoord persp_coord(const double *m, int x, double c1, double c4,
                                                                      double c7) {
    const double fx = (double)x;
    double sx = m[0] * fx + c1 + m[2];
    double sy = m[3] * fx + c4 + m[5];
    const double w = m[6] * fx + c7 + m[8];
    const double iw = FAST ? recip_f64(w) : (w != 0 ? 1.0 / w : 0.0);
    sx *= iw;
    sy *= iw;
    if (!FAST) {
        sx = clampd(sx, -1e9, 1e9);
        sy = clampd(sy, -1e9, 1e9);
    }
    const double x0f = floor(sx), y0f = floor(sy);
    return BlkCoord{(int)x0f, (int)y0f, (float)(sx - x0f), (float)(sy - y0f)};
}

template <int NW, bool FAST>
__global__ __launch_bounds__(64 * NW) void warp_persp_blk_u8(PerspParams p, TileOrder o, uint32_t lim) {
    __shared__ uint32_t lds[NW][64];
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int bx0 = bx * 4 * kBlkW, by0 = (by * NW + wv) * kBlkH;
    const double fy = (double)(by0 + l / kBlkW);
    const double c1 = p.m[1] * fy, c4 = p.m[4] * fy, c7 = p.m[7] * fy;
    BlkCoord c[4];
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) c[j] = persp_coord<FAST>(p.m, bx0 + kBlkW * j + l % kBlkW, c1, c4, c7);
    blk_sample<true>(p.s, (float)p.cvd, lim, c, v);
    blk_store(p.d, bx0, by0, v, lds[wv]);
}

// The eager form: each wave T block rows of one 64-pixel column.
template <int NW, int T, bool FAST>
__global__ __launch_bounds__(64 * NW) void warp_persp_blk_pipe_u8(PerspParams p, TileOrder o, uint32_t lim) {
    __shared__ uint32_t lds[NW][64];
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int bx0 = bx * 4 * kBlkW, by0 = (by * NW + wv) * T * kBlkH;
    if (by0 >= p.d.h) return;
    const int rows_left = (p.d.h - by0 + kBlkH - 1) / kBlkH;
    auto make_coord = [&](int t) {
        const double fy = (double)(by0 + t * kBlkH + l / kBlkW);
        const double c1 = p.m[1] * fy, c4 = p.m[4] * fy, c7 = p.m[7] * fy;
        return [&p, bx0, l, c1, c4, c7](int j) { return persp_coord<FAST>(p.m, bx0 + kBlkW * j + l % kBlkW, c1, c4, c7); };
    };
    blk_pipeline<true, T>(p.s, (float)p.cvd, lim, rows_left, make_coord,
                          [&](int t, const float (&v)[4]) { blk_store_wave(p.d, bx0, by0 + t * kBlkH, v, lds[wv]); });
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

inline dim3 grid2d(int w, int h, int px) { return dim3((unsigned)((w + 64 * px - 1) / (64 * px)), (unsigned)((h + 3) / 4)); }

}  // namespace

// Block rows per wave of the eager block kernels (tuning "warp_blocks"):
// one by default. Two, with the second block's loads issued before the first
// is blended, measured slower (1080p: affine 5.6 -> 6.3 us, perspective 6.4
// -> 7.3 us): the kernels are bound by VALU issue with every wave of the
// launch resident at once, and the pipelined form needs about twice the
// registers.
inline int warp_blocks_per_wave() { return tuning().warp_blocks == 2 ? 2 : 1; }

inline bool persp_finite(const PerspParams &p) {
    for (int i = 0; i < 9; i++)
        if (!std::isfinite(p.m[i])) return false;
    return std::isfinite(p.cvd);
}

int launch_warp_affine(const AffineParams &p, hipStream_t st) {
    const dim3 b(kTileThreads);
    const uint32_t lim = blk_limit(p.s);
    if (p.depth == 0 && p.d.cn == 1 && p.interp == 1 && p.border == BORDER_CONSTANT && lim && affine_bounded(p) &&
        tuning().warp_layout != 1) {
        constexpr int NW = 2;
        if (p.s.w >= 2 && p.s.h >= 2 && tuning().warp_layout != 2) {
            const int T = warp_blocks_per_wave();
            const TileOrder o = tile_order((p.d.w + 63) / 64, (p.d.h + 4 * NW * T - 1) / (4 * NW * T), true);
            const dim3 g(tile_grid(o.ntiles)), t(64 * NW);
            if (T == 1) hipLaunchKernelGGL((warp_affine_blk_pipe_u8<NW, 1>), g, t, 0, st, p, o, lim);
            else hipLaunchKernelGGL((warp_affine_blk_pipe_u8<NW, 2>), g, t, 0, st, p, o, lim);
        } else {
            const TileOrder o = tile_order((p.d.w + 63) / 64, (p.d.h + 4 * NW - 1) / (4 * NW), true);
            hipLaunchKernelGGL((warp_affine_blk_u8<NW>), dim3(tile_grid(o.ntiles)), dim3(64 * NW), 0, st, p, o, lim);
        }
        return launched();
    }
    if (p.depth == 0 && p.d.cn == 3 && p.interp == 1 && p.border == BORDER_CONSTANT && tap_limit(p.s)) {
        const TileOrder o = tile_order((p.d.w + 63) / 64, (p.d.h + 3) / 4, true);
        hipLaunchKernelGGL(warp_affine_c3_u8, dim3(tile_grid(o.ntiles)), b, 0, st, p, o, tap_limit(p.s));
        return launched();
    }
    if (p.depth == 0 && p.d.cn == 1 && p.interp == 1 && p.border == BORDER_CONSTANT) {
        const TileOrder o = tile_order((p.d.w + kWarpTW - 1) / kWarpTW, (p.d.h + kWarpTH - 1) / kWarpTH, true);
        hipLaunchKernelGGL(warp_affine_box_u8, dim3(tile_grid(o.ntiles)), b, 0, st, p, o);
        return launched();
    }
    if (p.depth == 0 && p.d.cn == 1 && p.interp == 1) {
        const TileOrder o = tile_order((p.d.w + 255) / 256, (p.d.h + 3) / 4);
        hipLaunchKernelGGL(warp_affine_lin4_u8, dim3(tile_grid(o.ntiles)), b, 0, st, p, o);
        return launched();
    }
    const dim3 g = grid2d(p.d.w, p.d.h, 1);
    switch (p.depth) {
    case 0: hipLaunchKernelGGL(warp_affine_k<uint8_t>, g, b, 0, st, p); break;
    case 2: hipLaunchKernelGGL(warp_affine_k<uint16_t>, g, b, 0, st, p); break;
    case 3: hipLaunchKernelGGL(warp_affine_k<int16_t>, g, b, 0, st, p); break;
    case 5: hipLaunchKernelGGL(warp_affine_k<float>, g, b, 0, st, p); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_warp_perspective(const PerspParams &p, hipStream_t st) {
    const dim3 b(kTileThreads);
    const uint32_t lim = blk_limit(p.s);
    if (p.depth == 0 && p.d.cn == 1 && p.interp == 1 && p.border == BORDER_CONSTANT && lim && persp_finite(p) &&
        tuning().warp_layout != 1) {
        constexpr int NW = 4;
        // XCD bands: neighbouring tiles' footprints overlap, and in row-major
        // order each XCD's L2 fetches them again (3.07x the input against 1.01x).
        const TileOrder o = tile_order((p.d.w + 63) / 64, (p.d.h + 4 * NW - 1) / (4 * NW), true);
        // w over the launched pixels (the tiles cover x < 64 tiles_x, y < 16 tiles_y).
        const double xs[2] = {0.0, 64.0 * o.tiles_x - 1}, ys[2] = {0.0, 16.0 * (o.ntiles / o.tiles_x) - 1};
        // Every corner well away from 0 relative to the terms' size, so that
        // the device's rounded w keeps the sign and range of the exact one.
        const double scale = std::fabs(p.m[6]) * xs[1] + std::fabs(p.m[7]) * ys[1] + std::fabs(p.m[8]);
        bool fast = true;
        int sign = 0;
        for (double x : xs)
            for (double y : ys) {
                const double w = p.m[6] * x + p.m[7] * y + p.m[8];
                const int sg = w > 0 ? 1 : w < 0 ? -1 : 0;
                fast = fast && sg != 0 && (sign == 0 || sg == sign) && std::fabs(w) >= 0x1p-60 &&
                       std::fabs(w) <= 0x1p60 && std::fabs(w) >= 1e-9 * scale;
                sign = sg;
            }
        const bool eager = p.s.w >= 2 && p.s.h >= 2 && tuning().warp_layout != 2;
        const dim3 t(64 * NW);
        if (eager) {
            const int T = warp_blocks_per_wave();
            const TileOrder ot = tile_order((p.d.w + 63) / 64, (p.d.h + 4 * NW * T - 1) / (4 * NW * T), true);
            const dim3 g(tile_grid(ot.ntiles));
#define PERSP_PIPE(TT, F) hipLaunchKernelGGL((warp_persp_blk_pipe_u8<NW, TT, F>), g, t, 0, st, p, ot, lim)
            if (T == 1) { if (fast) PERSP_PIPE(1, true); else PERSP_PIPE(1, false); }
            else { if (fast) PERSP_PIPE(2, true); else PERSP_PIPE(2, false); }
#undef PERSP_PIPE
        } else {
            const dim3 g(tile_grid(o.ntiles));
            if (fast) hipLaunchKernelGGL((warp_persp_blk_u8<NW, true>), g, t, 0, st, p, o, lim);
            else hipLaunchKernelGGL((warp_persp_blk_u8<NW, false>), g, t, 0, st, p, o, lim);
        }
        return launched();
    }
    if (p.depth == 0 && p.d.cn == 1 && p.interp == 1) {
        const TileOrder o = tile_order((p.d.w + 255) / 256, (p.d.h + 3) / 4, true);
        hipLaunchKernelGGL(warp_persp_lin4_u8, dim3(tile_grid(o.ntiles)), b, 0, st, p, o);
        return launched();
    }
    const dim3 g = grid2d(p.d.w, p.d.h, 1);
    switch (p.depth) {
    case 0: hipLaunchKernelGGL(warp_persp_k<uint8_t>, g, b, 0, st, p); break;
    case 2: hipLaunchKernelGGL(warp_persp_k<uint16_t>, g, b, 0, st, p); break;
    case 3: hipLaunchKernelGGL(warp_persp_k<int16_t>, g, b, 0, st, p); break;
    case 5: hipLaunchKernelGGL(warp_persp_k<float>, g, b, 0, st, p); break;
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
