// k_undistort.hip -- lens undistortion for gfx950 (calib3d/undistort.zig).
//
//   undistort_u8          undistort.zig:205-272: the f64 map of every
//                         destination pixel, rounded to 1/32 px, then the
//                         15-bit fixed-point bilinear blend, constant border 0
//   undistort_map_f32     undistort.zig:82-103 initUndistortRectifyMap into
//                         f32 maps
//
// Every f64 operation is rounded separately (the library builds with
// -ffp-contract=off) and 1 / x is IEEE division, so the map is the
// reference's bit for bit.
#include "tiles.hpp"
#include "undistort.hpp"

namespace visar {

namespace {

constexpr int kUdThreads = 256;
constexpr int kUdPx = 4;  // pixels per lane, 64 apart along the row
constexpr int kUdTileW = 64 * kUdPx, kUdTileH = kUdThreads / 64;

// One destination row of MapParams.row (undistort.zig:44-78): the terms
// that depend on the row only, then map(j) per pixel.
struct RowMap {
    double ir0, ir3, ir6, rx, ry, rw;
    bool proj;

    __device__ __attribute__((always_inline)) RowMap(const double ir[9], double i) {
        ir0 = ir[0];
        ir3 = ir[3];
        ir6 = ir[6];
        rx = i * ir[1] + ir[2];
        ry = i * ir[4] + ir[5];
        rw = i * ir[7] + ir[8];
        proj = !(ir[6] == 0 && ir[7] == 0 && ir[8] == 1);
    }

    __device__ __attribute__((always_inline)) void map(const UndistortLens &L, double j, double &u, double &v) const {
        const double _x = rx + j * ir0, _y = ry + j * ir3;
        double x = _x, y = _y;
        if (proj) {
            const double w = 1 / (rw + j * ir6);
            x = _x * w;
            y = _y * w;
        }
        const double *k = L.k;
        const double x2 = x * x, y2 = y * y, r2 = x2 + y2, _2xy = 2 * x * y;
        double kr = 1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2;
        if (L.rational) kr /= 1 + ((k[7] * r2 + k[6]) * r2 + k[5]) * r2;
        const double xd = x * kr + k[2] * _2xy + k[3] * (r2 + 2 * x2);
        const double yd = y * kr + k[2] * (r2 + 2 * y2) + k[3] * _2xy;
        u = L.fx * xd + L.u0;
        v = L.fy * yd + L.v0;
    }
};

// undistort.zig:106-115 fix: 1/32 px, clamped to +-(2^31 - 1) with max
// before min (a NaN becomes -lim), rounded half to even.
__device__ __attribute__((always_inline)) inline int32_t fix32(double v) {
    const double lim = 2147483647.0;
    return (int32_t)rint(fmin(fmax(v * 32.0, -lim), lim));
}

// undistort.zig:120-157 sample: the inside, outside and per-tap border
// branches are one formula, sum over taps of (tap inside ? pixel : 0) * w,
// so every lane takes the same path: tap addresses are clamped into the
// image (always in bounds) and the taps outside contribute 0. ALIGNED
// (destination rows 4-byte aligned): each wave stages its row of the tile in
// LDS and stores it as 4 * CN contiguous bytes per lane (dword stores)
// instead of byte stores. WIN (2-4 channels, source rows 4-byte aligned and
// at least 12 bytes): a tap row's two pixels (2 CN bytes) come from two or
// three aligned dword loads, unpacked with alignbyte, instead of 2 CN byte
// loads; pixels whose taps reach past the left or right edge, or whose
// window would cross the row's end, are loaded byte by byte afterwards.
template <int CN, bool ALIGNED, bool WIN>
__global__ __launch_bounds__(kUdThreads) void undistort_u8(UndistortArgs p, TileOrder o) {
    __shared__ uint32_t stage[kUdTileH][kUdTileW * CN / 4];
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int y = by * kUdTileH + tile_ty();
    if (y >= p.d.h) return;
    // The stripe's own inverse (undistort.zig:241-247): new_k with its
    // principal point moved up by the stripe's first row.
    const int y0 = y / p.stripe * p.stripe;
    double ar[9], ir[9];
#pragma unroll
    for (int q = 0; q < 9; q++) ar[q] = p.ar[q];
    ar[5] = p.ar[5] - (double)y0;
    if (!inv3(ar, ir)) return;
    const RowMap rm(ir, (double)(y - y0));
    const int W = p.s.w, H = p.s.h;
    int32_t iu[kUdPx], iv[kUdPx];
    const int xb = bx * kUdTileW + tile_tx();
#pragma unroll
    for (int q = 0; q < kUdPx; q++) {
        double u, v;
        rm.map(p.lens, (double)(xb + 64 * q), u, v);
        iu[q] = fix32(u);
        iv[q] = fix32(v);
    }
    uint8_t t[kUdPx][4][CN];
    int32_t wq[kUdPx][4];
    long cxs[kUdPx][2];
    gptr<const uint8_t> rs[kUdPx][2];
    bool slow[kUdPx];
#pragma unroll
    for (int q = 0; q < kUdPx; q++) {
        const int32_t sx = iu[q] >> 5, sy = iv[q] >> 5, a = iu[q] & 31, b = iv[q] & 31;
        const bool inx0 = sx >= 0 && sx < W, inx1 = sx + 1 >= 0 && sx + 1 < W;
        const bool iny0 = sy >= 0 && sy < H, iny1 = sy + 1 >= 0 && sy + 1 < H;
        wq[q][0] = inx0 && iny0 ? (32 - a) * (32 - b) : 0;
        wq[q][1] = inx1 && iny0 ? a * (32 - b) : 0;
        wq[q][2] = inx0 && iny1 ? (32 - a) * b : 0;
        wq[q][3] = inx1 && iny1 ? a * b : 0;
        cxs[q][0] = min(max(sx, 0), W - 1);
        cxs[q][1] = min(max(sx + 1, 0), W - 1);
        const long cy0 = min(max(sy, 0), H - 1), cy1 = min(max(sy + 1, 0), H - 1);
        rs[q][0] = grow<const uint8_t>(p.s, cy0);
        rs[q][1] = grow<const uint8_t>(p.s, cy1);
        slow[q] = true;
        if constexpr (WIN) {
            // Window: dwords from the aligned-down start of tap 0's bytes.
            constexpr int ND = CN <= 2 ? 2 : 3;
            const long b0 = cxs[q][0] * CN, a0 = b0 & ~3l;
            const long lim = (long)W * CN - 4 * ND;
            slow[q] = !(inx0 && inx1) || a0 > lim;
            const long aw = min(a0, lim & ~3l);  // in bounds for every lane
            const int sh = (int)(b0 - a0);
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const gptr<const uint32_t> wp = (gptr<const uint32_t>)(rs[q][r] + aw);
                const uint32_t w0 = wp[0], w1 = wp[1], w2 = ND == 3 ? wp[2] : 0u;
                const uint32_t lo = __builtin_amdgcn_alignbyte(w1, w0, sh), hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
                const uint64_t v = (uint64_t)hi << 32 | lo;
#pragma unroll
                for (int c = 0; c < CN; c++) {
                    t[q][2 * r][c] = (uint8_t)(v >> (8 * c));
                    t[q][2 * r + 1][c] = (uint8_t)(v >> (8 * (CN + c)));
                }
            }
        }
    }
    bool any_slow = false;
#pragma unroll
    for (int q = 0; q < kUdPx; q++) any_slow |= slow[q];
    if (!WIN || __any(any_slow)) {
#pragma unroll
        for (int q = 0; q < kUdPx; q++) {
            if (!slow[q]) continue;
#pragma unroll
            for (int c = 0; c < CN; c++) {
                t[q][0][c] = rs[q][0][cxs[q][0] * CN + c];
                t[q][1][c] = rs[q][0][cxs[q][1] * CN + c];
                t[q][2][c] = rs[q][1][cxs[q][0] * CN + c];
                t[q][3][c] = rs[q][1][cxs[q][1] * CN + c];
            }
        }
    }
    uint8_t res[kUdPx][CN];
#pragma unroll
    for (int q = 0; q < kUdPx; q++)
#pragma unroll
        for (int c = 0; c < CN; c++)
            res[q][c] = (uint8_t)((t[q][0][c] * wq[q][0] + t[q][1][c] * wq[q][1] + t[q][2][c] * wq[q][2] +
                                   t[q][3][c] * wq[q][3] + 512) >> 10);
    const gptr<uint8_t> out = grow<uint8_t>(p.d, y);
    if constexpr (ALIGNED) {
        uint8_t *row = reinterpret_cast<uint8_t *>(stage[tile_ty()]);
#pragma unroll
        for (int q = 0; q < kUdPx; q++)
#pragma unroll
            for (int c = 0; c < CN; c++) row[(tile_tx() + 64 * q) * CN + c] = res[q][c];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Lane l stores bytes [4 CN l, 4 CN (l + 1)) of the wave's row span.
        const int b0 = tile_tx() * 4 * CN, valid = (p.d.w - bx * kUdTileW) * CN;
        const gptr<uint8_t> dst = out + (long)bx * kUdTileW * CN + b0;
        if (b0 + 4 * CN <= valid) {
#pragma unroll
            for (int k = 0; k < CN; k++) ((gptr<uint32_t>)dst)[k] = stage[tile_ty()][tile_tx() * CN + k];
        } else {
            for (int k = 0; k < 4 * CN && b0 + k < valid; k++) dst[k] = row[b0 + k];
        }
    } else {
#pragma unroll
        for (int q = 0; q < kUdPx; q++) {
            const int x = xb + 64 * q;
            if (x >= p.d.w) break;
#pragma unroll
            for (int c = 0; c < CN; c++) out[(long)x * CN + c] = res[q][c];
        }
    }
}

__global__ __launch_bounds__(kUdThreads) void undistort_map_f32(UndistortMapArgs p, TileOrder o) {
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int y = by * kUdTileH + tile_ty();
    if (y >= p.mx.h) return;
    const RowMap rm(p.ir, (double)y);
    const gptr<float> ox = (gptr<float>)grow<uint8_t>(p.mx, y), oy = (gptr<float>)grow<uint8_t>(p.my, y);
    const int xb = bx * kUdTileW + tile_tx();
#pragma unroll
    for (int q = 0; q < kUdPx; q++) {
        const int x = xb + 64 * q;
        if (x >= p.mx.w) break;
        double u, v;
        rm.map(p.lens, (double)x, u, v);
        ox[x] = (float)u;
        oy[x] = (float)v;
    }
}

inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

}  // namespace

int launch_undistort(const UndistortArgs &p, hipStream_t st) {
    const TileOrder o = tile_order((p.d.w + kUdTileW - 1) / kUdTileW, (p.d.h + kUdTileH - 1) / kUdTileH, true);
    const dim3 g(tile_grid(o.ntiles)), b(kUdThreads);
    const bool al = (uintptr_t)p.d.p % 4 == 0 && p.d.stride % 4 == 0;
    const bool win = (uintptr_t)p.s.p % 4 == 0 && p.s.stride % 4 == 0 && (long)p.s.w * p.s.cn >= 12;
    switch (p.s.cn * 4 + al * 2 + win) {
    case 4: case 5: hipLaunchKernelGGL((undistort_u8<1, false, false>), g, b, 0, st, p, o); break;
    case 6: case 7: hipLaunchKernelGGL((undistort_u8<1, true, false>), g, b, 0, st, p, o); break;
    case 8: hipLaunchKernelGGL((undistort_u8<2, false, false>), g, b, 0, st, p, o); break;
    case 9: hipLaunchKernelGGL((undistort_u8<2, false, true>), g, b, 0, st, p, o); break;
    case 10: hipLaunchKernelGGL((undistort_u8<2, true, false>), g, b, 0, st, p, o); break;
    case 11: hipLaunchKernelGGL((undistort_u8<2, true, true>), g, b, 0, st, p, o); break;
    case 12: hipLaunchKernelGGL((undistort_u8<3, false, false>), g, b, 0, st, p, o); break;
    case 13: hipLaunchKernelGGL((undistort_u8<3, false, true>), g, b, 0, st, p, o); break;
    case 14: hipLaunchKernelGGL((undistort_u8<3, true, false>), g, b, 0, st, p, o); break;
    case 15: hipLaunchKernelGGL((undistort_u8<3, true, true>), g, b, 0, st, p, o); break;
    case 16: hipLaunchKernelGGL((undistort_u8<4, false, false>), g, b, 0, st, p, o); break;
    case 17: hipLaunchKernelGGL((undistort_u8<4, false, true>), g, b, 0, st, p, o); break;
    case 18: hipLaunchKernelGGL((undistort_u8<4, true, false>), g, b, 0, st, p, o); break;
    case 19: hipLaunchKernelGGL((undistort_u8<4, true, true>), g, b, 0, st, p, o); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_undistort_map(const UndistortMapArgs &p, hipStream_t st) {
    const TileOrder o = tile_order((p.mx.w + kUdTileW - 1) / kUdTileW, (p.mx.h + kUdTileH - 1) / kUdTileH);
    hipLaunchKernelGGL(undistort_map_f32, dim3(tile_grid(o.ntiles)), dim3(kUdThreads), 0, st, p, o);
    return launched();
}

}  // namespace visar
