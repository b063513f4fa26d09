// tiles.hpp -- per-tile device functions shared by the plain kernels and the
// persistent batch kernel, so an image gives the same bytes whichever way
// it is launched.
//
// Every tile function is executed by one 256-thread workgroup, viewed as 64
// columns x 4 rows of threads (tx = threadIdx.x & 63, ty = threadIdx.x >> 6),
// and owns tile (bx, by) of its image. Functions that stage data in LDS take
// the workgroup's LDS buffer and end with no outstanding LDS use, so a
// persistent workgroup may run another tile right after a __syncthreads().
#pragma once

#include "kernels.hpp"

namespace visar {

constexpr int kTileThreads = 256;

__device__ inline int tile_tx() { return threadIdx.x & 63; }
__device__ inline int tile_ty() { return threadIdx.x >> 6; }

// ------------------------------------------------------------ stencils
//
// u8 sources, square 3/5/7 kernels (GaussianBlur u8 and Sobel u8 -> i16).
// Exact integer sums make the pass order free (reference filter.zig:163-167).
// The tile is 256 consecutive samples of the flat row x 4*TH output rows.
// The workgroup stages the rows it needs plus the column halo in LDS with the
// border rules applied while loading (border.zig:22-39): rows are mapped once,
// whole in-row dwords are coalesced dword loads, only dwords that straddle
// the row ends are assembled byte by byte. Each thread then runs the
// horizontal taps of its 4 samples on each staged row and the vertical taps
// over those results, with full-rate 24-bit multiplies (taps and partial
// sums stay below 2^23 in magnitude).

struct StencilArgs {
    DImg src, dst;
    int border, fin;
    float scale;
    int kx[7], ky[7];
};

template <int K, int CN, int TH>
struct StencilShape {
    static constexpr int R = K / 2;
    static constexpr int HW = (R * CN + 3) / 4;  // halo dwords on each side
    static constexpr int NW = 2 * HW + 1;        // dwords per row per thread
    static constexpr int PW = 64 + 2 * HW;       // dwords per staged row
    static constexpr int TR = 4 * TH + 2 * R;    // staged rows
    static constexpr int LDS_WORDS = TR * PW;
};

// Byte c of a source row with column border mapping (c may lie outside
// [0, n)): pixel floor(c / CN), channel c - CN * pixel.
template <int CN>
__device__ inline uint32_t mapped_byte(const uint8_t *row, int c, int W, int border) {
    const int px = (c + 64 * CN) / CN - 64;
    const int ch = c - px * CN;
    const int mx = border_map(px, W, border);
    return mx < 0 ? 0u : (uint32_t)row[mx * CN + ch];
}

template <typename D>
__device__ inline int stencil_finish(int a, int fin, float scale) {
    if (fin == FIN_FIX8) {
        int r = (a + (1 << 15)) >> 16;
        return r > 255 ? 255 : r;
    }
    if (fin == FIN_SCALED) return (int)sat_v<D>((float)a * scale);
    int lo = (int)Lim<D>::lo, hi = (int)Lim<D>::hi;
    return a < lo ? lo : a > hi ? hi : a;
}

template <typename D>
__device__ inline void store4(D *p, const int (&v)[4]);
template <>
__device__ inline void store4<uint8_t>(uint8_t *p, const int (&v)[4]) {
    *reinterpret_cast<uint32_t *>(p) = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
}
template <>
__device__ inline void store4<int16_t>(int16_t *p, const int (&v)[4]) {
    uint32_t a = (uint32_t)(uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16);
    uint32_t b = (uint32_t)(uint16_t)v[2] | ((uint32_t)(uint16_t)v[3] << 16);
    *reinterpret_cast<uint2 *>(p) = make_uint2(a, b);
}

template <int K, int CN, int TH, typename D>
__device__ inline void stencil_tile(const StencilArgs &p, int bx, int by, uint32_t *lds) {
    using S = StencilShape<K, CN, TH>;
    constexpr int R = S::R, HW = S::HW, NW = S::NW, PW = S::PW, TR = S::TR;
    uint32_t(*tile)[PW] = reinterpret_cast<uint32_t(*)[PW]>(lds);
    const int W = p.src.w, H = p.src.h;
    const int n = W * CN;
    const int x0 = bx * 256;
    const int y0 = by * 4 * TH;
    const bool aligned = ((p.src.stride | (uintptr_t)p.src.p) & 3) == 0;
    for (int idx = threadIdx.x; idx < TR * PW; idx += kTileThreads) {
        const int r = idx / PW, k = idx - r * PW;
        const int sy = border_map(y0 - R + r, H, p.border);
        const int c = x0 - 4 * HW + 4 * k;
        uint32_t v = 0;
        if (sy >= 0) {
            const uint8_t *row = p.src.p + (long)sy * p.src.stride;
            if (aligned && c >= 0 && c + 4 <= n) {
                v = *reinterpret_cast<const uint32_t *>(row + c);
            } else {
                for (int b = 0; b < 4; b++) v |= mapped_byte<CN>(row, c + b, W, p.border) << (8 * b);
            }
        }
        tile[r][k] = v;
    }
    __syncthreads();
    const int tx = tile_tx();
    const int x = x0 + tx * 4;
    const int yt = tile_ty() * TH;
    if (x >= n || y0 + yt >= H) return;
    int hz[TH + 2 * R][4];
#pragma unroll
    for (int r = 0; r < TH + 2 * R; r++) {
        uint32_t w[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = tile[yt + r][tx + k];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int a = 0;
#pragma unroll
            for (int i = 0; i < K; i++) {
                const int o = 4 * HW + j + (i - R) * CN;
                a += __mul24(p.kx[i], (int)((w[o >> 2] >> ((o & 3) * 8)) & 255u));
            }
            hz[r][j] = a;
        }
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;
        int o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int a = 0;
#pragma unroll
            for (int k = 0; k < K; k++) a += __mul24(p.ky[k], hz[t + k][j]);
            o[j] = stencil_finish<D>(a, p.fin, p.scale);
        }
        D *out = reinterpret_cast<D *>(p.dst.p + (long)y * p.dst.stride) + x;
        if (x + 4 <= n && (((uintptr_t)out | (uintptr_t)p.dst.stride) & (4 * sizeof(D) - 1)) == 0) {
            store4<D>(out, o);
        } else {
            for (int j = 0; j < 4 && x + j < n; j++) out[j] = (D)o[j];
        }
    }
}

// ---------------------------------------------------------- morphology
//
// u8 elements up to 7 x 7 in a KW x KW frame with a centred anchor. The
// element's on-cells are a per-row bit mask (byte ky of mask0/mask1), so the
// tap loop is fully unrolled and its conditions are uniform across the wave.
// Out-of-image cells are staged as the identity (0 for dilate, 255 for
// erode), which ignores them exactly (morph.zig:2-3, 257).

struct MorphSmallArgs {
    DImg src, dst;
    uint32_t mask0, mask1;
};

template <int KW, int CN, int TH>
struct MorphShape {
    static constexpr int HL = ((KW / 2) * CN + 3) / 4 * 4;   // left halo bytes (>= anchor * CN)
    static constexpr int HW = ((KW - 1) * CN + 3) / 4 + HL / 4;
    static constexpr int PW = 64 + HW;
    static constexpr int TR = 4 * TH + KW - 1;
    static constexpr int NWT = (HL + 4 + (KW - 1) * CN + 3) / 4 + 1;
    static constexpr int LDS_WORDS = TR * PW;
};

template <int KW, int CN, int TH, bool MAX>
__device__ inline void morph_small_tile(const MorphSmallArgs &p, int bx, int by, uint32_t *lds) {
    using S = MorphShape<KW, CN, TH>;
    constexpr int HL = S::HL, PW = S::PW, TR = S::TR, NWT = S::NWT, A = KW / 2;
    uint32_t(*tile)[PW] = reinterpret_cast<uint32_t(*)[PW]>(lds);
    const DImg &s = p.src, &d = p.dst;
    const int H = s.h, n = s.w * CN;
    const int x0 = bx * 256, y0 = by * 4 * TH;
    const uint32_t id = MAX ? 0u : 0xffffffffu;
    const bool aligned = ((s.stride | (uintptr_t)s.p) & 3) == 0;
    for (int idx = threadIdx.x; idx < TR * PW; idx += kTileThreads) {
        const int r = idx / PW, k = idx - r * PW;
        const int sy = y0 - A + r;
        const int c = x0 - HL + 4 * k;
        uint32_t v = id;
        if (sy >= 0 && sy < H) {
            const uint8_t *row = s.p + (long)sy * s.stride;
            if (aligned && c >= 0 && c + 4 <= n) {
                v = *reinterpret_cast<const uint32_t *>(row + c);
            } else {
                v = 0;
                for (int b = 0; b < 4; b++) {
                    const int cc = c + b;
                    const uint32_t byte = (cc >= 0 && cc < n) ? row[cc] : (MAX ? 0u : 255u);
                    v |= byte << (8 * b);
                }
            }
        }
        tile[r][k] = v;
    }
    __syncthreads();
    const int tx = tile_tx();
    const int x = x0 + tx * 4;
    const int yt = tile_ty() * TH;
    if (x >= n || y0 + yt >= H) return;
    uint32_t acc[TH][4];
#pragma unroll
    for (int t = 0; t < TH; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[t][j] = MAX ? 0u : 255u;
#pragma unroll
    for (int r = 0; r < TH + KW - 1; r++) {
        uint32_t w[NWT];
#pragma unroll
        for (int k = 0; k < NWT; k++) w[k] = (tx + k < PW) ? tile[yt + r][tx + k] : 0u;
#pragma unroll
        for (int t = 0; t < TH; t++) {
            const int ky = r - t;
            if (ky < 0 || ky >= KW) continue;
            const uint32_t bits = ky < 4 ? (p.mask0 >> (8 * ky)) & 255u : (p.mask1 >> (8 * (ky - 4))) & 255u;
#pragma unroll
            for (int i = 0; i < KW; i++) {
                if (!((bits >> i) & 1u)) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int o = HL + j + (i - A) * CN;
                    const uint32_t v = (w[o >> 2] >> ((o & 3) * 8)) & 255u;
                    acc[t][j] = MAX ? max(acc[t][j], v) : min(acc[t][j], v);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;
        uint8_t *out = d.p + (long)y * d.stride + x;
        if (x + 4 <= n && (((uintptr_t)out | (uintptr_t)d.stride) & 3) == 0) {
            *reinterpret_cast<uint32_t *>(out) = acc[t][0] | (acc[t][1] << 8) | (acc[t][2] << 16) | (acc[t][3] << 24);
        } else {
            for (int j = 0; j < 4 && x + j < n; j++) out[j] = (uint8_t)acc[t][j];
        }
    }
}

// -------------------------------------------------------------- warps
//
// The reference warps are float kernels (geometry.zig:410-533 warpAffine:
// f32 inverse, fma coordinates, fma blend; :615-672 warpPerspective: f64
// coordinates, f32 weights, no fma), reproduced operation for operation.
// Tiles are 64*PX pixels x 4 rows.

__device__ inline long to_idx(float f) { return (long)min_num(max_num(f, -1.0e9f), 1.0e9f); }
__device__ inline double clampd(double v, double lo, double hi) { return max_num(lo, min_num(v, hi)); }

template <typename T>
__device__ inline float ld_f(const DImg &s, long y, long i) {
    return (float)row_ptr<const T>(s, y)[i];
}

// Round half to even and clamp to u8 for a blend of u8 samples: such a value
// lies in [0, 255] (up to rounding), where the 1.5*2^23 trick in f32 equals
// sat.cast(u8, v) (geometry.zig:563-566 uses the same trick).
__device__ inline uint32_t round_u8(float v) {
    float r = rint_magic(v);
    r = min_num(max_num(r, 0.0f), 255.0f);
    return (uint32_t)r;
}

// Bilinear tap coordinates of one pixel: border-mapped (-1 = constant
// border tap); `out` = constant border and the footprint entirely outside.
struct Taps {
    int x0, x1, y0, y1;
    bool out;
};

__device__ inline Taps map_taps(long ix, long iy, int W, int H, int border) {
    Taps t;
    t.out = border == BORDER_CONSTANT && (ix < -1 || iy < -1 || ix >= W || iy >= H);
    if (ix >= 0 && ix + 1 < W && iy >= 0 && iy + 1 < H) {
        t.x0 = (int)ix, t.x1 = (int)ix + 1, t.y0 = (int)iy, t.y1 = (int)iy + 1;
    } else {
        t.x0 = border_map((int)ix, W, border), t.x1 = border_map((int)ix + 1, W, border);
        t.y0 = border_map((int)iy, H, border), t.y1 = border_map((int)iy + 1, H, border);
    }
    return t;
}

struct Quad {
    float p00, p01, p10, p11;
};

// Four u8 taps of a one-channel image, loaded from clamped addresses so the
// loads of several pixels are issued together, then masked to the border value.
__device__ inline Quad load_quad_u8(const DImg &s, const Taps &t, float cv) {
    const uint8_t *r0 = s.p + (long)max(t.y0, 0) * s.stride, *r1 = s.p + (long)max(t.y1, 0) * s.stride;
    const int xa = max(t.x0, 0), xb = max(t.x1, 0);
    uint32_t a = r0[xa], b = r0[xb], c = r1[xa], d = r1[xb];
    Quad q;
    q.p00 = (t.x0 >= 0 && t.y0 >= 0) ? (float)a : cv;
    q.p01 = (t.x1 >= 0 && t.y0 >= 0) ? (float)b : cv;
    q.p10 = (t.x0 >= 0 && t.y1 >= 0) ? (float)c : cv;
    q.p11 = (t.x1 >= 0 && t.y1 >= 0) ? (float)d : cv;
    return q;
}

// Four consecutive bilinear pixels of a one-channel u8 warpAffine.
__device__ inline uint32_t affine4_u8(const AffineParams &p, int x, float bx, float by) {
    const int W = p.s.w, H = p.s.h;
    Taps t[4];
    float ax[4], ay[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float xf = (float)(x + j);
        const float sx = fmaf(p.m[0], xf, bx), sy = fmaf(p.m[3], xf, by);
        const float fx = floorf(sx), fy = floorf(sy);
        ax[j] = sx - fx;
        ay[j] = sy - fy;
        t[j] = map_taps(to_idx(fx), to_idx(fy), W, H, p.border);
    }
    Quad q[4];
#pragma unroll
    for (int j = 0; j < 4; j++) q[j] = load_quad_u8(p.s, t[j], p.cv);
    uint32_t out = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float v0 = fmaf(ax[j], q[j].p01 - q[j].p00, q[j].p00);
        const float v1 = fmaf(ax[j], q[j].p11 - q[j].p10, q[j].p10);
        uint32_t v = round_u8(fmaf(ay[j], v1 - v0, v0));
        if (t[j].out) v = (uint32_t)p.cvd;
        out |= v << (8 * j);
    }
    return out;
}

// One output pixel of warpAffine, any depth and channel count.
template <typename T>
__device__ inline void affine_pixel(const AffineParams &p, int x, float bx, float by, T *o) {
    const DImg &s = p.s;
    const int cn = s.cn;
    const long W = s.w, H = s.h;
    float xf = (float)x;
    float sx = fmaf(p.m[0], xf, bx);
    float sy = fmaf(p.m[3], xf, by);
    if (p.interp == 0) {
        // geometry.zig:496-502 nearestPixel
        long ix = to_idx(rint_magic(min_num(max_num(sx, -4.0e6f), 4.0e6f)));
        long iy = to_idx(rint_magic(min_num(max_num(sy, -4.0e6f), 4.0e6f)));
        long mx = border_map((int)ix, (int)W, p.border), my = border_map((int)iy, (int)H, p.border);
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
    Taps t = map_taps(to_idx(fx), to_idx(fy), (int)W, (int)H, p.border);
    if (t.out) {
        for (int c = 0; c < cn; c++) o[c] = (T)p.cvd;
        return;
    }
    for (int c = 0; c < cn; c++) {
        float p00 = (t.x0 >= 0 && t.y0 >= 0) ? ld_f<T>(s, t.y0, (long)t.x0 * cn + c) : p.cv;
        float p01 = (t.x1 >= 0 && t.y0 >= 0) ? ld_f<T>(s, t.y0, (long)t.x1 * cn + c) : p.cv;
        float p10 = (t.x0 >= 0 && t.y1 >= 0) ? ld_f<T>(s, t.y1, (long)t.x0 * cn + c) : p.cv;
        float p11 = (t.x1 >= 0 && t.y1 >= 0) ? ld_f<T>(s, t.y1, (long)t.x1 * cn + c) : p.cv;
        float v0 = fmaf(ax, p01 - p00, p00);
        float v1 = fmaf(ax, p11 - p10, p10);
        o[c] = sat_cast<T>((double)fmaf(ay, v1 - v0, v0));
    }
}

// One-channel u8 bilinear warpAffine tile: 256 x 4 pixels, 4 per thread; a
// partial group at the row end computes all four, stores those in the row.
__device__ inline void affine_lin4_tile(const AffineParams &p, int bx, int by) {
    const DImg &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || x >= d.w) return;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2];
    const float byv = p.m[4] * yf + p.m[5];
    const uint32_t v = affine4_u8(p, x, bxv, byv);
    uint8_t *o = row_ptr<uint8_t>(d, y) + x;
    if (x + 4 <= d.w && ((uintptr_t)o & 3) == 0) {
        *reinterpret_cast<uint32_t *>(o) = v;
    } else {
        for (int j = 0; j < 4 && x + j < d.w; j++) o[j] = (uint8_t)(v >> (8 * j));
    }
}

// Any depth / channel count: 64 x 4 pixels per tile, one per thread.
template <typename T>
__device__ inline void affine_px_tile(const AffineParams &p, int bx, int by) {
    const DImg &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x = bx * 64 + tile_tx();
    if (y >= d.h || x >= d.w) return;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2];
    const float byv = p.m[4] * yf + p.m[5];
    affine_pixel<T>(p, x, bxv, byv, row_ptr<T>(d, y) + (long)x * d.cn);
}

// ------------------------------------------------------------- resize
//
// Tiles of 256 output samples x 4 rows (4 samples per thread; the exact
// half on one channel takes 16 samples per thread, 1024 x 4 per tile).

// resize_u8.zig:381-447: (a + b + c + d + 2) >> 2 per channel.
__device__ inline void half_tile(const DImg &s, const DImg &d, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const uint8_t *a = row_ptr<const uint8_t>(s, 2 * y), *b = row_ptr<const uint8_t>(s, 2 * y + 1);
    uint8_t *o = row_ptr<uint8_t>(d, y);
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        const int i = 2 * x * cn + ch;
        const uint32_t v = (uint32_t)a[i] + a[i + cn] + b[i] + b[i + cn];
        o[e] = (uint8_t)((v + 2) >> 2);
    }
}

__device__ inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

__device__ inline void half_c1_tile(const DImg &s, const DImg &d, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int x = (bx * 64 + tile_tx()) * 16;
    if (y >= d.h || x >= d.w) return;
    const uint8_t *a = row_ptr<const uint8_t>(s, 2 * y) + 2 * x, *b = row_ptr<const uint8_t>(s, 2 * y + 1) + 2 * x;
    uint8_t *o = row_ptr<uint8_t>(d, y) + x;
    if (x + 16 <= d.w && aligned16(a) && aligned16(b) && aligned16(o)) {
        uint4 a0 = reinterpret_cast<const uint4 *>(a)[0], a1 = reinterpret_cast<const uint4 *>(a)[1];
        uint4 b0 = reinterpret_cast<const uint4 *>(b)[0], b1 = reinterpret_cast<const uint4 *>(b)[1];
        uint32_t aw[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        uint32_t bw[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        uint32_t out[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int px = q * 4 + k;  // output pixel px reads input bytes 2px, 2px+1
                const uint32_t wa = aw[px / 2], wb = bw[px / 2];
                const int sh = (px & 1) * 16;
                const uint32_t sum = ((wa >> sh) & 255) + ((wa >> (sh + 8)) & 255) + ((wb >> sh) & 255) + ((wb >> (sh + 8)) & 255);
                v |= ((sum + 2) >> 2) << (8 * k);
            }
            out[q] = v;
        }
        *reinterpret_cast<uint4 *>(o) = make_uint4(out[0], out[1], out[2], out[3]);
        return;
    }
    const int cnt = min(16, d.w - x);
    for (int i = 0; i < cnt; i++) {
        const uint32_t v = (uint32_t)a[2 * i] + a[2 * i + 1] + b[2 * i] + b[2 * i + 1];
        o[i] = (uint8_t)((v + 2) >> 2);
    }
}

// resize_u8.zig:651-711: H = (s[o0]*w0 + s[o1]*w1) >> 4 as i16 per source
// row, out = sat_u8((((H0*b0) >> 16) + ((H1*b1) >> 16) + 2) >> 2).
__device__ inline void linear_tile(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const LinTap ty = yt[y];
    const uint8_t *r0 = row_ptr<const uint8_t>(s, ty.o0), *r1 = row_ptr<const uint8_t>(s, ty.o1);
    uint8_t *o = row_ptr<uint8_t>(d, y);
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        const LinTap tx = xt[x];
        const int32_t h0 = (int16_t)(((int32_t)r0[tx.o0 + ch] * tx.w0 + (int32_t)r0[tx.o1 + ch] * tx.w1) >> 4);
        const int32_t h1 = (int16_t)(((int32_t)r1[tx.o0 + ch] * tx.w0 + (int32_t)r1[tx.o1 + ch] * tx.w1) >> 4);
        const int32_t t = ((h0 * ty.w0) >> 16) + ((h1 * ty.w1) >> 16);
        const int32_t v = (t + 2) >> 2;
        o[e] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
}

// resize.zig:527-557 (areaGeneric, bit-exact with resize_u8.zig:1058-1208):
// per source row of the cell, the f32 sum of s*alpha over the cell's columns
// in table order, then the sum of those times beta, no FMA, then sat.cast.
__device__ inline void area_frac_tile(const DImg &s, const DImg &d, const AreaEntry *xt, const int32_t *xs,
                                      const AreaEntry *yt, const int32_t *ys, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    uint8_t *o = row_ptr<uint8_t>(d, y);
    const int ya = ys[y], yb = ys[y + 1];
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        const int xa = xs[x], xb = xs[x + 1];
        float sum = 0.0f;
        for (int k = ya; k < yb; k++) {
            const AreaEntry ye = yt[k];
            const uint8_t *r = row_ptr<const uint8_t>(s, ye.si);
            float h = 0.0f;
            for (int q = xa; q < xb; q++) h = h + (float)r[(long)xt[q].si * cn + ch] * xt[q].alpha;
            sum = sum + h * ye.alpha;
        }
        o[e] = sat_cast<uint8_t>((double)sum);
    }
}

// resize_u8.zig:1213-1277: integer block sum times the f32 reciprocal, satF32.
__device__ inline void area_int_tile(const DImg &s, const DImg &d, int ix, int iy, float scale, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    uint8_t *o = row_ptr<uint8_t>(d, y);
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        uint32_t sum = 0;
        for (int dy = 0; dy < iy; dy++) {
            const uint8_t *r = row_ptr<const uint8_t>(s, (long)y * iy + dy);
            for (int dx = 0; dx < ix; dx++) sum += r[(long)(x * ix + dx) * cn + ch];
        }
        o[e] = sat_f32<uint8_t>((float)sum * scale);
    }
}

}  // namespace visar
