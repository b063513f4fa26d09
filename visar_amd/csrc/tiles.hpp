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

// Global-memory views. Image pointers that reach a kernel through memory
// (the batch's job table) are generic to the compiler, which would then
// emit flat loads and stores; casting them to the global address space
// gets global ones.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __attribute__((always_inline)) inline gptr<T> gp(T *p) {
    return (gptr<T>)p;
}
template <typename T>
__device__ __attribute__((always_inline)) inline gptr<T> grow(const DImg &im, long y) {
    return (gptr<T>)(im.p + y * im.stride);
}

__device__ __attribute__((always_inline)) inline int tile_tx() { return threadIdx.x & 63; }
__device__ __attribute__((always_inline)) inline int tile_ty() { return threadIdx.x >> 6; }

// ------------------------------------------------------- row staging
//
// Tiles stage the source rows they need in LDS as whole, aligned dwords, so
// every global load is a coalesced dword load whatever the row's stride or
// base alignment. Staged dword k of staged row r holds the bytes at row
// offsets c0 - a + 4k .. c0 - a + 4k + 3, where a = (row address) & 3; a is
// kept per row (the last word of the row block), and the consumer recovers
// the logical dword m (bytes c0 + 4m ..) as alignbyte(staged[m + 1],
// staged[m], a). All loads of a thread are issued before any is used; dwords
// that reach past either end of the row are then rebuilt byte by byte with
// the tile's edge rule (border mapping for filters, the identity for
// morphology).

template <int TR, int PW2>
struct Staged {
    uint32_t (*tile)[PW2];
    const uint32_t *shift;
};

// row_of(r): source row of staged row r, or -1 for a row of `fill` words.
// edge_pos(c): where byte c of a row comes from (the identity inside
// [0, n), a border-mapped index outside, or -1 for the byte `fill_byte`).
template <int TR, int PW2, typename RowOf, typename EdgePos>
__device__ __attribute__((always_inline)) inline Staged<TR, PW2> stage_rows(const DImg &s, int n, int c0, uint32_t fill, uint32_t fill_byte,
                                             uint32_t *lds, RowOf row_of, EdgePos edge_pos) {
    constexpr int TOT = TR * PW2;
    constexpr int NI = (TOT + kTileThreads - 1) / kTileThreads;
    uint32_t(*tile)[PW2] = reinterpret_cast<uint32_t(*)[PW2]>(lds);
    uint32_t *shift = lds + TOT;
    uint32_t v[NI];
    // Whether any staged dword crosses a row end (the same for the whole
    // workgroup). Inner tiles only issue aligned dword loads; edge tiles load
    // every byte from its border-mapped position. Either way each thread
    // issues all its loads before it waits for any, and no load sits in a
    // branch of its own (the compiler would wait inside such a branch).
    const bool edge_tile = c0 - 3 < 0 || c0 + 4 * PW2 > n;
    if (!edge_tile) {
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int idx = threadIdx.x + i * kTileThreads;
            const int r = idx / PW2, k = idx - r * PW2;
            const int sy = idx < TOT ? row_of(r) : -1;
            const gptr<const uint8_t> row = grow<const uint8_t>(s, max(sy, 0));
            const int c = c0 - (int)((uintptr_t)row & 3) + 4 * k;
            v[i] = fill;
            if (sy >= 0) v[i] = *(gptr<const uint32_t>)(row + c);
        }
    } else {
        uint32_t eb[NI][4];
        int q[NI][4];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int idx = threadIdx.x + i * kTileThreads;
            const int r = idx / PW2, k = idx - r * PW2;
            const int sy = idx < TOT ? row_of(r) : -1;
            const gptr<const uint8_t> row = grow<const uint8_t>(s, max(sy, 0));
            const int c = c0 - (int)((uintptr_t)row & 3) + 4 * k;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                q[i][j] = sy >= 0 ? edge_pos(c + j) : -2;
                eb[i][j] = row[max(q[i][j], 0)];
            }
        }
#pragma unroll
        for (int i = 0; i < NI; i++) {
            uint32_t w = 0;
#pragma unroll
            for (int j = 0; j < 4; j++)
                w |= (q[i][j] >= 0 ? eb[i][j] : (q[i][j] == -1 ? fill_byte : (fill >> (8 * j)) & 255u)) << (8 * j);
            v[i] = w;
        }
    }
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int idx = threadIdx.x + i * kTileThreads;
        if (idx < TOT) {
            const int r = idx / PW2, k = idx - r * PW2;
            tile[r][k] = v[i];
            if (k == 0) {
                const int sy = row_of(r);
                shift[r] = sy >= 0 ? (uint32_t)((uintptr_t)(s.p + (long)sy * s.stride) & 3) : 0u;
            }
        }
    }
    __syncthreads();
    return Staged<TR, PW2>{tile, shift};
}

// Writes lane tx's four bytes v (row offsets x0 + 4tx ..) of a 256-byte tile
// row ending at `end` with whole aligned dwords where they lie inside the
// tile, so every byte of [x0, end) is written exactly once: lane tx writes
// the aligned word that starts b bytes before its own bytes (its
// predecessor's last b bytes and its own first 4 - b), bytes of words that
// cross the tile's ends go one at a time, and lane 63 writes its last b
// bytes. Every lane of the wave must call this (it shuffles).
__device__ __attribute__((always_inline)) inline void store_row_u8(uint8_t *row_, int x0, int end, uint32_t v) {
    const gptr<uint8_t> row = gp(row_);
    const int tx = tile_tx();
    const int b = (int)((uintptr_t)(row + x0) & 3);
    const uint32_t prev = __shfl_up(v, 1);
    const int x = x0 + 4 * tx;
    const int ws = x - b;
    const uint32_t word = b ? __builtin_amdgcn_alignbyte(v, prev, (uint32_t)(4 - b)) : v;
    if (ws >= x0 && ws + 4 <= end) {
        *(gptr<uint32_t>)(row + ws) = word;
    } else {
        for (int q = 0; q < 4; q++) {
            const int p = ws + q;
            if (p >= x0 && p < end) row[p] = (uint8_t)(word >> (8 * q));
        }
    }
    if (b && tx == 63) {
        for (int q = 4 - b; q < 4; q++) {
            const int p = x + q;
            if (p < end) row[p] = (uint8_t)(v >> (8 * q));
        }
    }
}

// ------------------------------------------------------------ stencils
//
// u8 sources, square 3/5/7 kernels (GaussianBlur u8 and Sobel u8 -> i16).
// Exact integer sums make the pass order free (reference filter.zig:163-167).
// The tile is 256 consecutive samples of the flat row x 4*TH output rows.
// The workgroup stages the rows it needs plus the column halo in LDS with the
// border rules applied while loading (border.zig:22-39). Each thread then
// runs the horizontal taps of its 4 samples on each staged row and the
// vertical taps over those results, with full-rate 24-bit multiplies (taps
// and partial sums stay below 2^23 in magnitude).

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
    static constexpr int PW2 = 64 + 2 * HW + 1;  // staged dwords per row (one more for the alignment shift)
    static constexpr int TR = 4 * TH + 2 * R;    // staged rows
    static constexpr int LDS_WORDS = TR * PW2 + TR;
};

// Index of byte c of a source row under column border mapping (c may lie
// outside [0, n)): pixel floor(c / CN), channel c - CN * pixel; -1 for a
// constant border.
template <int CN>
__device__ __attribute__((always_inline)) inline int mapped_pos(int c, int W, int border) {
    const int px = (c + 64 * CN) / CN - 64;
    const int ch = c - px * CN;
    const int mx = border_map_halo(px, W, border);
    return mx < 0 ? -1 : mx * CN + ch;
}

template <typename D>
__device__ __attribute__((always_inline)) inline int stencil_finish(int a, int fin, float scale) {
    if (fin == FIN_FIX8) {
        int r = (a + (1 << 15)) >> 16;
        return r > 255 ? 255 : r;
    }
    if (fin == FIN_SCALED) return (int)sat_v<D>((float)a * scale);
    int lo = (int)Lim<D>::lo, hi = (int)Lim<D>::hi;
    return a < lo ? lo : a > hi ? hi : a;
}

template <typename D>
__device__ __attribute__((always_inline)) inline void stencil_store(D *row, int x0, int x, int n, const int (&v)[4]);
template <>
__device__ __attribute__((always_inline)) inline void stencil_store<uint8_t>(uint8_t *row, int x0, int x, int n, const int (&v)[4]) {
    const uint32_t w = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
    store_row_u8(row, x0, min(x0 + 256, n), w);
}
template <>
__device__ __attribute__((always_inline)) inline void stencil_store<int16_t>(int16_t *row, int x0, int x, int n, const int (&v)[4]) {
    if (x >= n) return;
    const gptr<int16_t> out = gp(row + x);
    if (x + 4 <= n && ((uintptr_t)out & 7) == 0) {
        uint32_t a = (uint32_t)(uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16);
        uint32_t b = (uint32_t)(uint16_t)v[2] | ((uint32_t)(uint16_t)v[3] << 16);
        *(gptr<u32x2>)out = u32x2{a, b};
    } else {
        for (int j = 0; j < 4 && x + j < n; j++) out[j] = (int16_t)v[j];
    }
}

template <int K, int CN, int TH, typename D>
__device__ __attribute__((always_inline)) inline void stencil_tile(const StencilArgs &p, int bx, int by, uint32_t *lds) {
    using S = StencilShape<K, CN, TH>;
    constexpr int R = S::R, HW = S::HW, NW = S::NW, PW2 = S::PW2, TR = S::TR;
    const int W = p.src.w, H = p.src.h;
    const int n = W * CN;
    const int x0 = bx * 256;
    const int y0 = by * 4 * TH;
    const int border = p.border;
    const auto st = stage_rows<TR, PW2>(
        p.src, n, x0 - 4 * HW, 0u, 0u, lds, [&](int r) { return border_map_halo(y0 - R + r, H, border); },
        [&](int c) { return mapped_pos<CN>(c, W, border); });
    const int tx = tile_tx();
    const int x = x0 + tx * 4;
    const int yt = tile_ty() * TH;
    if (y0 + yt >= H) return;  // the same for the whole wave
    int hz[TH + 2 * R][4];
#pragma unroll
    for (int r = 0; r < TH + 2 * R; r++) {
        const uint32_t a = st.shift[yt + r];
        uint32_t d[NW + 1];
#pragma unroll
        for (int k = 0; k <= NW; k++) d[k] = st.tile[yt + r][tx + k];
        uint32_t w[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], a);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int acc = 0;
#pragma unroll
            for (int i = 0; i < K; i++) {
                const int o = 4 * HW + j + (i - R) * CN;
                acc += __mul24(p.kx[i], (int)((w[o >> 2] >> ((o & 3) * 8)) & 255u));
            }
            hz[r][j] = acc;
        }
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;  // uniform across the wave
        int o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int acc = 0;
#pragma unroll
            for (int k = 0; k < K; k++) acc += __mul24(p.ky[k], hz[t + k][j]);
            o[j] = stencil_finish<D>(acc, p.fin, p.scale);
        }
        stencil_store<D>(reinterpret_cast<D *>(p.dst.p + (long)y * p.dst.stride), x0, x, n, o);
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
    static constexpr int NWT = (HL + 4 + (KW - 1) * CN + 3) / 4 + 1;  // dwords a thread reads per row
    static constexpr int PW2 = 64 + NWT + 1;
    static constexpr int TR = 4 * TH + KW - 1;
    static constexpr int LDS_WORDS = TR * PW2 + TR;
};

template <int KW, int CN, int TH, bool MAX>
__device__ __attribute__((always_inline)) inline void morph_small_tile(const MorphSmallArgs &p, int bx, int by, uint32_t *lds) {
    using S = MorphShape<KW, CN, TH>;
    constexpr int HL = S::HL, PW2 = S::PW2, TR = S::TR, NWT = S::NWT, A = KW / 2;
    const DImg &s = p.src, &d = p.dst;
    const int H = s.h, n = s.w * CN;
    const int x0 = bx * 256, y0 = by * 4 * TH;
    const uint32_t id = MAX ? 0u : 0xffffffffu;
    const auto st = stage_rows<TR, PW2>(
        s, n, x0 - HL, id, id & 255u, lds,
        [&](int r) {
            const int sy = y0 - A + r;
            return (sy >= 0 && sy < H) ? sy : -1;
        },
        [&](int c) { return (c >= 0 && c < n) ? c : -1; });
    const int tx = tile_tx();
    const int yt = tile_ty() * TH;
    if (y0 + yt >= H) return;  // the same for the whole wave
    uint32_t acc[TH][4];
#pragma unroll
    for (int t = 0; t < TH; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[t][j] = MAX ? 0u : 255u;
#pragma unroll
    for (int r = 0; r < TH + KW - 1; r++) {
        const uint32_t a = st.shift[yt + r];
        uint32_t dd[NWT + 1];
#pragma unroll
        for (int k = 0; k <= NWT; k++) dd[k] = st.tile[yt + r][tx + k];
        uint32_t w[NWT];
#pragma unroll
        for (int k = 0; k < NWT; k++) w[k] = __builtin_amdgcn_alignbyte(dd[k + 1], dd[k], a);
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
        if (y >= H) break;  // uniform across the wave
        const uint32_t w = acc[t][0] | (acc[t][1] << 8) | (acc[t][2] << 16) | (acc[t][3] << 24);
        store_row_u8(d.p + (long)y * d.stride, x0, min(x0 + 256, n), w);
    }
}

// -------------------------------------------------------------- warps
//
// The reference warps are float kernels (geometry.zig:410-533 warpAffine:
// f32 inverse, fma coordinates, fma blend; :615-672 warpPerspective: f64
// coordinates, f32 weights, no fma), reproduced operation for operation.
// Tiles are 64*PX pixels x 4 rows.

__device__ __attribute__((always_inline)) inline long to_idx(float f) { return (long)min_num(max_num(f, -1.0e9f), 1.0e9f); }
__device__ __attribute__((always_inline)) inline double clampd(double v, double lo, double hi) { return max_num(lo, min_num(v, hi)); }

template <typename T>
__device__ __attribute__((always_inline)) inline float ld_f(const DImg &s, long y, long i) {
    return (float)grow<const T>(s, y)[i];
}

// Round half to even and clamp to u8 for a blend of u8 samples: such a value
// lies in [0, 255] (up to rounding), where the 1.5*2^23 trick in f32 equals
// sat.cast(u8, v) (geometry.zig:563-566 uses the same trick).
__device__ __attribute__((always_inline)) inline uint32_t round_u8(float v) {
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

__device__ __attribute__((always_inline)) inline Taps map_taps(long ix, long iy, int W, int H, int border) {
    Taps t;
    t.out = border == BORDER_CONSTANT && (ix < -1 || iy < -1 || ix >= W || iy >= H);
    if (ix >= 0 && ix + 1 < W && iy >= 0 && iy + 1 < H) {
        t.x0 = (int)ix, t.x1 = (int)ix + 1, t.y0 = (int)iy, t.y1 = (int)iy + 1;
    } else {
        t.x0 = border_map_inl((int)ix, W, border), t.x1 = border_map_inl((int)ix + 1, W, border);
        t.y0 = border_map_inl((int)iy, H, border), t.y1 = border_map_inl((int)iy + 1, H, border);
    }
    return t;
}

struct Quad {
    float p00, p01, p10, p11;
};

// Four u8 taps of a one-channel image, loaded from clamped addresses so the
// loads of several pixels are issued together, then masked to the border value.
__device__ __attribute__((always_inline)) inline Quad load_quad_u8(const DImg &s, const Taps &t, float cv) {
    const gptr<const uint8_t> r0 = grow<const uint8_t>(s, max(t.y0, 0)), r1 = grow<const uint8_t>(s, max(t.y1, 0));
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
__device__ __attribute__((always_inline)) inline uint32_t affine4_u8(const AffineParams &p, int x, float bx, float by) {
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
__device__ __attribute__((always_inline)) inline void affine_pixel(const AffineParams &p, int x, float bx, float by, gptr<T> o) {
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
        long mx = border_map_inl((int)ix, (int)W, p.border), my = border_map_inl((int)iy, (int)H, p.border);
        if (mx >= 0 && my >= 0) {
            const gptr<const T> sp = grow<const T>(s, my) + mx * cn;
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
__device__ __attribute__((always_inline)) inline void affine_lin4_tile(const AffineParams &p, int bx, int by) {
    const DImg &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || x >= d.w) return;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2];
    const float byv = p.m[4] * yf + p.m[5];
    const uint32_t v = affine4_u8(p, x, bxv, byv);
    const gptr<uint8_t> o = grow<uint8_t>(d, y) + x;
    if (x + 4 <= d.w && ((uintptr_t)o & 3) == 0) {
        *(gptr<uint32_t>)o = v;
    } else {
        for (int j = 0; j < 4 && x + j < d.w; j++) o[j] = (uint8_t)(v >> (8 * j));
    }
}

// Any depth / channel count: 64 x 4 pixels per tile, one per thread.
template <typename T>
__device__ __attribute__((always_inline)) inline void affine_px_tile(const AffineParams &p, int bx, int by) {
    const DImg &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x = bx * 64 + tile_tx();
    if (y >= d.h || x >= d.w) return;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2];
    const float byv = p.m[4] * yf + p.m[5];
    affine_pixel<T>(p, x, bxv, byv, grow<T>(d, y) + (long)x * d.cn);
}

// ------------------------------------------------------------- resize
//
// Tiles of 256 output samples x 4 rows (4 samples per thread; the exact
// half on one channel takes 16 samples per thread, 1024 x 4 per tile).

// resize_u8.zig:381-447: (a + b + c + d + 2) >> 2 per channel.
__device__ __attribute__((always_inline)) inline void half_tile(const DImg &s, const DImg &d, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const gptr<const uint8_t> a = grow<const uint8_t>(s, 2 * y), b = grow<const uint8_t>(s, 2 * y + 1);
    const gptr<uint8_t> o = grow<uint8_t>(d, y);
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        const int i = 2 * x * cn + ch;
        const uint32_t v = (uint32_t)a[i] + a[i + cn] + b[i] + b[i + cn];
        o[e] = (uint8_t)((v + 2) >> 2);
    }
}

__device__ __attribute__((always_inline)) inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
template <typename T>
__device__ __attribute__((always_inline)) inline bool aligned16(gptr<T> p) { return ((uintptr_t)p & 15) == 0; }

__device__ __attribute__((always_inline)) inline void half_c1_tile(const DImg &s, const DImg &d, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int x = (bx * 64 + tile_tx()) * 16;
    if (y >= d.h || x >= d.w) return;
    const gptr<const uint8_t> a = grow<const uint8_t>(s, 2 * y) + 2 * x, b = grow<const uint8_t>(s, 2 * y + 1) + 2 * x;
    const gptr<uint8_t> o = grow<uint8_t>(d, y) + x;
    if (x + 16 <= d.w && aligned16(a) && aligned16(b) && aligned16(o)) {
        const u32x4 a0 = ((gptr<const u32x4>)a)[0], a1 = ((gptr<const u32x4>)a)[1];
        const u32x4 b0 = ((gptr<const u32x4>)b)[0], b1 = ((gptr<const u32x4>)b)[1];
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
        *(gptr<u32x4>)o = u32x4{out[0], out[1], out[2], out[3]};
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
__device__ __attribute__((always_inline)) inline void linear_tile(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const LinTap ty = yt[y];
    const gptr<const uint8_t> r0 = grow<const uint8_t>(s, ty.o0), r1 = grow<const uint8_t>(s, ty.o1);
    const gptr<uint8_t> o = grow<uint8_t>(d, y);
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
__device__ __attribute__((always_inline)) inline void area_frac_tile(const DImg &s, const DImg &d, const AreaEntry *xt, const int32_t *xs,
                                      const AreaEntry *yt, const int32_t *ys, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const gptr<uint8_t> o = grow<uint8_t>(d, y);
    const int ya = ys[y], yb = ys[y + 1];
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        const int xa = xs[x], xb = xs[x + 1];
        float sum = 0.0f;
        for (int k = ya; k < yb; k++) {
            const AreaEntry ye = yt[k];
            const gptr<const uint8_t> r = grow<const uint8_t>(s, ye.si);
            float h = 0.0f;
            for (int q = xa; q < xb; q++) h = h + (float)r[(long)xt[q].si * cn + ch] * xt[q].alpha;
            sum = sum + h * ye.alpha;
        }
        o[e] = sat_cast<uint8_t>((double)sum);
    }
}

// resize_u8.zig:1213-1277: integer block sum times the f32 reciprocal, satF32.
__device__ __attribute__((always_inline)) inline void area_int_tile(const DImg &s, const DImg &d, int ix, int iy, float scale, int bx, int by) {
    const int y = by * 4 + tile_ty();
    const int cn = d.cn;
    const int n = d.w * cn;
    const int e0 = (bx * 64 + tile_tx()) * 4;
    if (y >= d.h || e0 >= n) return;
    const gptr<uint8_t> o = grow<uint8_t>(d, y);
    for (int e = e0; e < e0 + 4 && e < n; e++) {
        const int x = e / cn, ch = e - x * cn;
        uint32_t sum = 0;
        for (int dy = 0; dy < iy; dy++) {
            const gptr<const uint8_t> r = grow<const uint8_t>(s, (long)y * iy + dy);
            for (int dx = 0; dx < ix; dx++) sum += r[(long)(x * ix + dx) * cn + ch];
        }
        o[e] = sat_f32<uint8_t>((float)sum * scale);
    }
}

}  // namespace visar
