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

// Runs fn(integral_constant<int, N>) with N = n when n is odd and 3 <= n <=
// MaxN, else with N = 0, so that a window loop whose length is known at
// compile time can be unrolled in full: its reads then issue together and
// its weights or counters stay out of the scalar unit (a runtime-length
// loop spends about four scalar instructions per tap, and the one scalar
// unit of a CU then bounds the kernel). n is uniform: one scalar compare
// per candidate length, once per call site.
template <int N, int MaxN, typename Fn>
__device__ inline void with_odd_len_from(int n, Fn &fn) {
    if constexpr (N > MaxN) {
        fn(std::integral_constant<int, 0>{});
    } else {
        if (n == N) fn(std::integral_constant<int, N>{});
        else with_odd_len_from<N + 2, MaxN>(n, fn);
    }
}
template <int MaxN, typename Fn>
__device__ inline void with_odd_len(int n, Fn fn) {
    with_odd_len_from<3, MaxN>(n, fn);
}

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

// Tile (bx, by) of a plain call launched as a 1-D grid of tile_grid(ntiles)
// workgroups. Workgroup b runs on XCD b % 8. In XCD order it takes tile
// (b % 8) * ceil(ntiles / 8) + b / 8 of the row-major tile list, so each XCD
// works through one contiguous band of the image and the halo rows or
// source boxes that neighbouring tiles share stay in that XCD's L2; in
// row-major order it takes tile b. False for the padding workgroups.
struct TileOrder {
    int tiles_x, ntiles, xcd;
};
inline unsigned tile_grid(int ntiles) { return (unsigned)((ntiles + 7) / 8 * 8); }
// xcd_default: the kernel's default order. XCD bands for the warps (the
// perspective warp runs faster in them, and the affine box warp, whose
// neighbouring tiles' boxes overlap, reads a quarter of the bytes from
// memory at the same speed), the box filter and the pyramids (neighbouring
// tiles share halo rows); row-major for the others.
inline TileOrder tile_order(int tiles_x, int tiles_y, bool xcd_default = false) {
    const int k = tuning().tile_order;
    return TileOrder{tiles_x, tiles_x * tiles_y, k == 2 || (k == 0 && xcd_default)};
}
__device__ __attribute__((always_inline)) inline bool tile_of(const TileOrder &o, int &bx, int &by) {
    const int b = (int)blockIdx.x;
    const int t = o.xcd ? (b & 7) * ((o.ntiles + 7) >> 3) + (b >> 3) : b;
    if (t >= o.ntiles) return false;
    by = t / o.tiles_x;
    bx = t - by * o.tiles_x;
    return true;
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
// EW: dwords at each end of a staged row that may cross the row's ends and
// are still read by a thread with output in the row (left: the dwords
// before byte 0; right: from the first dword that reaches past byte n - 1). Dwords wholly inside
// the row are one aligned load each; the crossing ones (at most EW per end
// and row, one per thread) are four byte loads from their mapped positions,
// issued right after the dword loads so the workgroup waits once.
template <int TR, int PW2, int EW, typename RowOf, typename EdgePos>
__device__ __attribute__((always_inline)) inline Staged<TR, PW2> stage_rows(const DImg &s, int n, int c0, uint32_t fill,
                                                                            uint32_t fill_byte, uint32_t *lds,
                                                                            RowOf row_of, EdgePos edge_pos) {
    constexpr int TOT = TR * PW2;
    constexpr int NI = (TOT + kTileThreads - 1) / kTileThreads;
    constexpr int SLOTS = TR * 2 * EW;
    static_assert(SLOTS <= kTileThreads, "one edge slot per thread");
    uint32_t(*tile)[PW2] = reinterpret_cast<uint32_t(*)[PW2]>(lds);
    uint32_t *shift = lds + TOT;
    uint32_t v[NI];
    bool put[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int idx = threadIdx.x + i * kTileThreads;
        const int r = idx / PW2, k = idx - r * PW2;
        const int sy = idx < TOT ? row_of(r) : -1;
        const gptr<const uint8_t> row = grow<const uint8_t>(s, max(sy, 0));
        const int c = c0 - (int)((uintptr_t)row & 3) + 4 * k;
        const bool in = sy >= 0 && c >= 0 && c + 4 <= n;
        put[i] = idx < TOT && (in || sy < 0);
        v[i] = fill;
        if (in) v[i] = *(gptr<const uint32_t>)(row + c);
    }
    // Edge slot of this thread: row r, left dword j or right dword j.
    const int slot = threadIdx.x;
    int er = -1, ek = 0;
    uint32_t ev = 0;
    if (slot < SLOTS) {
        const int r = slot / (2 * EW), j = slot - r * 2 * EW;
        const int sy = row_of(r);
        if (sy >= 0) {
            const gptr<const uint8_t> row = grow<const uint8_t>(s, sy);
            const int a = (int)((uintptr_t)row & 3);
            int k = j;
            if (j >= EW) k = ((n - 4 - c0 + a) >> 2) + 1 + (j - EW);  // first dword past the row end, onwards
            const int c = c0 - a + 4 * k;
            const bool mine = k >= 0 && k < PW2 && (j < EW ? c < 0 : (c >= 0 && c + 4 > n));
            if (mine) {
                int q[4];
                uint32_t b[4];
#pragma unroll
                for (int t = 0; t < 4; t++) q[t] = edge_pos(c + t);
#pragma unroll
                for (int t = 0; t < 4; t++) b[t] = row[max(q[t], 0)];
#pragma unroll
                for (int t = 0; t < 4; t++) ev |= (q[t] >= 0 ? b[t] : fill_byte) << (8 * t);
                er = r;
                ek = k;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int idx = threadIdx.x + i * kTileThreads;
        if (put[i]) {
            const int r = idx / PW2, k = idx - r * PW2;
            tile[r][k] = v[i];
        }
    }
    if (er >= 0) tile[er][ek] = ev;
    if (threadIdx.x < TR) {
        const int sy = row_of(threadIdx.x);
        shift[threadIdx.x] = sy >= 0 ? (uint32_t)((uintptr_t)(s.p + (long)sy * s.stride) & 3) : 0u;
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
    const auto st = stage_rows<TR, PW2, HW + 2>(
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

// -------------------------------------------------- packed 16-bit stencils
//
// The stencils of the benchmark shapes in pairs of 16-bit lanes: GaussianBlur
// u8 7x7 with sigma 0 (below), 3x3 and 5x5 with sigma 0 (binomial taps 64 * [1 2 1] and
// 16 * [1 4 6 4 1], which reduce to [1 2 1] and [1 4 6 4 1]; the full sum is
// 4096 or 256 times the reduced one, so the reference's (sum + 2^15) >> 16 is
// (v + 8) >> 4 or (v + 128) >> 8 exactly, and every partial sum stays below
// 2^16), and Sobel 3x3 dx or dy to i16 with scale 1 (partial sums within
// +-1020, exact in wrapping 16-bit arithmetic). A thread's 4 samples are
// handled as two pairs, samples (0, 2) and (1, 3); a pair of source bytes
// (b[e], b[e + 2]) comes from the staged dwords with one v_perm.

enum SwarMode : int { SWAR_G3 = 0, SWAR_G5 = 1, SWAR_SOBEL_DX = 2, SWAR_SOBEL_DY = 3, SWAR_G7 = 4 };

// The packed mode whose taps a u8 stencil call has exactly (k taps per
// axis, destination depth ddepth), or -1 (host side, k_stencil.hip).
int swar_mode(const StencilArgs &a, int k, int ddepth);

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
typedef int16_t i16x2 __attribute__((ext_vector_type(2)));

template <int MODE>
struct SwarTaps;
template <>
struct SwarTaps<SWAR_G3> {
    static constexpr int K = 3;
    static constexpr int h[3] = {1, 2, 1}, v[3] = {1, 2, 1};
};
template <>
struct SwarTaps<SWAR_G5> {
    static constexpr int K = 5;
    static constexpr int h[5] = {1, 4, 6, 4, 1}, v[5] = {1, 4, 6, 4, 1};
};
// 7 x 7 with sigma 0: taps 4 * {2, 7, 14, 18, 14, 7, 2}, so the result is
// (R + 2048) >> 12 for the reduced sum R < 2^20. The horizontal sums stay
// below 2^14 in 16-bit lanes; the vertical pass adds the mirrored rows in
// 16 bits and takes the weighted sum in 32 bits with v_dot2_u32_u16.
template <>
struct SwarTaps<SWAR_G7> {
    static constexpr int K = 7;
    static constexpr int h[7] = {2, 7, 14, 18, 14, 7, 2}, v[7] = {2, 7, 14, 18, 14, 7, 2};
};
template <>
struct SwarTaps<SWAR_SOBEL_DX> {
    static constexpr int K = 3;
    static constexpr int h[3] = {-1, 0, 1}, v[3] = {1, 2, 1};
};
template <>
struct SwarTaps<SWAR_SOBEL_DY> {
    static constexpr int K = 3;
    static constexpr int h[3] = {1, 2, 1}, v[3] = {-1, 0, 1};
};

// (b[e], b[e + 2]) as two zero-extended 16-bit lanes; byte e is relative to
// the thread's own dword w[HW], so it lives at index e + 4 HW of w[].
template <int HW, int NW>
__device__ __attribute__((always_inline)) inline uint32_t byte_pair(const uint32_t (&w)[NW], int e) {
    const int q = e + 4 * HW, i = q >> 2, o = q & 3;
    if (o <= 1) return __builtin_amdgcn_perm(w[i], w[i], 0x0c000c00u | ((uint32_t)(o + 2) << 16) | (uint32_t)o);
    return __builtin_amdgcn_perm(w[i + 1], w[i], 0x0c000c00u | ((uint32_t)(o + 2) << 16) | (uint32_t)o);
}

template <int MODE>
__device__ __attribute__((always_inline)) inline uint32_t swar_mac(uint32_t acc, uint32_t x, int k) {
    if (k == 0) return acc;
    if constexpr (MODE == SWAR_SOBEL_DX || MODE == SWAR_SOBEL_DY) {
        const i16x2 a = __builtin_bit_cast(i16x2, acc), b = __builtin_bit_cast(i16x2, x);
        i16x2 r;
        if (k == 1) r = a + b;
        else if (k == -1) r = a - b;
        else r = a + b * (int16_t)k;
        return __builtin_bit_cast(uint32_t, r);
    } else {
        const u16x2 a = __builtin_bit_cast(u16x2, acc), b = __builtin_bit_cast(u16x2, x);
        const u16x2 r = k == 1 ? a + b : a + b * (uint16_t)k;
        return __builtin_bit_cast(uint32_t, r);
    }
}

template <int MODE, int CN, int TH, typename D>
__device__ __attribute__((always_inline)) inline void stencil_swar_tile(const StencilArgs &p, int bx, int by,
                                                                        uint32_t *lds) {
    using T = SwarTaps<MODE>;
    constexpr int K = T::K;
    using S = StencilShape<K, CN, TH>;
    constexpr int R = S::R, HW = S::HW, NW = S::NW, PW2 = S::PW2, TR = S::TR;
    const int W = p.src.w, H = p.src.h;
    const int n = W * CN;
    const int x0 = bx * 256;
    const int y0 = by * 4 * TH;
    const int border = p.border;
    const auto st = stage_rows<TR, PW2, HW + 2>(
        p.src, n, x0 - 4 * HW, 0u, 0u, lds, [&](int r) { return border_map_halo(y0 - R + r, H, border); },
        [&](int c) { return mapped_pos<CN>(c, W, border); });
    const int tx = tile_tx();
    const int x = x0 + tx * 4;
    const int yt = tile_ty() * TH;
    if (y0 + yt >= H) return;  // the same for the whole wave
    uint32_t he[TH + 2 * R], ho[TH + 2 * R];  // horizontal sums of samples (0, 2) and (1, 3)
#pragma unroll
    for (int r = 0; r < TH + 2 * R; r++) {
        const uint32_t a = st.shift[yt + r];
        uint32_t d[NW + 1];
#pragma unroll
        for (int k = 0; k <= NW; k++) d[k] = st.tile[yt + r][tx + k];
        uint32_t w[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], a);
        uint32_t se = 0, so = 0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            se = swar_mac<MODE>(se, byte_pair<HW>(w, (i - R) * CN), T::h[i]);
            so = swar_mac<MODE>(so, byte_pair<HW>(w, 1 + (i - R) * CN), T::h[i]);
        }
        he[r] = se;
        ho[r] = so;
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;  // uniform across the wave
        D *row = reinterpret_cast<D *>(p.dst.p + (long)y * p.dst.stride);
        if constexpr (MODE == SWAR_G7) {
            // per lane: 2 (r0 + r6) + 7 (r1 + r5) + 14 (r2 + r4) + 18 r3
            uint32_t f[4];  // samples 0, 2 (even pair) and 1, 3 (odd pair)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint32_t *hv = q == 0 ? he : ho;
                const u16x2 a = __builtin_bit_cast(u16x2, hv[t]) + __builtin_bit_cast(u16x2, hv[t + 6]);
                const u16x2 b = __builtin_bit_cast(u16x2, hv[t + 1]) + __builtin_bit_cast(u16x2, hv[t + 5]);
                const u16x2 c = __builtin_bit_cast(u16x2, hv[t + 2]) + __builtin_bit_cast(u16x2, hv[t + 4]);
                const uint32_t ua = __builtin_bit_cast(uint32_t, a), ub = __builtin_bit_cast(uint32_t, b),
                               uc = __builtin_bit_cast(uint32_t, c), um = hv[t + 3];
#pragma unroll
                for (int l = 0; l < 2; l++) {
                    const uint32_t sel = l == 0 ? 0x05040100u : 0x07060302u;
                    const u16x2 ab = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(ub, ua, sel));
                    const u16x2 cm = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(um, uc, sel));
                    uint32_t r = __builtin_amdgcn_udot2(ab, (u16x2){2, 7}, 2048u, false);
                    r = __builtin_amdgcn_udot2(cm, (u16x2){14, 18}, r, false);
                    f[q + 2 * l] = r >> 12;
                }
            }
            // f[0] = sample 0, f[1] = sample 1, f[2] = sample 2, f[3] = sample 3
            const uint32_t v4 = f[0] | (f[1] << 8) | (f[2] << 16) | (f[3] << 24);
            store_row_u8(reinterpret_cast<uint8_t *>(row), x0, min(x0 + 256, n), v4);
            continue;
        }
        uint32_t ve = 0, vo = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            ve = swar_mac<MODE>(ve, he[t + k], T::v[k]);
            vo = swar_mac<MODE>(vo, ho[t + k], T::v[k]);
        }
        if constexpr (MODE == SWAR_G3 || MODE == SWAR_G5) {
            constexpr uint16_t add = MODE == SWAR_G3 ? 8 : 128, sh = MODE == SWAR_G3 ? 4 : 8;
            const u16x2 fe = (__builtin_bit_cast(u16x2, ve) + add) >> sh;
            const u16x2 fo = (__builtin_bit_cast(u16x2, vo) + add) >> sh;
            // bytes: sample 0 = fe.lo, 1 = fo.lo, 2 = fe.hi, 3 = fo.hi
            const uint32_t v4 = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, fo), __builtin_bit_cast(uint32_t, fe),
                                                      0x06020400u);
            store_row_u8(reinterpret_cast<uint8_t *>(row), x0, min(x0 + 256, n), v4);
        } else {
            // i16 samples: dword 0 = (sample 0, sample 1), dword 1 = (2, 3)
            const uint32_t lo = __builtin_amdgcn_perm(vo, ve, 0x05040100u);
            const uint32_t hi = __builtin_amdgcn_perm(vo, ve, 0x07060302u);
            if (x >= n) continue;
            const gptr<int16_t> out = gp(reinterpret_cast<int16_t *>(row) + x);
            if (x + 4 <= n && ((uintptr_t)out & 7) == 0) {
                *(gptr<u32x2>)out = u32x2{lo, hi};
            } else {
                const int16_t vals[4] = {(int16_t)lo, (int16_t)(lo >> 16), (int16_t)hi, (int16_t)(hi >> 16)};
                for (int j = 0; j < 4 && x + j < n; j++) out[j] = vals[j];
            }
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
    const auto st = stage_rows<TR, PW2, (NWT - HL / 4 + 1 > HL / 4 + 1 ? NWT - HL / 4 + 1 : HL / 4 + 1)>(
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
    // v is a blend of u8 samples, never NaN, so IEEE min/max equal the
    // reference's NaN-aware ones here.
    const float r = fminf(fmaxf(rint_magic(v), 0.0f), 255.0f);
    return (uint32_t)r;
}

// round_u8(v) written into byte `byte` of acc, other bytes kept
// (v_cvt_pk_u8_f32: round to nearest even and saturate to [0, 255], equal
// to round_u8 for every finite v).
__device__ __attribute__((always_inline)) inline uint32_t pack_u8(float v, uint32_t byte, uint32_t acc) {
    return __builtin_amdgcn_cvt_pk_u8_f32(v, byte, acc);
}

// Bilinear tap coordinates of one pixel: border-mapped (-1 = constant
// border tap); `out` = constant border and the footprint entirely outside.
struct Taps {
    int x0, x1, y0, y1;
    bool out;
};

__device__ __attribute__((always_inline)) inline Taps map_taps(int ix, int iy, int W, int H, int border) {
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

// toIdx (geometry.zig:492-494) for the coordinates a warp produces, which
// are finite: IEEE min/max equal the reference's NaN-aware ones on them.
__device__ __attribute__((always_inline)) inline int to_i32(float f) { return (int)fminf(fmaxf(f, -1.0e9f), 1.0e9f); }

// Whether a bilinear footprint at (ix, iy) lies wholly inside the image.
__device__ __attribute__((always_inline)) inline bool inner_tap(int ix, int iy, int W, int H) {
    return (unsigned)ix < (unsigned)(W - 1) && (unsigned)iy < (unsigned)(H - 1);
}

// The taps of four one-channel u8 pixels. When every footprint of the wave
// lies inside the image (the common case), the taps are plain loads at
// 32-bit offsets from the image base; otherwise each goes through the
// border rules.
__device__ __attribute__((always_inline)) inline void quads_u8(const DImg &s, int border, float cv, const int (&ix)[4],
                                                               const int (&iy)[4], Quad (&q)[4], bool (&out)[4]) {
    const int W = s.w, H = s.h;
    bool inner = (unsigned long)s.h * (unsigned long)s.stride < (1ul << 32);
#pragma unroll
    for (int j = 0; j < 4; j++) inner = inner && inner_tap(ix[j], iy[j], W, H);
    if (border == BORDER_CONSTANT) {
        // Every pixel of the wave outside by the constant-border rule: no tap is read.
        bool none = true;
#pragma unroll
        for (int j = 0; j < 4; j++) none = none && (ix[j] < -1 || iy[j] < -1 || ix[j] >= W || iy[j] >= H);
        if (__all(none)) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                q[j].p00 = q[j].p01 = q[j].p10 = q[j].p11 = cv;
                out[j] = true;
            }
            return;
        }
    }
    if (__all(inner)) {
        const gptr<const uint8_t> base = gp((const uint8_t *)s.p);
        const uint32_t st = (uint32_t)s.stride;
        uint32_t a[4], b[4], c[4], d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t o = (uint32_t)iy[j] * st + (uint32_t)ix[j];
            a[j] = base[o];
            b[j] = base[o + 1];
            c[j] = base[o + st];
            d[j] = base[o + st + 1];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            q[j].p00 = (float)a[j];
            q[j].p01 = (float)b[j];
            q[j].p10 = (float)c[j];
            q[j].p11 = (float)d[j];
            out[j] = false;
        }
        return;
    }
    Taps t[4];
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = map_taps(ix[j], iy[j], W, H, border);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        q[j] = load_quad_u8(s, t[j], cv);
        out[j] = t[j].out;
    }
}

// Four bilinear pixels of a one-channel u8 warpAffine, at x, x + dx,
// x + 2dx, x + 3dx of one row.
__device__ __attribute__((always_inline)) inline uint32_t affine4_u8(const AffineParams &p, int x, int dx, float bx,
                                                                     float by) {
    int ix[4], iy[4];
    float ax[4], ay[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float xf = (float)(x + j * dx);
        const float sx = fmaf(p.m[0], xf, bx), sy = fmaf(p.m[3], xf, by);
        const float fx = floorf(sx), fy = floorf(sy);
        ax[j] = sx - fx;
        ay[j] = sy - fy;
        ix[j] = to_i32(fx);
        iy[j] = to_i32(fy);
    }
    Quad q[4];
    bool out[4];
    quads_u8(p.s, p.border, p.cv, ix, iy, q, out);
    uint32_t res = 0;
    const float cvd = (float)p.cvd;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float v0 = fmaf(ax[j], q[j].p01 - q[j].p00, q[j].p00);
        const float v1 = fmaf(ax[j], q[j].p11 - q[j].p10, q[j].p10);
        res = pack_u8(out[j] ? cvd : fmaf(ay[j], v1 - v0, v0), (uint32_t)j, res);
    }
    return res;
}

// Four output bytes at (x, y): one dword store when it fits and is aligned.
__device__ __attribute__((always_inline)) inline void store4_u8(const DImg &d, int x, int y, uint32_t v) {
    const gptr<uint8_t> o = grow<uint8_t>(d, y) + x;
    if (x + 4 <= d.w && ((uintptr_t)o & 3) == 0) {
        *(gptr<uint32_t>)o = v;
    } else {
        for (int j = 0; j < 4 && x + j < d.w; j++) o[j] = (uint8_t)(v >> (8 * j));
    }
}

// ---------------------------------------------- LDS-staged bilinear warps
//
// One-channel u8 bilinear warps with BORDER_CONSTANT, 64 x 16 output
// pixels per workgroup, 4 adjacent pixels per thread. Gathering taps from
// global memory costs one address per byte in the texture addresser, which
// bounds these kernels; instead the workgroup stages the source box its
// taps can reach (the corners' taps, widened by one pixel, clipped to the
// image) in LDS with dword loads, and gathers from there. Every thread then
// checks that each of its in-image taps lies in the box; if one does not
// (or the box is larger than the LDS tile), the workgroup gathers from
// global memory instead, so the result never depends on the box estimate.

constexpr int kWarpTW = 64, kWarpTH = 16;            // output pixels per workgroup
constexpr int kWarpBoxRows = 64, kWarpBoxPitch = 128;  // staged source rows, bytes per staged row
constexpr int kWarpBoxBytes = 124;                     // widest staged source span (+3 alignment bytes)
constexpr int kWarpLdsWords = kWarpBoxRows * kWarpBoxPitch / 4;

__device__ __attribute__((always_inline)) inline int wtile_x(int bx) { return bx * kWarpTW + (threadIdx.x & 15) * 4; }
__device__ __attribute__((always_inline)) inline int wtile_y(int by) { return by * kWarpTH + (threadIdx.x >> 4); }

struct WarpCoord {
    int ix, iy;
    float ax, ay;
};

// warpAffine coordinates of pixel (x, y) (geometry.zig:432-533). Without
// the clamp of to_i32 for pixels whose coordinates are known to lie near
// the image (the clamp then changes nothing).
struct AffineCoord {
    const float *m;
    template <bool CLAMP = true>
    __device__ __attribute__((always_inline)) WarpCoord at(int x, int y) const {
        const float yf = (float)y;
        const float bxv = m[1] * yf + m[2];
        const float byv = m[4] * yf + m[5];
        const float xf = (float)x;
        const float sx = fmaf(m[0], xf, bxv), sy = fmaf(m[3], xf, byv);
        const float fx = floorf(sx), fy = floorf(sy);
        if constexpr (CLAMP) return WarpCoord{to_i32(fx), to_i32(fy), sx - fx, sy - fy};
        else return WarpCoord{(int)fx, (int)fy, sx - fx, sy - fy};
    }
    __device__ __attribute__((always_inline)) WarpCoord operator()(int x, int y) const { return at<true>(x, y); }
};

// One-channel u8 bilinear warpAffine, BORDER_CONSTANT (value cv): a 64 x 16
// tile, 4 adjacent pixels per thread.
//
// The box is the corners' taps widened by one pixel (the per-pixel
// coordinates differ from the corners' linear extremes by rounding only),
// clipped to the image. Every live thread checks that each of its in-image
// taps lies in the box, and the workgroup takes the global gathers if one
// does not or the box is larger than the LDS tile. Staged row r holds source
// bytes xlo .. xlo + 123 of row ylo + r from byte 0 (each lane aligns its
// dword with its neighbour's). A tile whose widened box lies inside the
// image ("interior") has only in-image taps: a pixel then reads its four
// taps at one LDS address plus immediate offsets. Within a thread's 4
// adjacent pixels, each coordinate is a single rounding of a linear function
// of x, hence monotone, so the end pixels bound the others.
__device__ __attribute__((always_inline)) inline void affine_box_tile(const AffineParams &p, int bx, int by,
                                                                      uint32_t *lds) {
    const DImg &d = p.d, &s = p.s;
    const int W = s.w, H = s.h;
    const int x = wtile_x(bx), y = wtile_y(by);
    const bool live = x < d.w && y < d.h;
    const AffineCoord coord{p.m};
    int xlo = 0x7fffffff, xhi = -0x7fffffff, ylo = 0x7fffffff, yhi = -0x7fffffff;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const WarpCoord e = coord(bx * kWarpTW + (k & 1) * (kWarpTW - 1), by * kWarpTH + (k >> 1) * (kWarpTH - 1));
        xlo = min(xlo, e.ix);
        xhi = max(xhi, e.ix);
        ylo = min(ylo, e.iy);
        yhi = max(yhi, e.iy);
    }
    const bool interior = xlo - 1 >= 0 && ylo - 1 >= 0 && xhi + 2 <= W - 1 && yhi + 2 <= H - 1;
    xlo = max(xlo - 1, 0);
    ylo = max(ylo - 1, 0);
    xhi = min(xhi + 2, W - 1);
    yhi = min(yhi + 2, H - 1);
    const int nb = xhi - xlo + 1, nr = yhi - ylo + 1;
    const bool fits = nb > 0 && nr > 0 && nb <= kWarpBoxBytes && nr <= kWarpBoxRows &&
                      (unsigned long)s.h * (unsigned long)s.stride < (1ul << 31);
    // Interior tile: its pixels' coordinates lie within two pixels of the
    // corners' range, inside the image, so the clamp is not needed.
    WarpCoord c[4];
    if (interior) {
#pragma unroll
        for (int j = 0; j < 4; j++) c[j] = coord.at<false>(x + j, y);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) c[j] = coord(x + j, y);
    }
    // A tile whose every live pixel lies outside by the constant-border rule
    // (ix < -1, iy < -1, ix >= W or iy >= H: the value without reading a tap)
    // is written at once.
    bool none = true;
#pragma unroll
    for (int j = 0; j < 4; j++) none = none && (c[j].ix < -1 || c[j].iy < -1 || c[j].ix >= W || c[j].iy >= H);
    if (__syncthreads_and(!live || none)) {
        if (live) {
            const uint32_t v = (uint32_t)p.cvd;
            store4_u8(d, x, y, v * 0x01010101u);
        }
        return;
    }
    bool bad = false;
    if (live) {
        if (interior) {
            bad = min(c[0].ix, c[3].ix) < xlo || max(c[0].ix, c[3].ix) + 1 > xhi || min(c[0].iy, c[3].iy) < ylo ||
                  max(c[0].iy, c[3].iy) + 1 > yhi;
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x0 = c[j].ix, y0 = c[j].iy;
                bad = bad || ((unsigned)x0 < (unsigned)W && (x0 < xlo || x0 > xhi)) ||
                      ((unsigned)(x0 + 1) < (unsigned)W && (x0 + 1 < xlo || x0 + 1 > xhi)) ||
                      ((unsigned)y0 < (unsigned)H && (y0 < ylo || y0 > yhi)) ||
                      ((unsigned)(y0 + 1) < (unsigned)H && (y0 + 1 < ylo || y0 + 1 > yhi));
            }
        }
    }
    Quad q[4];
    bool out[4];
    const bool use = __syncthreads_or(bad) == 0 && fits;
    if (!use) {
        int ix[4], iy[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ix[j] = c[j].ix;
            iy[j] = c[j].iy;
        }
        quads_u8(s, BORDER_CONSTANT, p.cv, ix, iy, q, out);
    } else {
        // Stage rows ylo..yhi, 8 rows x 32 dwords per round: lane k of a
        // 32-lane group loads the k-th aligned dword at or after the row's
        // byte xlo - a, then aligns it with lane k + 1's.
        const int k = threadIdx.x & 31;
        uint32_t v[kWarpBoxRows / 8];
#pragma unroll
        for (int t = 0; t < kWarpBoxRows / 8; t++) {
            const int r = t * 8 + (int)(threadIdx.x >> 5);
            v[t] = 0;
            if (r < nr) {
                const gptr<const uint8_t> row = grow<const uint8_t>(s, ylo + r);
                const int a = (int)((uintptr_t)(row + xlo) & 3);
                const int c0 = xlo - a + 4 * k;  // row offset of this dword
                if (4 * k < nb + a) {
                    if (c0 >= 0 && c0 + 4 <= W) {
                        v[t] = *(gptr<const uint32_t>)(row + c0);
                    } else {
                        uint32_t w = 0;
                        for (int b = 0; b < 4; b++)
                            if (c0 + b >= 0 && c0 + b < W) w |= (uint32_t)row[c0 + b] << (8 * b);
                        v[t] = w;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < kWarpBoxRows / 8; t++) {
            const int r = t * 8 + (int)(threadIdx.x >> 5);
            const uint32_t nx = __shfl_down(v[t], 1);
            const int a = (int)((uintptr_t)(grow<const uint8_t>(s, ylo + r) + xlo) & 3);
            if (r < nr && k < 31) lds[r * (kWarpBoxPitch / 4) + k] = __builtin_amdgcn_alignbyte(nx, v[t], (uint32_t)a);
        }
        __syncthreads();
        const uint8_t *box = reinterpret_cast<const uint8_t *>(lds);
        if (interior) {
            // Only live threads' taps were checked against the box.
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint8_t *b = box + (live ? (c[j].iy - ylo) * kWarpBoxPitch + (c[j].ix - xlo) : 0);
                q[j].p00 = (float)b[0];
                q[j].p01 = (float)b[1];
                q[j].p10 = (float)b[kWarpBoxPitch];
                q[j].p11 = (float)b[kWarpBoxPitch + 1];
                out[j] = false;
            }
        } else {
            const float cv = p.cv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x0 = c[j].ix, y0 = c[j].iy;
                const int r0 = min(max(y0 - ylo, 0), nr - 1), r1 = min(max(y0 + 1 - ylo, 0), nr - 1);
                const int k0 = min(max(x0 - xlo, 0), nb - 1), k1 = min(max(x0 + 1 - xlo, 0), nb - 1);
                const uint32_t b00 = box[r0 * kWarpBoxPitch + k0], b01 = box[r0 * kWarpBoxPitch + k1];
                const uint32_t b10 = box[r1 * kWarpBoxPitch + k0], b11 = box[r1 * kWarpBoxPitch + k1];
                const bool vx0 = (unsigned)x0 < (unsigned)W, vx1 = (unsigned)(x0 + 1) < (unsigned)W;
                const bool vy0 = (unsigned)y0 < (unsigned)H, vy1 = (unsigned)(y0 + 1) < (unsigned)H;
                q[j].p00 = vx0 && vy0 ? (float)b00 : cv;
                q[j].p01 = vx1 && vy0 ? (float)b01 : cv;
                q[j].p10 = vx0 && vy1 ? (float)b10 : cv;
                q[j].p11 = vx1 && vy1 ? (float)b11 : cv;
                out[j] = x0 < -1 || y0 < -1 || x0 >= W || y0 >= H;
            }
        }
    }
    if (!live) return;
    uint32_t res = 0;
    const float cvd = (float)p.cvd;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float v0 = fmaf(c[j].ax, q[j].p01 - q[j].p00, q[j].p00);
        const float v1 = fmaf(c[j].ax, q[j].p11 - q[j].p10, q[j].p10);
        const float v = fmaf(c[j].ay, v1 - v0, v0);
        res = pack_u8(out[j] ? cvd : v, (uint32_t)j, res);
    }
    store4_u8(d, x, y, res);
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

// Three-channel u8 bilinear warpAffine with a constant border
// (geometry.zig:508-533, the same arithmetic as affine_pixel): 64 x 4 pixels
// per tile, one per thread. A pixel whose footprint lies inside the image
// loads each tap row's two pixels (6 bytes) with one 12-byte buffer load at
// the dword at or below them, v_alignbyte picking the bytes, instead of six
// byte gathers; the range check of the buffer (lim bytes from the dword
// below the image base) makes every offset safe, and pixels whose footprint
// leaves the image (or whose second row's load would pass the image's last
// byte) take the byte-by-byte rules of affine_pixel.
__device__ __attribute__((always_inline)) inline void affine_c3_u8_tile(const AffineParams &p, uint32_t lim, int bx, int by) {
    const DImg &s = p.s, &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x = bx * 64 + tile_tx();
    if (y >= d.h || x >= d.w) return;
    const float yf = (float)y, xf = (float)x;
    const float sx = fmaf(p.m[0], xf, p.m[1] * yf + p.m[2]);
    const float sy = fmaf(p.m[3], xf, p.m[4] * yf + p.m[5]);
    const float fx = floorf(sx), fy = floorf(sy);
    const float ax = sx - fx, ay = sy - fy;
    const int ix = (int)to_idx(fx), iy = (int)to_idx(fy);
    const int W = s.w, H = s.h;
    const gptr<uint8_t> o = grow<uint8_t>(d, y) + 3L * x;
    if (ix < -1 || iy < -1 || ix >= W || iy >= H) {
        const uint8_t c = (uint8_t)p.cvd;
        o[0] = c;
        o[1] = c;
        o[2] = c;
        return;
    }
    const uint32_t st = (uint32_t)s.stride, b3 = (uint32_t)((uintptr_t)s.p & 3);
    const uint32_t o0 = (uint32_t)iy * st + 3u * (uint32_t)ix + b3, o1 = o0 + st;
    if ((unsigned)ix < (unsigned)(W - 1) && (unsigned)iy < (unsigned)(H - 1) && (o1 & ~3u) + 12 <= lim) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(s.p - b3, 0, (int)lim, 0x00020000);
        typedef unsigned int vu32x3 __attribute__((ext_vector_type(3)));
        const vu32x3 r0 = __builtin_bit_cast(vu32x3, __builtin_amdgcn_raw_buffer_load_b96(rs, o0 & ~3u, 0, 0));
        const vu32x3 r1 = __builtin_bit_cast(vu32x3, __builtin_amdgcn_raw_buffer_load_b96(rs, o1 & ~3u, 0, 0));
        // bytes 0..5 of each tap row: pixel ix (0-2) and ix + 1 (3-5)
        const uint32_t a0 = __builtin_amdgcn_alignbyte(r0.y, r0.x, o0 & 3), a1 = __builtin_amdgcn_alignbyte(r0.z, r0.y, o0 & 3);
        const uint32_t b0 = __builtin_amdgcn_alignbyte(r1.y, r1.x, o1 & 3), b1 = __builtin_amdgcn_alignbyte(r1.z, r1.y, o1 & 3);
        const uint64_t ra = (uint64_t)a1 << 32 | a0, rb = (uint64_t)b1 << 32 | b0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float p00 = (float)((ra >> (8 * c)) & 255u), p01 = (float)((ra >> (8 * c + 24)) & 255u);
            const float p10 = (float)((rb >> (8 * c)) & 255u), p11 = (float)((rb >> (8 * c + 24)) & 255u);
            const float v0 = fmaf(ax, p01 - p00, p00);
            const float v1 = fmaf(ax, p11 - p10, p10);
            o[c] = (uint8_t)pack_u8(fmaf(ay, v1 - v0, v0), 0, 0);
        }
        return;
    }
    const bool vx0 = (unsigned)ix < (unsigned)W, vx1 = (unsigned)(ix + 1) < (unsigned)W;
    const bool vy0 = (unsigned)iy < (unsigned)H, vy1 = (unsigned)(iy + 1) < (unsigned)H;
    const gptr<const uint8_t> ra = grow<const uint8_t>(s, vy0 ? iy : 0), rb = grow<const uint8_t>(s, vy1 ? iy + 1 : 0);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float p00 = vx0 && vy0 ? (float)ra[3 * ix + c] : p.cv;
        const float p01 = vx1 && vy0 ? (float)ra[3 * (ix + 1) + c] : p.cv;
        const float p10 = vx0 && vy1 ? (float)rb[3 * ix + c] : p.cv;
        const float p11 = vx1 && vy1 ? (float)rb[3 * (ix + 1) + c] : p.cv;
        const float v0 = fmaf(ax, p01 - p00, p00);
        const float v1 = fmaf(ax, p11 - p10, p10);
        o[c] = (uint8_t)pack_u8(fmaf(ay, v1 - v0, v0), 0, 0);
    }
}

// Four output bytes of lane tx at x0 + tx + 64j (j = 0..3) of row y: each
// store instruction writes 64 consecutive bytes.
__device__ __attribute__((always_inline)) inline void store4_strided_u8(const DImg &d, int x0, int y, uint32_t v) {
    const gptr<uint8_t> o = grow<uint8_t>(d, y);
    const int x = x0 + tile_tx();
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (x + 64 * j < d.w) o[x + 64 * j] = (uint8_t)(v >> (8 * j));
}

// One-channel u8 bilinear warp tiles are 256 x 4 pixels; lane tx of a wave
// computes pixels x0 + tx + 64j, so one gather instruction covers 64
// neighbouring output pixels, whose taps lie on few source rows under a
// rotation (a lane with 4 adjacent pixels would spread one instruction over
// 256 pixels and many more cache lines).
__device__ __attribute__((always_inline)) inline void affine_lin4_tile(const AffineParams &p, int bx, int by) {
    const DImg &d = p.d;
    const int y = by * 4 + tile_ty();
    const int x0 = bx * 256;
    if (y >= d.h || x0 + tile_tx() >= d.w) return;
    const float yf = (float)y;
    const float bxv = p.m[1] * yf + p.m[2];
    const float byv = p.m[4] * yf + p.m[5];
    store4_strided_u8(d, x0, y, affine4_u8(p, x0 + tile_tx(), 64, bxv, byv));
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
__device__ __attribute__((always_inline)) inline int32_t lin_h(uint32_t s0, uint32_t s1, int32_t w0, int32_t w1) {
    return (int16_t)(((int32_t)s0 * w0 + (int32_t)s1 * w1) >> 4);
}
__device__ __attribute__((always_inline)) inline uint32_t lin_v(int32_t h0, int32_t h1, int32_t b0, int32_t b1) {
    const int32_t t = ((h0 * b0) >> 16) + ((h1 * b1) >> 16);
    const int32_t v = (t + 2) >> 2;
    return (uint32_t)(v < 0 ? 0 : v > 255 ? 255 : v);
}

// Stages source rows ylo .. ylo + nr - 1, bytes c_lo .. c_lo + 4 nd - 4 of
// each, as aligned dwords: LDS word r * PITCH + k holds the row's bytes from
// c_lo - a + 4k on, where a = box_shift(s, ylo + r, c_lo). Dwords wholly
// inside the row are single loads; bytes outside the row stay zero. Slots
// are spread over the workgroup (the float quotient is exact for these
// sizes and corrected all the same). The caller synchronises.
__device__ __attribute__((always_inline)) inline int box_shift(const DImg &s, int y, int c_lo) {
    return (int)((uintptr_t)(grow<const uint8_t>(s, y) + c_lo) & 3);
}
template <int PITCH>
__device__ __attribute__((always_inline)) inline void stage_box(const DImg &s, int ylo, int nr, int c_lo, int nd, uint32_t *lds) {
    const int ns = s.w * s.cn;
    const float rcp = 1.0f / (float)nd;
    const int total = nr * nd;
    // Four slots per thread per round, their loads issued together.
    for (int base = threadIdx.x; base < total; base += 4 * kTileThreads) {
        uint32_t v[4];
        int slot[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = base + u * kTileThreads;
            slot[u] = -1;
            v[u] = 0;
            if (i < total) {
                int r = (int)(((float)i + 0.5f) * rcp);
                int k = i - r * nd;
                if (k >= nd) r++, k -= nd;
                if (k < 0) r--, k += nd;
                const gptr<const uint8_t> row = grow<const uint8_t>(s, ylo + r);
                const int c = c_lo - box_shift(s, ylo + r, c_lo) + 4 * k;
                if (c >= 0 && c + 4 <= ns) {
                    v[u] = *(gptr<const uint32_t>)(row + c);
                } else {
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (c + t >= 0 && c + t < ns) v[u] |= (uint32_t)row[c + t] << (8 * t);
                }
                slot[u] = r * PITCH + k;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (slot[u] >= 0) lds[slot[u]] = v[u];
    }
}

// Direct gathers: 256 output samples x 4 rows.
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
        const int32_t h0 = lin_h(r0[tx.o0 + ch], r0[tx.o1 + ch], tx.w0, tx.w1);
        const int32_t h1 = lin_h(r1[tx.o0 + ch], r1[tx.o1 + ch], tx.w0, tx.w1);
        o[e] = (uint8_t)lin_v(h0, h1, ty.w0, ty.w1);
    }
}

__device__ __attribute__((always_inline)) inline LinTap load_tap(gptr<const LinTap> p, int i) {
    const u32x4 v = *(gptr<const u32x4>)(p + i);
    LinTap t;
    t.o0 = (int32_t)v.x;
    t.o1 = (int32_t)v.y;
    t.w0 = (int32_t)v.z;
    t.w1 = (int32_t)v.w;
    return t;
}

// e / cn for the channel counts 1-4 without an integer division (cn 3:
// the high half of e * ceil(2^33 / 3), exact for every 32-bit e).
__device__ __attribute__((always_inline)) inline int div_cn(int e, int cn) {
    switch (cn) {
    case 1: return e;
    case 2: return e >> 1;
    case 3: return (int)(__umulhi((uint32_t)e, 0xAAAAAAABu) >> 1);
    case 4: return e >> 2;
    default: return e / cn;
    }
}

// Through LDS: 256 output samples x kLinTH rows, for tap tables that
// linear_box_fits() accepted (taps that never decrease, every tile's source
// box within kLinRows x kLinWords). The box of a tile, rows yt[y0].o0 ..
// yt[y_last].o1 and samples xt[x_first].o0 .. xt[x_last].o1 + cn - 1, then
// holds every tap of the tile. It is staged as aligned dwords; pass 1 forms
// H for every staged row and the tile's 256 samples (as i16 in LDS), pass 2
// blends the two H rows of each output row and stores whole dwords.
__device__ __attribute__((always_inline)) inline void linear_box_tile(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt,
                                                                      int bx, int by, uint32_t *lds) {
    const int cn = d.cn, n = d.w * cn;
    const int e_first = bx * 256, y0 = by * kLinTH;
    const int e_last = min(n, e_first + 256) - 1, y_last = min(d.h, y0 + kLinTH) - 1;
    const gptr<const LinTap> xg = gp(xt), yg = gp(yt);
    const int g = tile_tx(), wv = __builtin_amdgcn_readfirstlane(tile_ty());
    // Every tap this thread needs, loaded before the box so the loads fly
    // together: its 4 samples' column taps and its wave's 4 row taps.
    LinTap tx[4], ty[4];
#pragma unroll
    for (int j = 0; j < 4; j++) tx[j] = load_tap(xg, div_cn(min(e_first + 4 * g + j, e_last), cn));
#pragma unroll
    for (int q = 0; q < 4; q++) ty[q] = load_tap(yg, min(y0 + wv * 4 + q, y_last));
    const int c_lo = load_tap(xg, div_cn(e_first, cn)).o0;
    const int ylo = load_tap(yg, y0).o0, yhi = load_tap(yg, y_last).o1;
    const int nr = yhi - ylo + 1, nd = (load_tap(xg, div_cn(e_last, cn)).o1 + cn - 1 - c_lo + 7) >> 2;
    uint8_t *box = reinterpret_cast<uint8_t *>(lds);
    uint32_t *hb = lds + kLinRows * kLinWords;
    stage_box<kLinWords>(s, ylo, nr, c_lo, nd, lds);
    int32_t b0[4], b1[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int e = min(e_first + 4 * g + j, e_last);
        const int ch = e - div_cn(e, cn) * cn;
        b0[j] = tx[j].o0 + ch - c_lo;
        b1[j] = tx[j].o1 + ch - c_lo;
    }
    __syncthreads();
    // Pass 1: H rows.
    for (int r = wv; r < nr; r += 4) {
        const uint8_t *rb = box + r * (kLinWords * 4) + box_shift(s, ylo + r, c_lo);
        uint32_t p[2] = {0u, 0u};
#pragma unroll
        for (int j = 0; j < 4; j++)
            p[j >> 1] |= ((uint32_t)lin_h(rb[b0[j]], rb[b1[j]], tx[j].w0, tx[j].w1) & 0xffffu) << (16 * (j & 1));
        *reinterpret_cast<u32x2 *>(hb + r * 128 + 2 * g) = u32x2{p[0], p[1]};
    }
    __syncthreads();
    // Pass 2: rows y0 + 4 wv .. + 3.
    const int e = e_first + 4 * g;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int y = y0 + wv * 4 + q;
        if (y > y_last) break;
        const u32x2 h0 = *reinterpret_cast<const u32x2 *>(hb + (ty[q].o0 - ylo) * 128 + 2 * g);
        const u32x2 h1 = *reinterpret_cast<const u32x2 *>(hb + (ty[q].o1 - ylo) * 128 + 2 * g);
        uint32_t out = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t a = j < 2 ? h0.x : h0.y, b = j < 2 ? h1.x : h1.y;
            const int32_t v0 = (int16_t)(a >> (16 * (j & 1))), v1 = (int16_t)(b >> (16 * (j & 1)));
            out |= lin_v(v0, v1, ty[q].w0, ty[q].w1) << (8 * j);
        }
        if (e > e_last) continue;
        const gptr<uint8_t> o = grow<uint8_t>(d, y) + e;
        if (e + 3 <= e_last && ((uintptr_t)o & 3) == 0) {
            *(gptr<uint32_t>)o = out;
        } else {
            for (int j = 0; j < 4 && e + j <= e_last; j++) o[j] = (uint8_t)(out >> (8 * j));
        }
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

// area_frac_tile through LDS, for tables that area_box_fits() accepted
// (every cell with 1 .. kAreaMaxCell entries, each tile's entries within
// the span of its first and last cell, and that span within kAreaRows x
// kAreaWords): 256 output samples x kAreaTH rows. The box is staged as
// aligned dwords; pass 1 forms, for every staged row and the tile's 256
// samples, the f32 sum of s * alpha over the sample's cell in table order
// (one sample per thread, the cell's entries in registers); pass 2 sums
// those rows times beta in table order, the same operations in the same
// order as area_frac_tile.
__device__ __attribute__((always_inline)) inline AreaEntry load_entry(gptr<const AreaEntry> p, int i) {
    const u32x2 v = *(gptr<const u32x2>)(p + i);
    AreaEntry e;
    e.si = (int32_t)v.x;
    e.alpha = __uint_as_float(v.y);
    return e;
}

__device__ __attribute__((always_inline)) inline void area_box_tile(const DImg &s, const DImg &d, const AreaEntry *xt, const int32_t *xs,
                                                                    const AreaEntry *yt, const int32_t *ys, int bx, int by, uint32_t *lds) {
    const int cn = d.cn, n = d.w * cn;
    const int e_first = bx * 256, y0 = by * kAreaTH;
    const int e_last = min(n, e_first + 256) - 1, y_last = min(d.h, y0 + kAreaTH) - 1;
    const gptr<const AreaEntry> xg = gp(xt), yg = gp(yt);
    const gptr<const int32_t> xsg = gp(xs), ysg = gp(ys);
    const int c_lo = xg[xsg[e_first / cn]].si * cn;
    const int c_hi = xg[xsg[e_last / cn + 1] - 1].si * cn + cn - 1;
    const int ylo = yg[ysg[y0]].si, yhi = yg[ysg[y_last + 1] - 1].si;
    const int nr = yhi - ylo + 1, nd = (c_hi - c_lo + 7) >> 2;
    stage_box<kAreaWords>(s, ylo, nr, c_lo, nd, lds);
    // This thread's sample of pass 1 and its cell's entries.
    const int e = min(e_first + (int)threadIdx.x, e_last);
    const int x = e / cn, ch = e - x * cn;
    const int q0 = xsg[x], nq = xsg[x + 1] - q0;
    int xo[kAreaMaxCell];
    float al[kAreaMaxCell];
#pragma unroll
    for (int q = 0; q < kAreaMaxCell; q++) {
        const AreaEntry en = load_entry(xg, q0 + min(q, nq - 1));
        xo[q] = en.si * cn + ch - c_lo;
        al[q] = en.alpha;
    }
    __syncthreads();
    const uint8_t *box = reinterpret_cast<const uint8_t *>(lds);
    float *hb = reinterpret_cast<float *>(lds + kAreaRows * kAreaWords);
    for (int r = 0; r < nr; r++) {
        const uint8_t *rb = box + r * (kAreaWords * 4) + box_shift(s, ylo + r, c_lo);
        float h = 0.0f;
#pragma unroll
        for (int q = 0; q < kAreaMaxCell; q++)
            if (q < nq) h = h + (float)rb[xo[q]] * al[q];
        hb[r * 256 + threadIdx.x] = h;
    }
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(tile_ty()), g = tile_tx();
    const int e4 = e_first + 4 * g;
#pragma unroll 1
    for (int y = y0 + wv; y <= y_last; y += 4) {
        const int k0 = ysg[y], k1 = ysg[y + 1];
        float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = k0; k < k1; k++) {
            const AreaEntry en = load_entry(yg, k);
            const float4 hv = *reinterpret_cast<const float4 *>(hb + (en.si - ylo) * 256 + 4 * g);
            sum[0] = sum[0] + hv.x * en.alpha;
            sum[1] = sum[1] + hv.y * en.alpha;
            sum[2] = sum[2] + hv.z * en.alpha;
            sum[3] = sum[3] + hv.w * en.alpha;
        }
        if (e4 > e_last) continue;
        uint32_t out = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) out |= (uint32_t)sat_cast<uint8_t>((double)sum[j]) << (8 * j);
        const gptr<uint8_t> o = grow<uint8_t>(d, y) + e4;
        if (e4 + 3 <= e_last && ((uintptr_t)o & 3) == 0) {
            *(gptr<uint32_t>)o = out;
        } else {
            for (int j = 0; j < 4 && e4 + j <= e_last; j++) o[j] = (uint8_t)(out >> (8 * j));
        }
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
