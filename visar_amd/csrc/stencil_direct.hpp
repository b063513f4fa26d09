// stencil_direct.hpp -- the packed 16-bit stencils (GaussianBlur u8 3x3,
// 5x5, 7x7 with sigma 0; Sobel 3x3 dx or dy, u8 -> i16, scale 1) on
// one-channel images, with the rows streamed from memory into registers
// and no LDS, for gfx950.
//
// The sums are the reference's exact integer ones (filter.zig:120-341 with
// the taps of :350-483; the same reduced taps as stencil_swar_tile in
// tiles.hpp), taken in the other order: the vertical taps first, on the
// lane's own samples, then the horizontal taps on those column sums. The
// sums are exact integers either way (every partial sum stays below 2^16,
// or is taken in 32 bits for 7x7), so both kernels give the same bytes.
// Out-of-image columns take the border rule on the column sums, which is
// the rule on the samples: a column's sum depends only on which column
// the border maps it to.
//
// Each wave owns TB output rows of a column strip: lane i holds bytes
// X0 + 4 (i - 1) .. + 3 of every row, lanes 1..62 produce output (a strip
// is 248 bytes) and lanes 0 and 63 only lend their column sums to their
// neighbours (DPP wave shifts), so a row is one dword load per lane. At the
// image's left and right edges the lane holding the edge dword rebuilds its
// outer neighbours from its own column sums with one permute per border
// kind. Source rows outside the image are border-mapped on the scalar unit.
// Requires rows of a multiple of 4 bytes (at least 8), each starting on a
// dword, in source and destination, and a border other than WRAP.
#pragma once

#include "tiles.hpp"

namespace visar {

struct StencilDirectArgs {
    StencilArgs s;
    uint32_t span_s, span_d;  // bytes from each image's first byte to its last + 1 (< 2^31)
    int strips, ntiles, xcd;  // the tile order: tile (strip, group of wpg bands)
    int bands;                // output row bands of TB rows
    int wpg;                  // waves per workgroup (1, 2 or 4), on consecutive bands of one strip
};

namespace sdirect {

constexpr int kStrip = 248;  // output bytes per strip (lanes 1..62)
// DPP controls: lane i takes lane i - 1's value (shr) or lane i + 1's (shl);
// the lane without a source reads 0 (bound_ctrl), so no old value has to be
// set up in the destination first.
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));

__device__ __attribute__((always_inline)) inline uint32_t dpp_shr(uint32_t v) {
    return __builtin_amdgcn_mov_dpp(v, kDppWaveShr1, 0xf, 0xf, true);
}
__device__ __attribute__((always_inline)) inline uint32_t dpp_shl(uint32_t v) {
    return __builtin_amdgcn_mov_dpp(v, kDppWaveShl1, 0xf, 0xf, true);
}
__device__ __attribute__((always_inline)) inline uint32_t perm(uint32_t a, uint32_t b, uint32_t sel) {
    return __builtin_amdgcn_perm(a, b, sel);
}

// 16-bit lanes of a u32 (pairs of column sums): add, subtract, scale.
__device__ __attribute__((always_inline)) inline uint32_t padd(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __attribute__((always_inline)) inline uint32_t psub(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b));
}
__device__ __attribute__((always_inline)) inline uint32_t pmad(uint32_t a, uint16_t k, uint32_t acc) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) * k + __builtin_bit_cast(u16x2, acc));
}
// Weighted sum of pairs x[0..K) with taps TAP(i) (0, +-1 and small
// positive integers known at compile time), in wrapping 16-bit lanes.
template <typename TAP, int K, int I = 0>
__device__ __attribute__((always_inline)) inline uint32_t pdot(const uint32_t (&x)[K], uint32_t acc, bool first) {
    if constexpr (I == K) {
        return acc;
    } else {
        constexpr int w = TAP::at(I);
        if constexpr (w == 0) {
            return pdot<TAP, K, I + 1>(x, acc, first);
        } else {
            uint32_t v;
            if (first) v = w == 1 ? x[I] : w == -1 ? psub(0u, x[I]) : pmad(x[I], (uint16_t)w, 0u);
            else v = w == 1 ? padd(acc, x[I]) : w == -1 ? psub(acc, x[I]) : pmad(x[I], (uint16_t)w, acc);
            return pdot<TAP, K, I + 1>(x, v, false);
        }
    }
}
template <int MODE>
struct HTap {
    static constexpr int at(int i) { return SwarTaps<MODE>::h[i]; }
};
template <int MODE>
struct VTap {
    static constexpr int at(int i) { return SwarTaps<MODE>::v[i]; }
};

// border.zig:22-39 for a row index p (a border other than WRAP): the row it
// reads, or -1 for a constant border. Reflects repeatedly (images shorter
// than the halo).
__device__ __attribute__((always_inline)) inline int row_map(int p, int n, int border) {
    if ((unsigned)p < (unsigned)n) return p;
    if (border == BORDER_CONSTANT) return -1;
    if (border == BORDER_REPLICATE || n == 1) return p < 0 ? 0 : n - 1;
    const int d = border == BORDER_REFLECT101 ? 1 : 0;
    while (p < 0 || p >= n) p = p < 0 ? -p - 1 + d : 2 * n - 1 - d - p;
    return p;
}

// Column sums C0..C3 of the lane's 4 columns are held as pairs lo = (C0,
// C1), hi = (C2, C3). An edge lane rebuilds its neighbouring pairs by the
// border rule on columns (border.zig:22-39 within 3 columns of a row of at
// least 8 bytes):
//   left  l1 = (C-2, C-1), l2 = (C-4, C-3)   right r1 = (C4, C5), r2 = (C6, C7)
//   reflect-101  C-1 = C1, C-2 = C2, C-3 = C3   C4 = C2, C5 = C1, C6 = C0
//   reflect      C-1 = C0, C-2 = C1, C-3 = C2   C4 = C3, C5 = C2, C6 = C1
//   replicate    C-k = C0                       C4.. = C3
//   constant     0 (a column outside the image sums zeros)
// The selectors are for v_perm(hi, lo): byte codes 0-3 pick bytes of lo,
// 4-7 bytes of hi, 12 a zero byte; the unused halves (C-4, C7) get any value.
struct EdgeSel {
    uint32_t l1, l2, r1, r2;
};
__device__ __attribute__((always_inline)) inline EdgeSel edge_sel(int border) {
    switch (border) {
    case BORDER_REFLECT101: return {0x03020504u, 0x07060706u, 0x03020504u, 0x01000100u};
    case BORDER_REFLECT: return {0x01000302u, 0x05040504u, 0x05040706u, 0x03020302u};
    case BORDER_REPLICATE: return {0x01000100u, 0x01000100u, 0x07060706u, 0x07060706u};
    default: return {0x0c0c0c0cu, 0x0c0c0c0cu, 0x0c0c0c0cu, 0x0c0c0c0cu};
    }
}

template <int MODE, int TB>
__device__ __attribute__((always_inline)) inline void direct_wave(const StencilDirectArgs &a, int strip, int band) {
    using T = SwarTaps<MODE>;
    constexpr int K = T::K, R = K / 2, ROWS = TB + 2 * R;
    constexpr bool SOBEL = MODE == SWAR_SOBEL_DX || MODE == SWAR_SOBEL_DY;
    const StencilArgs &p = a.s;
    const int lane = (int)threadIdx.x & 63;
    const int n = p.src.w, H = p.src.h, border = p.border;
    const int xc = strip * kStrip + 4 * (lane - 1);  // the lane's first byte
    const int Y0 = band * TB;                        // first output row
    if (Y0 >= H) return;
    // Buffer loads and stores: a scalar row offset plus the lane's column
    // offset. Every address lies inside the row it reads or writes: lanes
    // outside the row read its first or last dword (those values are
    // replaced by the edge rule or never used).
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(p.src.p, 0, a.span_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(p.dst.p, 0, a.span_d, kBufWord3);
    const int ss = (int)p.src.stride, ds = (int)p.dst.stride;
    const int xl = min(max(xc, 0), n - 4);
    // Row offsets first (scalar), then every load back to back.
    int so[ROWS];
    uint64_t zero_rows = 0;  // rows outside the image under a constant border (uniform)
    if (Y0 - R >= 0 && Y0 - R + ROWS <= H) {
#pragma unroll
        for (int r = 0; r < ROWS; r++) so[r] = (Y0 - R + r) * ss;
    } else {
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const int y = row_map(Y0 - R + r, H, border);
            if (y < 0) zero_rows |= 1ull << r;
            so[r] = max(y, 0) * ss;
        }
    }
    uint32_t own[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) own[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, xl, so[r], 0);
    // Samples as pairs (b0, b1) and (b2, b3) in 16-bit lanes.
    uint32_t lo[ROWS], hi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const uint32_t w = (zero_rows >> r) & 1 ? 0u : own[r];
        lo[r] = perm(0u, w, 0x0c010c00u);
        hi[r] = perm(0u, w, 0x0c030c02u);
    }
    const bool e_left = xc == 0, e_right = xc == n - 4;
    const bool edge = __builtin_amdgcn_ballot_w64(e_left || e_right) != 0;  // uniform
    const EdgeSel es = edge_sel(border);
    const bool live = lane >= 1 && lane <= 62 && xc < n;
#pragma unroll
    for (int t = 0; t < TB; t++) {
        const int y = Y0 + t;
        if (y >= H) break;  // uniform
        uint32_t cl, ch;    // column sums (C0, C1), (C2, C3)
        {
            uint32_t xs[K], ys[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                xs[k] = lo[t + k];
                ys[k] = hi[t + k];
            }
            cl = pdot<VTap<MODE>, K>(xs, 0u, true);
            ch = pdot<VTap<MODE>, K>(ys, 0u, true);
        }
        // The neighbouring lanes' column sums.
        uint32_t l1 = dpp_shr(ch), r1 = dpp_shl(cl), l2 = 0, r2 = 0;
        if constexpr (K == 7) {
            l2 = dpp_shr(cl);
            r2 = dpp_shl(ch);
        }
        if (edge) {
            if (e_left) {
                l1 = perm(ch, cl, es.l1);
                if constexpr (K == 7) l2 = perm(ch, cl, es.l2);
            }
            if (e_right) {
                r1 = perm(ch, cl, es.r1);
                if constexpr (K == 7) r2 = perm(ch, cl, es.r2);
            }
        }
        // Pairs P(d) = (C_d, C_d+1): for even d a neighbour's or the lane's
        // own pair, for odd d the halves of two consecutive pairs.
        const uint32_t am1 = __builtin_amdgcn_alignbyte(cl, l1, 2);  // (C-1, C0)
        const uint32_t a1 = __builtin_amdgcn_alignbyte(ch, cl, 2);   // (C1, C2)
        const uint32_t a3 = __builtin_amdgcn_alignbyte(r1, ch, 2);   // (C3, C4)
        uint32_t o01, o23;  // outputs (0, 1) and (2, 3)
        if constexpr (MODE == SWAR_G7) {
            // 2 (C-3 + C3) + 7 (C-2 + C2) + 14 (C-1 + C1) + 18 C0 in 32 bits
            const uint32_t am3 = __builtin_amdgcn_alignbyte(l1, l2, 2);  // (C-3, C-2)
            const uint32_t a5 = __builtin_amdgcn_alignbyte(r2, r1, 2);   // (C5, C6)
            uint32_t f[4];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                // P(d) relative to the pair's first output, sample 2q
                const uint32_t pm3 = q == 0 ? am3 : am1, pm2 = q == 0 ? l1 : cl, pm1 = q == 0 ? am1 : a1;
                const uint32_t p0 = q == 0 ? cl : ch, p1 = q == 0 ? a1 : a3, p2 = q == 0 ? ch : r1;
                const uint32_t p3 = q == 0 ? a3 : a5;
                const uint32_t s3 = padd(pm3, p3), s2 = padd(pm2, p2), s1 = padd(pm1, p1);
#pragma unroll
                for (int l = 0; l < 2; l++) {
                    const uint32_t sel = l == 0 ? 0x05040100u : 0x07060302u;
                    const u16x2 ab = __builtin_bit_cast(u16x2, perm(s2, s3, sel));
                    const u16x2 cm = __builtin_bit_cast(u16x2, perm(p0, s1, sel));
                    uint32_t v = __builtin_amdgcn_udot2(ab, (u16x2){2, 7}, 2048u, false);
                    v = __builtin_amdgcn_udot2(cm, (u16x2){14, 18}, v, false);
                    f[2 * q + l] = v >> 12;
                }
            }
            const uint32_t v4 = f[0] | (f[1] << 8) | (f[2] << 16) | (f[3] << 24);
            if (live) __builtin_amdgcn_raw_buffer_store_b32(v4, rd, xc, y * ds, 0);
            continue;
        } else if constexpr (K == 3) {
            const uint32_t x01[3] = {am1, cl, a1}, x23[3] = {a1, ch, a3};
            o01 = pdot<HTap<MODE>, 3>(x01, 0u, true);
            o23 = pdot<HTap<MODE>, 3>(x23, 0u, true);
        } else {
            const uint32_t x01[5] = {l1, am1, cl, a1, ch}, x23[5] = {cl, a1, ch, a3, r1};
            o01 = pdot<HTap<MODE>, 5>(x01, 0u, true);
            o23 = pdot<HTap<MODE>, 5>(x23, 0u, true);
        }
        if constexpr (SOBEL) {
            if (live) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vu32x2, u32x2{o01, o23}), rd, 2 * xc, y * ds, 0);
        } else {
            constexpr uint16_t add = MODE == SWAR_G3 ? 8 : 128, sh = MODE == SWAR_G3 ? 4 : 8;
            const u16x2 f01 = (__builtin_bit_cast(u16x2, o01) + add) >> sh;
            const u16x2 f23 = (__builtin_bit_cast(u16x2, o23) + add) >> sh;
            const uint32_t v4 = perm(__builtin_bit_cast(uint32_t, f23), __builtin_bit_cast(uint32_t, f01), 0x06040200u);
            if (live) __builtin_amdgcn_raw_buffer_store_b32(v4, rd, xc, y * ds, 0);
        }
    }
}

// One workgroup: wpg waves on consecutive bands of one strip, tiles in XCD
// bands (neighbouring strips and bands share their halo lines in one L2).
template <int MODE, int TB>
__global__ __launch_bounds__(256) void sep_direct_u8(StencilDirectArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    direct_wave<MODE, TB>(a, strip, a.wpg * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

}  // namespace sdirect

}  // namespace visar
