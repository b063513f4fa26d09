// resize_stream.hpp -- LINEAR resize of u8 images (1-4 channels) with the
// source rows streamed from memory into registers and no LDS, for gfx950.
//
// Arithmetic: resize_u8.zig:651-711, as in linear_tile: per source row
// H = (s[o0] w0 + s[o1] w1) >> 4 (i16), out = sat_u8((((H0 b0) >> 16) +
// ((H1 b1) >> 16) + 2) >> 2). For tap tables whose weights are at least 0
// with w0 + w1 <= 2050 (linearTapU8 gives sums of 2048) every H lies in
// [0, 32671], so the i16 cast is the identity, every product fits 31 bits
// and out = (m0 + m1 + 2) >> 2 <= 255 needs no clamp. Each step is then one
// unsigned instruction: H << 8 = dot2((s0, s1), (16 w0, 16 w1)) & ~0xff (the
// dot is 16 times the reference's sum; the mask is the >> 4), and
// m = mul_hi_u24(H << 8, b << 8) = (H b) >> 16.
//
// Layout: each wave owns a run of bands of TB output rows of a strip of
// output samples, lane i holding samples X0 + 4i .. + 3; strips advance by
// 252 samples, so lane 63 recomputes the next strip's first four samples.
// The column taps are read once per wave. The source bytes of a lane's four
// samples' taps lie within 8 consecutive bytes of a row (checked on the
// host): a source row is one 12-byte load per lane at the dword at or below
// them, realigned with two alignbytes, and each sample's two bytes are
// gathered with one permute. A band's row taps sit in scalar registers and
// its output rows are unrolled: each output row loads the rows its taps
// bring in that the previous output row did not hold (all of a band's loads
// issued before its first use), forms their H, and reuses the others.
//
// Output dwords are aligned in memory whatever the row's alignment: with
// a = (address of sample X0) & 3, lane i stores the dword holding samples
// X0 + 4i - a .. + 3 - a, its low a bytes taken from lane i - 1 (a DPP
// shift). Lane 0 of every strip but the first leaves that dword to lane 63
// of the previous strip. Dwords that cross the row's ends are written byte
// by byte.
#pragma once

#include "tiles.hpp"

namespace visar {

namespace lstream {

constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access
constexpr int kDppWaveShr1 = 0x138;    // lane i takes lane i - 1's value (lane 0 keeps 0)
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int vu32x3 __attribute__((ext_vector_type(3)));

// (h * b) >> 16 for h = H << 8, b = B << 8 below 2^24: v_mul_hi_u32_u24.
__device__ __attribute__((always_inline)) inline uint32_t mulhi24(uint32_t h, uint32_t b) {
    return (uint32_t)(((uint64_t)(h & 0xffffffu) * (uint64_t)(b & 0xffffffu)) >> 32);
}

// A row's 12-byte window holding the lane's bytes: those at byte offset
// q = qx + r (r the row's offset) and the 7 after, loaded from the dword at
// or below q. On the image's last row a window that would cross the end of
// the range (where the load would return only zeros) starts lower, at
// lim - 12 (lim is a multiple of 4; rows of at least 12 bytes keep every
// other row's windows inside the range). AS: r is a multiple of 4.
template <bool AS>
__device__ __attribute__((always_inline)) inline vu32x3 load_window(__amdgpu_buffer_rsrc_t rs, uint32_t qx, uint32_t r, bool last,
                                                                     uint32_t lim) {
    uint32_t at;
    if (__builtin_expect(last, 0)) at = min((qx + r) & ~3u, lim - 12u);
    else if (AS) at = (qx & ~3u) + r;
    else at = (qx + r) & ~3u;
    return __builtin_bit_cast(vu32x3, __builtin_amdgcn_raw_buffer_load_b96(rs, at, 0, 0));
}
// The lane's 8 bytes from a window (w0 low, w1 high).
template <bool AS>
__device__ __attribute__((always_inline)) inline vu32x2 window_bytes(vu32x3 v, uint32_t qx, uint32_t r, bool last, uint32_t lim) {
    const uint32_t q = qx + r;
    if (__builtin_expect(last, 0)) {  // uniform
        const uint32_t dl = q - min(q & ~3u, lim - 12u);  // 0..11
        const uint32_t d0 = dl >= 8 ? v.z : dl >= 4 ? v.y : v.x;
        const uint32_t d1 = dl >= 8 ? 0u : dl >= 4 ? v.z : v.y;
        const uint32_t d2 = dl >= 4 ? 0u : v.z;
        return vu32x2{__builtin_amdgcn_alignbyte(d1, d0, dl), __builtin_amdgcn_alignbyte(d2, d1, dl)};
    }
    const uint32_t sh = AS ? qx : q;  // alignbyte takes its low 2 bits
    return vu32x2{__builtin_amdgcn_alignbyte(v.y, v.x, sh), __builtin_amdgcn_alignbyte(v.z, v.y, sh)};
}
// H << 8 of the lane's four samples from its 8 bytes of a row.
__device__ __attribute__((always_inline)) inline void h_row(vu32x2 w, const uint32_t (&sel)[4], const uint32_t (&wt)[4], uint32_t (&h)[4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const u16x2 pr = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(w.y, w.x, sel[j]));
        h[j] = __builtin_amdgcn_udot2(pr, __builtin_bit_cast(u16x2, wt[j]), 0u, false) & 0x00ffff00u;
    }
}

// One wave: bands [band0, band1) of TB output rows of one strip (RUN false:
// band0 only, with no loop, which keeps more registers free for TB 8). AD:
// every destination row starts on a dword (the lanes' dwords are their own
// samples); AS: every source row does (each window's dword offset is the
// lane's own plus the row's, with the same byte shift on every row).
template <int TB, bool AD, bool AS, bool RUN>
__device__ __attribute__((always_inline)) inline void lin_stream_wave(const LinStreamArgs &a, int strip, int band0, int band1) {
    const int lane = (int)threadIdx.x & 63;
    const int cn = a.d.cn, n = a.d.w * cn, dh = a.d.h;
    band1 = min(band1, (dh + TB - 1) / TB);
    if (band0 >= band1) return;  // uniform
    const int X0 = strip * kLinStreamStrip;
    const int e0 = X0 + 4 * lane;
    // This lane's column taps: its window start L, each sample's byte pair
    // selector and weights.
    const gptr<const LinTap> xg = gp(a.xt);
    LinTap tx[4];
#pragma unroll
    for (int j = 0; j < 4; j++) tx[j] = load_tap(xg, div_cn(min(e0 + j, n - 1), cn));
    int lo[4], hi[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int e = min(e0 + j, n - 1);
        const int ch = e - div_cn(e, cn) * cn;
        lo[j] = tx[j].o0 + ch;
        hi[j] = tx[j].o1 + ch;
    }
    const int L = min(min(lo[0], lo[1]), min(lo[2], lo[3]));
    uint32_t sel[4], wt[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        sel[j] = (uint32_t)(lo[j] - L) | ((uint32_t)(hi[j] - L) << 16) | 0x0c000c00u;
        wt[j] = ((uint32_t)tx[j].w0 << 4) | ((uint32_t)tx[j].w1 << 20);
    }
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, (int)a.lim_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, (int)a.lim_d, kBufWord3);
    const uint32_t ss = (uint32_t)a.s.stride, ds = (uint32_t)a.d.stride;
    const int ylast = a.s.h - 1;
    const uint32_t qx = a.off_s + (uint32_t)L;
    // Destination: interior strips store whole dwords from lanes 1..63 (lane
    // 0's offset is past the range, so the buffer drops its store); the first
    // and last strips take the per-lane rules.
    const bool edge = strip == 0 || X0 + 4 * 63 + 3 >= n;
    const uint32_t lane_off = lane == 0 ? 0x80000000u : 4u * (uint32_t)lane;
    const bool own = lane > 0 || strip == 0;
    const uint32_t qd0 = a.off_d + (uint32_t)X0;
    // A band's row taps and windows: row o1 of each output row, and row o0
    // where it is neither of the previous output row's (a downscale's
    // skipped row), all loads issued before the band is blended.
    struct Band {
        int ya[TB], yb[TB], wa[TB], wb[TB];
        vu32x3 sa[TB], sb[TB];
    };
    auto fetch = [&](Band &B, int band, int pa, int pb) {
        const int y0 = band * TB;
#pragma unroll
        for (int t = 0; t < TB; t++) {
            const LinTap ty = a.yt[min(y0 + t, dh - 1)];
            B.ya[t] = ty.o0;
            B.yb[t] = ty.o1;
            B.wa[t] = ty.w0;
            B.wb[t] = ty.w1;
        }
#pragma unroll
        for (int t = 0; t < TB; t++) {
            const int qa = t ? B.ya[t - 1] : pa, qb = t ? B.yb[t - 1] : pb;
            if (B.ya[t] != qa && B.ya[t] != qb && B.ya[t] != B.yb[t])
                B.sa[t] = load_window<AS>(rs, qx, (uint32_t)B.ya[t] * ss, B.ya[t] == ylast, a.lim_s);
            if (B.yb[t] != qb) B.sb[t] = load_window<AS>(rs, qx, (uint32_t)B.yb[t] * ss, B.yb[t] == ylast, a.lim_s);
        }
    };
    // H of the rows of the last output row done (pa, pb; -1 before the first).
    uint32_t ha[4] = {0u, 0u, 0u, 0u}, hb[4] = {0u, 0u, 0u, 0u};
    auto blend = [&](const Band &B, int band, int pa, int pb) {
        const int y0 = band * TB;
#pragma unroll
        for (int t = 0; t < TB; t++) {
            if (y0 + t >= dh) break;  // uniform
            // H of rows o0 and o1 from the previous row's (o1 - o0 is 0 or 1;
            // both never decrease): ha first, from the previous hb if o0 was
            // that row, then hb, then ha again if o0 is the new o1.
            const int qa = t ? B.ya[t - 1] : pa, qb = t ? B.yb[t - 1] : pb;
            const bool a_same = B.ya[t] == qa;
            const bool a_prev_b = !a_same && B.ya[t] == qb;
            const bool b_same = B.yb[t] == qb;
            const bool a_is_b = B.ya[t] == B.yb[t];
            if (a_prev_b) {
#pragma unroll
                for (int j = 0; j < 4; j++) ha[j] = hb[j];
            } else if (!a_same && !a_is_b) {
                h_row(window_bytes<AS>(B.sa[t], qx, (uint32_t)B.ya[t] * ss, B.ya[t] == ylast, a.lim_s), sel, wt, ha);
            }
            if (!b_same) h_row(window_bytes<AS>(B.sb[t], qx, (uint32_t)B.yb[t] * ss, B.yb[t] == ylast, a.lim_s), sel, wt, hb);
            if (!a_same && !a_prev_b && a_is_b) {
#pragma unroll
                for (int j = 0; j < 4; j++) ha[j] = hb[j];
            }
            // m0 + m1 + 2 <= 1022 per sample: pairs in 16-bit halves, >> 2
            // in both halves at once, the four low bytes picked by one permute.
            const uint32_t b0 = (uint32_t)B.wa[t] << 8, b1 = (uint32_t)B.wb[t] << 8;
            uint32_t m[4];
#pragma unroll
            for (int j = 0; j < 4; j++) m[j] = mulhi24(ha[j], b0) + mulhi24(hb[j], b1) + 2u;
            const u16x2 p01 = __builtin_bit_cast(u16x2, m[0] | (m[1] << 16)) >> (u16x2){2, 2};
            const u16x2 p23 = __builtin_bit_cast(u16x2, m[2] | (m[3] << 16)) >> (u16x2){2, 2};
            const uint32_t out =
                __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, p23), __builtin_bit_cast(uint32_t, p01), 0x06040200u);
            const uint32_t qd = qd0 + (uint32_t)(y0 + t) * ds;
            const int al = AD ? 0 : (int)(qd & 3u);  // uniform
            uint32_t dw = out;
            if constexpr (!AD) {
                const uint32_t left = __builtin_amdgcn_mov_dpp(out, kDppWaveShr1, 0xf, 0xf, true);
                dw = __builtin_amdgcn_perm(out, left, 0x07060504u - (uint32_t)al * 0x01010101u);
            }
            if (!edge) {
                __builtin_amdgcn_raw_buffer_store_b32(dw, rd, qd - (uint32_t)al + lane_off, 0, 0);
            } else {
                const int f = e0 - al;
                const uint32_t at = qd - (uint32_t)al + 4u * (uint32_t)lane;
                if (own && f >= 0 && f + 3 < n) {
                    __builtin_amdgcn_raw_buffer_store_b32(dw, rd, at, 0, 0);
                } else if (own && f + 3 >= 0 && f < n) {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (f + k >= 0 && f + k < n)
                            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(dw >> (8 * k)), rd, at + k, 0, 0);
                }
            }
        }
    };
    if constexpr (!RUN) {
        Band B;
        fetch(B, band0, -1, -1);
        blend(B, band0, -1, -1);
    } else {
        int pa = -1, pb = -1;
        for (int band = band0; band < band1; band++) {
            Band B;
            fetch(B, band, pa, pb);
            blend(B, band, pa, pb);
            pa = B.ya[TB - 1];
            pb = B.yb[TB - 1];
        }
    }
}

// One workgroup: 4 waves, each on a run of a.run bands (of TB rows) of one
// strip, the workgroup's 4 runs consecutive; tiles in XCD bands.
template <int TB, bool AD, bool AS, bool RUN>
__global__ __launch_bounds__(256) void linear_stream_u8(LinStreamArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    const int w = 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    lin_stream_wave<TB, AD, AS, RUN>(a, strip, w * a.run, (w + 1) * a.run);
}

}  // namespace lstream

}  // namespace visar
