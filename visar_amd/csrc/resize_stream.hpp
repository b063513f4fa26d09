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
// Layout: each wave owns TB output rows of a strip of output samples, lane i
// holding samples X0 + 4i .. + 3; strips advance by 252 samples, so lane 63
// recomputes the next strip's first four samples. The source bytes of a
// lane's four samples' taps lie within 8 consecutive bytes of a row (checked
// on the host): each source row the band needs is one 12-byte load per lane
// at the dword at or below them, realigned with two alignbytes, and each
// sample's two bytes are gathered with one permute. All of a band's loads
// are issued before the first use. Rows are visited in order; an output row
// is blended as soon as its second tap row's H is formed, its first tap row
// being that row or the previous one (o1 - o0 is 0 or 1).
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
typedef unsigned int vu32x3 __attribute__((ext_vector_type(3)));

// (h * b) >> 16 for h = H << 8, b = B << 8 below 2^24: v_mul_hi_u32_u24.
__device__ __attribute__((always_inline)) inline uint32_t mulhi24(uint32_t h, uint32_t b) {
    return (uint32_t)(((uint64_t)(h & 0xffffffu) * (uint64_t)(b & 0xffffffu)) >> 32);
}

template <int TB>
__device__ __attribute__((always_inline)) inline void lin_stream_wave(const LinStreamArgs &a, int strip, int band) {
    const int lane = (int)threadIdx.x & 63;
    const int cn = a.d.cn, n = a.d.w * cn;
    const int y0 = band * TB;
    if (y0 >= a.d.h) return;  // uniform
    const int rows = min(TB, a.d.h - y0);
    const int X0 = strip * kLinStreamStrip;
    const int e0 = X0 + 4 * lane;
    // This lane's column taps, and row y0 + lane's taps in lanes 0..TB-1.
    const gptr<const LinTap> xg = gp(a.xt);
    LinTap tx[4];
#pragma unroll
    for (int j = 0; j < 4; j++) tx[j] = load_tap(xg, div_cn(min(e0 + j, n - 1), cn));
    const LinTap ty = load_tap(gp(a.yt), y0 + min(lane, rows - 1));
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
    const int ylo = __builtin_amdgcn_readlane(ty.o0, 0);
    const int nr = __builtin_amdgcn_readlane(ty.o1, rows - 1) - ylo + 1;  // <= kLinStreamRows
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, (int)a.lim_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, (int)a.lim_d, kBufWord3);
    const uint32_t ss = (uint32_t)a.s.stride, ds = (uint32_t)a.d.stride;
    // A load that crosses the end of the buffer's range returns zeros in
    // every dword, so on the image's last row a window that would cross it
    // starts lower, at lim_s - 12 (lim_s is a multiple of 4; rows of at
    // least 12 bytes keep every other row's windows inside the range), and
    // the lane's bytes are picked from the dword they start in.
    const int ylast = a.s.h - 1;
    uint32_t q[kLinStreamRows];
    vu32x3 src[kLinStreamRows];
#pragma unroll
    for (int r = 0; r < kLinStreamRows; r++) {
        q[r] = a.off_s + (uint32_t)(ylo + min(r, nr - 1)) * ss + (uint32_t)L;
        if (r < nr) {
            const uint32_t at = ylo + r == ylast ? min(q[r] & ~3u, a.lim_s - 12u) : q[r] & ~3u;
            src[r] = __builtin_bit_cast(vu32x3, __builtin_amdgcn_raw_buffer_load_b96(rs, at, 0, 0));
        }
    }
    // Destination: the dword this lane stores holds samples f .. f + 3.
    const uint32_t qd0 = a.off_d + (uint32_t)X0;
    const bool own = lane > 0 || strip == 0;
    uint32_t hp[4] = {0u, 0u, 0u, 0u};
    int t = 0;
#pragma unroll
    for (int r = 0; r < kLinStreamRows; r++) {
        if (r < nr) {  // uniform
            uint32_t w0 = __builtin_amdgcn_alignbyte(src[r].y, src[r].x, q[r]);
            uint32_t w1 = __builtin_amdgcn_alignbyte(src[r].z, src[r].y, q[r]);
            if (ylo + r == ylast) {  // uniform
                const uint32_t dl = q[r] - min(q[r] & ~3u, a.lim_s - 12u);  // 0..11
                const uint32_t d0 = dl >= 8 ? src[r].z : dl >= 4 ? src[r].y : src[r].x;
                const uint32_t d1 = dl >= 8 ? 0u : dl >= 4 ? src[r].z : src[r].y;
                const uint32_t d2 = dl >= 4 ? 0u : src[r].z;
                w0 = __builtin_amdgcn_alignbyte(d1, d0, dl);
                w1 = __builtin_amdgcn_alignbyte(d2, d1, dl);
            }
            uint32_t hc[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const u16x2 pr = __builtin_bit_cast(u16x2, __builtin_amdgcn_perm(w1, w0, sel[j]));
                hc[j] = __builtin_amdgcn_udot2(pr, __builtin_bit_cast(u16x2, wt[j]), 0u, false) & 0x00ffff00u;
            }
            // Output rows whose second tap row is this one (uniform loop).
            while (t < rows && __builtin_amdgcn_readlane(ty.o1, t) == ylo + r) {
                const bool same = __builtin_amdgcn_readlane(ty.o0, t) == ylo + r;
                const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane(ty.w0, t) << 8;
                const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane(ty.w1, t) << 8;
                uint32_t out = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t m = mulhi24(same ? hc[j] : hp[j], b0) + mulhi24(hc[j], b1) + 2u;
                    out |= (m >> 2) << (8 * j);
                }
                const uint32_t left = __builtin_amdgcn_update_dpp(0u, out, kDppWaveShr1, 0xf, 0xf, false);
                const uint32_t qd = qd0 + (uint32_t)(y0 + t) * ds;
                const int al = (int)(qd & 3u);  // uniform
                const uint32_t dw = __builtin_amdgcn_perm(out, left, 0x07060504u - (uint32_t)al * 0x01010101u);
                const int f = e0 - al;
                const uint32_t at = qd - (uint32_t)al + 4u * (uint32_t)lane;
                if (own && f >= 0 && f + 3 < n) {
                    __builtin_amdgcn_raw_buffer_store_b32(dw, rd, at, 0, 0);
                } else if (own && f + 3 >= 0 && f < n) {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (f + k >= 0 && f + k < n) __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(dw >> (8 * k)), rd, at + k, 0, 0);
                }
                t++;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) hp[j] = hc[j];
        }
    }
}

// One workgroup: wpg waves on consecutive bands of one strip, tiles in XCD
// bands.
template <int TB>
__global__ __launch_bounds__(256) void linear_stream_u8(LinStreamArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    lin_stream_wave<TB>(a, strip, a.wpg * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

}  // namespace lstream

}  // namespace visar
