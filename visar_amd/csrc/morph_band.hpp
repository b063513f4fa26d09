// morph_band.hpp -- one erode or dilate pass on one-channel u8 images for
// elements that are unions of centred rectangles up to 7 x 7, the rows
// streamed from memory into registers with the vertical window first, for
// gfx950 (the single passes of morphologyEx; morph_direct.hpp keeps the
// two-pass launches).
//
// Semantics as in morph_direct.hpp (reference src/imgproc/morph.zig:489-519,
// the subtraction of morphologyEx at :532-593): the min / max over the
// element is the min / max over its rectangles, each the min / max over a
// horizontal window of the column-wise min / max over its rows, and cells
// outside the image read as the edge-replicated image (exact for these
// elements: each rectangle that holds an outside cell also holds the
// in-image cell it replicates).
//
// The layout is the streamed stencils' (stencil_direct.hpp): each wave owns
// TB output rows of a 248-byte column strip, lane i holding bytes
// X0 + 4 (i - 1) .. + 3 of every row (one dword load per row, all issued
// before the first use); lanes 1..62 produce output and lanes 0 and 63 only
// lend their column results to their neighbours (DPP wave shifts). Per
// rectangle, taken by growing height, the column window grows from the
// previous one by its extra rows; the horizontal window then takes the
// neighbours' column results. The lane holding a row's first or last dword
// rebuilds the outer columns by replication. Requires rows of a multiple of
// 4 bytes (at least 8), each starting on a dword, in every image.
#pragma once

#include "morph_chain.hpp"
#include "stencil_direct.hpp"

namespace visar {

namespace mband {

using mchain::for_rects;
using mchain::hi_pair;
using mchain::lo_pair;
using mchain::pack2;
using mchain::RectOf;
using sdirect::dpp_shl;
using sdirect::dpp_shr;
using sdirect::kStrip;

constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access

template <bool MAX>
__device__ __attribute__((always_inline)) inline uint32_t pop2(uint32_t a, uint32_t b) {
    const u16x2 x = __builtin_bit_cast(u16x2, a), y = __builtin_bit_cast(u16x2, b);
    return __builtin_bit_cast(uint32_t, MAX ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
}

template <uint32_t EL>
constexpr int el_height() {
    int v = 0;
    for (int k = 0; k < (int)(EL & 15u); k++) v = (int)((EL >> (6 + 4 * k)) & 3u) > v ? (int)((EL >> (6 + 4 * k)) & 3u) : v;
    return v;
}

// Columns (C0, C1), (C2, C3) of the lane: the min / max over the lane's
// pairs lo / hi of the horizontal window [-A, A] (A <= 3), the neighbours'
// columns from the neighbouring lanes, replicated past the row's ends.
template <bool MAX, int A>
__device__ __attribute__((always_inline)) inline void h_window(uint32_t cl, uint32_t ch, bool e_left, bool e_right, bool edge,
                                                               uint32_t &o01, uint32_t &o23) {
    if constexpr (A == 0) {
        o01 = cl;
        o23 = ch;
    } else {
        uint32_t l1 = dpp_shr(ch), r1 = dpp_shl(cl), l2 = 0, r2 = 0;
        if constexpr (A == 3) {
            l2 = dpp_shr(cl);
            r2 = dpp_shl(ch);
        }
        if (edge) {
            if (e_left) l1 = l2 = __builtin_amdgcn_perm(ch, cl, 0x01000100u);   // (C0, C0)
            if (e_right) r1 = r2 = __builtin_amdgcn_perm(ch, cl, 0x07060706u);  // (C3, C3)
        }
        // P(d) = (C_d, C_d+1)
        const uint32_t am1 = __builtin_amdgcn_alignbyte(cl, l1, 2), a1 = __builtin_amdgcn_alignbyte(ch, cl, 2);
        const uint32_t a3 = __builtin_amdgcn_alignbyte(r1, ch, 2);
        // outputs (0, 1): P(-A) .. P(A); outputs (2, 3): P(2 - A) .. P(2 + A)
        o01 = pop2<MAX>(pop2<MAX>(am1, cl), a1);
        o23 = pop2<MAX>(pop2<MAX>(a1, ch), a3);
        if constexpr (A >= 2) {
            o01 = pop2<MAX>(pop2<MAX>(o01, l1), ch);
            o23 = pop2<MAX>(pop2<MAX>(o23, cl), r1);
        }
        if constexpr (A >= 3) {
            const uint32_t am3 = __builtin_amdgcn_alignbyte(l1, l2, 2), a5 = __builtin_amdgcn_alignbyte(r2, r1, 2);
            o01 = pop2<MAX>(pop2<MAX>(o01, am3), a3);
            o23 = pop2<MAX>(pop2<MAX>(o23, am1), a5);
        }
    }
}

// POST 1: out = saturate(C - v), 2: out = saturate(v - C).
template <uint32_t EL, bool MAX, int POST, int TB>
__device__ __attribute__((always_inline)) inline void band_wave(const MorphDirectArgs &a, int strip, int band) {
    constexpr int B = el_height<EL>(), ROWS = TB + 2 * B;
    const int lane = (int)threadIdx.x & 63;
    const int n = a.s.w, H = a.s.h;
    const int xc = strip * kStrip + 4 * (lane - 1);
    const int Y0 = band * TB;
    if (Y0 >= H) return;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, a.span_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, a.span_d, kBufWord3);
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc(a.c.p, 0, a.span_c, kBufWord3);
    const int ss = (int)a.s.stride, ds = (int)a.d.stride, cs = (int)a.c.stride;
    const int xl = min(max(xc, 0), n - 4);
    int so[ROWS];
    if (Y0 - B >= 0 && Y0 - B + ROWS <= H) {
#pragma unroll
        for (int r = 0; r < ROWS; r++) so[r] = (Y0 - B + r) * ss;
    } else {
#pragma unroll
        for (int r = 0; r < ROWS; r++) so[r] = min(max(Y0 - B + r, 0), H - 1) * ss;
    }
    uint32_t own[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) own[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, xl, so[r], 0);
    uint32_t lo[ROWS], hi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        lo[r] = __builtin_bit_cast(uint32_t, lo_pair(own[r]));
        hi[r] = __builtin_bit_cast(uint32_t, hi_pair(own[r]));
    }
    const bool e_left = xc == 0, e_right = xc == n - 4;
    const bool edge = __builtin_amdgcn_ballot_w64(e_left || e_right) != 0;  // uniform
    const bool live = lane >= 1 && lane <= 62 && xc < n;
    const uint32_t id = MAX ? 0u : 0x00ff00ffu;
#pragma unroll
    for (int t = 0; t < TB; t++) {
        const int y = Y0 + t;
        if (y >= H) break;  // uniform
        uint32_t vl = lo[t + B], vh = hi[t + B], r01 = id, r23 = id;
        int vprev = 0;
        for_rects<EL>([&](auto kc) {
            using R = RectOf<EL, decltype(kc)::value>;
#pragma unroll
            for (int e = 1; e <= 3; e++) {
                if (e > vprev && e <= R::v) {
                    vl = pop2<MAX>(pop2<MAX>(vl, lo[t + B - e]), lo[t + B + e]);
                    vh = pop2<MAX>(pop2<MAX>(vh, hi[t + B - e]), hi[t + B + e]);
                }
            }
            vprev = R::v > vprev ? R::v : vprev;
            uint32_t o01, o23;
            h_window<MAX, R::w>(vl, vh, e_left, e_right, edge, o01, o23);
            r01 = pop2<MAX>(r01, o01);
            r23 = pop2<MAX>(r23, o23);
        });
        u16x2 v[2] = {__builtin_bit_cast(u16x2, r01), __builtin_bit_cast(u16x2, r23)};
        if constexpr (POST != 0) {
            if (live) {
                const uint32_t cw = __builtin_amdgcn_raw_buffer_load_b32(rc, xc, y * cs, 0);
                const u16x2 c0 = lo_pair(cw), c1 = hi_pair(cw);
                v[0] = POST == 1 ? __builtin_elementwise_sub_sat(c0, v[0]) : __builtin_elementwise_sub_sat(v[0], c0);
                v[1] = POST == 1 ? __builtin_elementwise_sub_sat(c1, v[1]) : __builtin_elementwise_sub_sat(v[1], c1);
            }
        }
        if (live) __builtin_amdgcn_raw_buffer_store_b32(pack2(v[0], v[1]), rd, xc, y * ds, 0);
    }
}

// One workgroup: 4 waves on 4 consecutive bands of one strip, tiles in XCD
// bands. a.bands4: groups of 4 bands of TB rows.
template <uint32_t EL, bool MAX, int POST, int TB>
__global__ __launch_bounds__(256) void morph_band_u8(MorphDirectArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    band_wave<EL, MAX, POST, TB>(a, strip, 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

}  // namespace mband

}  // namespace visar
