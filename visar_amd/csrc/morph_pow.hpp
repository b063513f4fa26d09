// morph_pow.hpp -- n iterated erode or dilate passes (n = 2, 3) of a union of
// centred rectangles on one-channel u8 images as one pass with the n-fold
// Minkowski sum of the element, column window first, for gfx950: the form
// of morph_band.hpp for sums up to half-width 11 and half-height 9 (the
// powers whose rectangles are too many for morph_stream.hpp's accumulators).
//
// Semantics: morph.zig:63-96 (iterations repeat the pass, each ignoring
// cells outside the image) and :489-519. For a union of centred rectangles
// with outside cells ignored, n passes equal one pass with the n-fold sum
// (morph_stream.hpp), and ignoring outside cells equals reading the
// edge-replicated image (each rectangle holding an outside cell holds the
// in-image cell it replicates).
//
// Layout as in morph_band.hpp, with HL = 1, 2 or 3 lanes of halo on each side:
// lane i holds bytes X0 + 4 (i - HL) .. + 3 of every row (one dword load per
// row, all issued before the first use); lanes HL..63-HL produce output.
// Lanes wholly left or right of the row hold its first or last byte in all
// four columns (the replication, applied to the source), so the horizontal
// windows take their neighbours' columns by DPP with no edge cases; rows
// outside the image are clamped. Per output row the column window grows
// through the sum's rectangles by height, and each rectangle's horizontal
// window takes the neighbouring columns. Requires rows of a multiple of 4
// bytes (at least 8), each starting on a dword, in every image.
#pragma once

#include "morph_band.hpp"

namespace visar {

namespace mpow {

using mband::kBufWord3;
using mband::pop2;
using mstream::RectSet;

template <uint32_t BASE, int N>
struct Pow {
    static constexpr RectSet R = mstream::power(BASE, N);
    static constexpr int A = mstream::max_a(R), B = mstream::max_b(R);
};
// Launches exist for the sums with A <= 11, B <= 9 that the accumulator
// kernel does not take.
template <uint32_t BASE, int N>
constexpr bool fits() {
    return Pow<BASE, N>::A <= 11 && Pow<BASE, N>::B <= 9 && mstream::acc_rows(Pow<BASE, N>::R) > mstream::kStreamMaxRows;
}
// Lanes of halo on each side: 4 HL bytes cover the half-width.
constexpr int halo_lanes(int a) { return (a + 3) / 4; }

template <int... I, typename F>
__device__ __attribute__((always_inline)) inline void for_seq(std::integer_sequence<int, I...>, F &&f) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Output pairs (C0, C1), (C2, C3) of the min / max over [-A, A] columns,
// from the even pairs P(e) = (C_e, C_e+1): the lane's own (lo = (C0, C1),
// hi = (C2, C3)) and its neighbours' (lanes i -+ m in ll, lh / rl, rh).
template <bool MAX, int A, int HL>
__device__ __attribute__((always_inline)) inline void h_win(uint32_t lo, uint32_t hi, const uint32_t (&ll)[HL], const uint32_t (&lh)[HL],
                                                            const uint32_t (&rl)[HL], const uint32_t (&rh)[HL], uint32_t &o01,
                                                            uint32_t &o23) {
    auto even = [&](int d) -> uint32_t {  // d even, -4 HL <= d <= 4 HL + 2 (A + 3 <= 4 HL + 2)
        if (d >= 0 && d <= 2) return d == 0 ? lo : hi;
        if (d < 0) {
            const int m = (-d + 3) / 4, r = d + 4 * m;  // lane i - m, its pair at r (0 or 2)
            return r == 0 ? ll[m - 1] : lh[m - 1];
        }
        const int m = d / 4, r = d - 4 * m;
        return r == 0 ? rl[m - 1] : rh[m - 1];
    };
    if constexpr (A == 0) {
        o01 = lo;
        o23 = hi;
    } else {
        // W(x): the pairwise min / max of the even pairs P(e), e in [x, x + 2A]:
        // its low halves cover the even columns of [x, x + 2A], its high
        // halves the odd columns of [x + 1, x + 2A + 1]. Each output half is
        // one window's even and odd columns, so two W's and the halves of a
        // third give it, with no odd pair assembled.
        auto W = [&](int x) -> uint32_t {
            const int e0 = x + (x & 1), e1 = x + 2 * A - ((x + 2 * A) & 1);
            uint32_t w = even(e0);
#pragma unroll
            for (int e = e0 + 2; e <= e1; e += 2) w = pop2<MAX>(w, even(e));
            return w;
        };
        const uint32_t w1 = W(-A), w2 = W(-A - 1), w3 = W(1 - A), w5 = W(2 - A), w7 = W(3 - A);
        // o01 = (C[-A..A], C[1-A..A+1]) = W(-A) with (W(-A-1)'s high half, W(1-A)'s low half).
        o01 = pop2<MAX>(w1, __builtin_amdgcn_alignbyte(w3, w2, 2));
        // o23 = (C[2-A..A+2], C[3-A..A+3]) = W(2-A) with (W(1-A)'s high half, W(3-A)'s low half).
        o23 = pop2<MAX>(w5, __builtin_amdgcn_alignbyte(w7, w3, 2));
    }
}

// POST 1: out = saturate(C - v), 2: out = saturate(v - C).
template <uint32_t BASE, int N, bool MAX, int POST, int TB>
__device__ __attribute__((always_inline)) inline void pow_wave(const MorphDirectArgs &a, int strip, int band) {
    using PW = Pow<BASE, N>;
    constexpr int A = PW::A, B = PW::B, ROWS = TB + 2 * B;
    constexpr int HL = halo_lanes(A);
    static_assert(HL >= 1 && HL <= 3, "half-width outside the kernel's limits");
    constexpr int SW = 4 * (64 - 2 * HL);  // output bytes per strip
    const int lane = (int)threadIdx.x & 63;
    const int n = a.s.w, H = a.s.h;
    const int xc = strip * SW + 4 * (lane - HL);
    const int Y0 = band * TB;
    if (Y0 >= H) return;  // uniform
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, a.span_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, a.span_d, kBufWord3);
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc(a.c.p, 0, a.span_c, kBufWord3);
    const int ss = (int)a.s.stride, ds = (int)a.d.stride, cs = (int)a.c.stride;
    const int xl = min(max(xc, 0), n - 4);
    int so[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) so[r] = min(max(Y0 - B + r, 0), H - 1) * ss;
    uint32_t own[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) own[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, xl, so[r], 0);
    // Lanes left of the row repeat its first byte, lanes right of it its last.
    const uint32_t rep = xc < 0 ? 0x00000000u : 0x03030303u;
    const bool outside = xc < 0 || xc >= n;
    uint32_t lo[ROWS], hi[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const uint32_t v = outside ? __builtin_amdgcn_perm(0u, own[r], rep) : own[r];
        lo[r] = __builtin_amdgcn_perm(0u, v, 0x0c010c00u);
        hi[r] = __builtin_amdgcn_perm(0u, v, 0x0c030c02u);
    }
    const bool live = lane >= HL && lane < 64 - HL && xc < n;
    const uint32_t id = MAX ? 0u : 0x00ff00ffu;
#pragma unroll
    for (int t = 0; t < TB; t++) {
        const int y = Y0 + t;
        if (y >= H) break;  // uniform
        uint32_t vl = lo[t + B], vh = hi[t + B], r01 = id, r23 = id;
        for_seq(std::make_integer_sequence<int, PW::R.n>{}, [&](auto kc) {
            constexpr int K = decltype(kc)::value;
            constexpr int bk = PW::R.b[K], ak = PW::R.a[K], bp = K == 0 ? 0 : PW::R.b[K == 0 ? 0 : K - 1];
#pragma unroll
            for (int e = bp + 1; e <= bk; e++) {
                vl = pop2<MAX>(pop2<MAX>(vl, lo[t + B - e]), lo[t + B + e]);
                vh = pop2<MAX>(pop2<MAX>(vh, hi[t + B - e]), hi[t + B + e]);
            }
            uint32_t ll[HL], lh[HL], rl[HL], rh[HL];
            ll[0] = mband::dpp_shr(vl);
            lh[0] = mband::dpp_shr(vh);
            rl[0] = mband::dpp_shl(vl);
            rh[0] = mband::dpp_shl(vh);
#pragma unroll
            for (int m = 1; m < HL; m++) {
                ll[m] = mband::dpp_shr(ll[m - 1]);
                lh[m] = mband::dpp_shr(lh[m - 1]);
                rl[m] = mband::dpp_shl(rl[m - 1]);
                rh[m] = mband::dpp_shl(rh[m - 1]);
            }
            uint32_t o01, o23;
            h_win<MAX, ak, HL>(vl, vh, ll, lh, rl, rh, o01, o23);
            r01 = pop2<MAX>(r01, o01);
            r23 = pop2<MAX>(r23, o23);
        });
        u16x2 v[2] = {__builtin_bit_cast(u16x2, r01), __builtin_bit_cast(u16x2, r23)};
        if constexpr (POST != 0) {
            if (live) {
                const uint32_t cw = __builtin_amdgcn_raw_buffer_load_b32(rc, xc, y * cs, 0);
                const u16x2 c0 = mchain::lo_pair(cw), c1 = mchain::hi_pair(cw);
                v[0] = POST == 1 ? __builtin_elementwise_sub_sat(c0, v[0]) : __builtin_elementwise_sub_sat(v[0], c0);
                v[1] = POST == 1 ? __builtin_elementwise_sub_sat(c1, v[1]) : __builtin_elementwise_sub_sat(v[1], c1);
            }
        }
        if (live) __builtin_amdgcn_raw_buffer_store_b32(mchain::pack2(v[0], v[1]), rd, xc, y * ds, 0);
    }
}

// One workgroup: 4 waves on 4 consecutive bands of one strip, tiles in XCD
// bands.
template <uint32_t BASE, int N, bool MAX, int POST, int TB>
__global__ __launch_bounds__(256) void morph_pow_u8(MorphDirectArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    pow_wave<BASE, N, MAX, POST, TB>(a, strip, 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

}  // namespace mpow

}  // namespace visar
