// stencil_stream.hpp -- the u8 GaussianBlur 3x3, 5x5, 7x7 with sigma 0 on
// one- and three-channel images of any row alignment and width, with the
// rows streamed from memory into registers and no LDS, for gfx950 (the
// general form of stencil_direct.hpp, whose kernel needs one-channel rows of
// a multiple of 4 bytes starting on dwords).
//
// The sums are the reference's exact integer ones (filter.zig:120-341 with
// the taps of :350-483), taken vertical taps first as in stencil_direct.hpp;
// every partial sum is exact, so the bytes equal those of the other kernels.
//
// Each wave owns TB output rows of a strip of 240 pixels: lane i holds
// pixels P0 + 4 (i - 2) .. + 3 of every row (4 CN bytes), lanes 2..61
// produce output and lanes 0, 1, 62, 63 lend their column sums (and, for
// the stores, their outputs) to their neighbours. The border acts on the
// source: a lane whose pixels are not all inside the row (or whose window
// would pass the row's end) loads each byte from its border-mapped column,
// so column sums of outside columns are those the border rule gives and the
// horizontal pass needs no edge cases. Other lanes load their bytes with one
// 8- or 16-byte load from the dword at or below them, realigned with
// alignbyte. Rows outside the image are border-mapped on the scalar unit
// (zero under a constant border of value 0).
//
// Three channels: the vertical pass works on the 12 bytes as 16-bit pairs;
// the column sums are then split into one pair set per channel (pixels 0-1
// and 2-3 of the lane), which the horizontal pass treats as three one-
// channel images; the outputs are interleaved back into 12 bytes.
//
// Stores: with a = (address of the strip's first output byte) & 3, lane i
// stores the dwords holding its bytes shifted down by a, the low a bytes
// from lane i - 1 (a DPP shift of its last dword); dwords crossing the row's
// ends are written byte by byte.
#pragma once

#include "morph_band.hpp"
#include "stencil_direct.hpp"

namespace visar {

struct SStreamArgs {
    StencilArgs s;            // src / dst with p rounded down to a dword
    uint32_t off_s, off_d;    // the bytes skipped
    uint32_t lim_s, lim_d;    // bytes from each rounded-down base to the image's last byte + 1 (< 2^31)
    int strips, ntiles, xcd;  // the tile order: tile (strip, group of 4 bands)
    int bands;                // output row bands of TB rows
};

// The streamed kernel's arguments for bands of tb rows (all but the tile
// order); false for other channel counts, mismatched images or byte spans
// beyond 32-bit offsets.
bool stencil_stream_args(const StencilArgs &p, int tb, SStreamArgs &a);

namespace sstream {

using sdirect::dpp_shl;
using sdirect::dpp_shr;
using sdirect::HTap;
using sdirect::padd;
using sdirect::pdot;
using sdirect::perm;
using sdirect::VTap;

constexpr int kStripPx = 240;          // output pixels per strip (lanes 2..61)
constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int vu32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int vu32x4 __attribute__((ext_vector_type(4)));

// One output row's four samples of one plane from the plane's column-sum
// pairs cl = (C0, C1), ch = (C2, C3) and its neighbours' (stencil_direct.hpp).
template <int MODE>
__device__ __attribute__((always_inline)) inline uint32_t h_pass(uint32_t cl, uint32_t ch) {
    constexpr int K = SwarTaps<MODE>::K;
    uint32_t l1 = dpp_shr(ch), r1 = dpp_shl(cl), l2 = 0, r2 = 0;
    if constexpr (K == 7) {
        l2 = dpp_shr(cl);
        r2 = dpp_shl(ch);
    }
    const uint32_t am1 = __builtin_amdgcn_alignbyte(cl, l1, 2);  // (C-1, C0)
    const uint32_t a1 = __builtin_amdgcn_alignbyte(ch, cl, 2);   // (C1, C2)
    const uint32_t a3 = __builtin_amdgcn_alignbyte(r1, ch, 2);   // (C3, C4)
    if constexpr (MODE == SWAR_G7) {
        const uint32_t am3 = __builtin_amdgcn_alignbyte(l1, l2, 2);  // (C-3, C-2)
        const uint32_t a5 = __builtin_amdgcn_alignbyte(r2, r1, 2);   // (C5, C6)
        uint32_t f[4];
#pragma unroll
        for (int q = 0; q < 2; q++) {
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
        return f[0] | (f[1] << 8) | (f[2] << 16) | (f[3] << 24);
    } else {
        uint32_t o01, o23;
        if constexpr (K == 3) {
            const uint32_t x01[3] = {am1, cl, a1}, x23[3] = {a1, ch, a3};
            o01 = pdot<HTap<MODE>, 3>(x01, 0u, true);
            o23 = pdot<HTap<MODE>, 3>(x23, 0u, true);
        } else {
            const uint32_t x01[5] = {l1, am1, cl, a1, ch}, x23[5] = {cl, a1, ch, a3, r1};
            o01 = pdot<HTap<MODE>, 5>(x01, 0u, true);
            o23 = pdot<HTap<MODE>, 5>(x23, 0u, true);
        }
        constexpr uint16_t add = MODE == SWAR_G3 ? 8 : 128, sh = MODE == SWAR_G3 ? 4 : 8;
        const u16x2 f01 = (__builtin_bit_cast(u16x2, o01) + add) >> sh;
        const u16x2 f23 = (__builtin_bit_cast(u16x2, o23) + add) >> sh;
        return perm(__builtin_bit_cast(uint32_t, f23), __builtin_bit_cast(uint32_t, f01), 0x06040200u);
    }
}

// The reductions a streamed wave runs: the vertical one on K rows of 16-bit
// pairs, the horizontal one on a plane's column results (cl, ch) giving the
// lane's four output bytes, and the border the source is read with.
template <int MODE>
struct GaussOp {
    static constexpr int K = SwarTaps<MODE>::K;
    __device__ static int border(int b) { return b; }
    __device__ static uint32_t vert(const uint32_t (&xs)[K]) { return pdot<VTap<MODE>, K>(xs, 0u, true); }
    __device__ static uint32_t horiz(uint32_t cl, uint32_t ch) { return h_pass<MODE>(cl, ch); }
};
// erode (MAX false) / dilate with a K x K rectangle: morph.zig:489-519 with
// cells outside the image ignored, which for a rectangle is the min / max
// over the edge-replicated image (every rectangle holding an outside cell
// holds the in-image cell it replicates).
template <int KK, bool MAX>
struct RectOp {
    static constexpr int K = KK;
    __device__ static int border(int) { return BORDER_REPLICATE; }
    __device__ static uint32_t vert(const uint32_t (&xs)[K]) {
        uint32_t v = xs[0];
#pragma unroll
        for (int k = 1; k < K; k++) v = mband::pop2<MAX>(v, xs[k]);
        return v;
    }
    __device__ static uint32_t horiz(uint32_t cl, uint32_t ch) {
        uint32_t o01, o23;
        mband::h_window<MAX, K / 2>(cl, ch, false, false, false, o01, o23);
        return perm(o23, o01, 0x06040200u);
    }
};

// AD: every destination row starts on a dword (stores need no shift).
template <class OP, int CN, int TB, bool AD>
__device__ __attribute__((always_inline)) inline void stream_wave(const SStreamArgs &a, int strip, int band) {
    constexpr int K = OP::K, R = K / 2, ROWS = TB + 2 * R, NB = 4 * CN;
    const StencilArgs &p = a.s;
    const int lane = (int)threadIdx.x & 63;
    const int w = p.src.w, H = p.src.h, border = OP::border(p.border);
    const int Y0 = band * TB;
    if (Y0 >= H) return;  // uniform
    const int P0 = strip * kStripPx;
    const int px0 = P0 + 4 * (lane - 2);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(p.src.p, 0, (int)a.lim_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(p.dst.p, 0, (int)a.lim_d, kBufWord3);
    const uint32_t ss = (uint32_t)p.src.stride, ds = (uint32_t)p.dst.stride;
    // Fast lanes: all 4 pixels inside the row and at least 4 more bytes of
    // the row after them (the load window stays inside the range).
    const bool fast = px0 >= 0 && (px0 + 4) * CN + 4 <= w * CN;
    // Slow lanes: each byte's border-mapped column (-1: the constant 0).
    int col[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int m = border_map_halo(px0 + b / CN, w, border);
        col[b] = m < 0 ? -1 : m * CN + b % CN;
    }
    // Row offsets (scalar), then every load back to back.
    int ro[ROWS];
    uint32_t zero_rows = 0;  // rows outside the image under a constant border (uniform)
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const int y = border_map_halo(Y0 - R + r, H, border);
        if (y < 0) zero_rows |= 1u << r;
        ro[r] = (int)a.off_s + max(y, 0) * (int)ss;
    }
    // Every lane's window of every row first, with no branch (a lane whose
    // window would leave the image reads its row's first dwords instead),
    // so all of them are in flight together; then, in waves holding edge
    // lanes only, the edge lanes' border-mapped bytes, a few rows at a time.
    constexpr int WD = CN == 1 ? 2 : 4;  // window dwords
    uint32_t win[ROWS][WD];
    const uint32_t lane_b = (uint32_t)(px0 * CN);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const uint32_t q = fast ? ((uint32_t)ro[r] + lane_b) & ~3u : (uint32_t)ro[r] & ~3u;
        if constexpr (CN == 1) {
            const vu32x2 v = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, q, 0, 0));
            win[r][0] = v.x;
            win[r][1] = v.y;
        } else {
            const vu32x4 v = __builtin_bit_cast(vu32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, q, 0, 0));
            win[r][0] = v.x;
            win[r][1] = v.y;
            win[r][2] = v.z;
            win[r][3] = v.w;
        }
    }
    uint32_t raw[ROWS][CN];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const uint32_t sh = (uint32_t)ro[r] + lane_b;  // alignbyte takes its low 2 bits
#pragma unroll
        for (int k = 0; k < CN; k++) raw[r][k] = __builtin_amdgcn_alignbyte(win[r][k + 1], win[r][k], sh);
    }
    if (__builtin_amdgcn_ballot_w64(!fast) != 0) {  // uniform
        constexpr int CH = CN == 1 ? ROWS : 2;      // rows of byte loads in flight together
#pragma unroll
        for (int r0 = 0; r0 < ROWS; r0 += CH) {
            uint32_t by[CH][NB];
#pragma unroll
            for (int r = r0; r < r0 + CH && r < ROWS; r++) {
#pragma unroll
                for (int b = 0; b < NB; b++)
                    by[r - r0][b] = __builtin_amdgcn_raw_buffer_load_b8(rs, (uint32_t)(ro[r] + max(col[b], 0)), 0, 0);
            }
#pragma unroll
            for (int r = r0; r < r0 + CH && r < ROWS; r++) {
#pragma unroll
                for (int k = 0; k < CN; k++) {
                    uint32_t v = 0;
#pragma unroll
                    for (int b = 4 * k; b < 4 * k + 4; b++) v |= (col[b] < 0 ? 0u : by[r - r0][b] & 0xffu) << (8 * (b % 4));
                    if (!fast) raw[r][k] = v;
                }
            }
        }
    }
    // Samples as 16-bit pairs: pair j of a row holds bytes 2j, 2j + 1.
    uint32_t pr[ROWS][2 * CN];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
#pragma unroll
        for (int k = 0; k < CN; k++) {
            const uint32_t v = (zero_rows >> r) & 1 ? 0u : raw[r][k];
            pr[r][2 * k] = perm(0u, v, 0x0c010c00u);
            pr[r][2 * k + 1] = perm(0u, v, 0x0c030c02u);
        }
    }
    const bool produce = lane >= 2 && lane <= 61;
    const int nb = w * CN;  // bytes per row
#pragma unroll
    for (int t = 0; t < TB; t++) {
        const int y = Y0 + t;
        if (y >= H) break;  // uniform
        // Column sums of the lane's bytes, pair by pair.
        uint32_t cs[2 * CN];
#pragma unroll
        for (int j = 0; j < 2 * CN; j++) {
            uint32_t xs[K];
#pragma unroll
            for (int k = 0; k < K; k++) xs[k] = pr[t + k][j];
            cs[j] = OP::vert(xs);
        }
        uint32_t out[CN];
        if constexpr (CN == 1) {
            out[0] = OP::horiz(cs[0], cs[1]);
        } else {
            // Per channel c: (S[c], S[c + 3]) and (S[c + 6], S[c + 9]), S[i]
            // the 16-bit sum of byte i (pair i / 2, half i % 2).
            const uint32_t c0l = perm(cs[1], cs[0], 0x07060100u), c0h = perm(cs[4], cs[3], 0x07060100u);
            const uint32_t c1l = perm(cs[2], cs[0], 0x05040302u), c1h = perm(cs[5], cs[3], 0x05040302u);
            const uint32_t c2l = perm(cs[2], cs[1], 0x07060100u), c2h = perm(cs[5], cs[4], 0x07060100u);
            const uint32_t o0 = OP::horiz(c0l, c0h), o1 = OP::horiz(c1l, c1h), o2 = OP::horiz(c2l, c2h);
            // Pixels 0-3 of channels 0-2 (oc.k) back to byte order:
            // o0.0 o1.0 o2.0 o0.1 | o1.1 o2.1 o0.2 o1.2 | o2.2 o0.3 o1.3 o2.3
            out[0] = perm(o2, perm(o1, o0, 0x01000400u), 0x03040100u);
            out[1] = perm(o2, perm(o1, o0, 0x06020005u), 0x03020500u);
            out[2] = perm(o2, perm(o1, o0, 0x00070300u), 0x07020106u);
        }
        // The dwords of the lane's bytes shifted down by al, the low al bytes
        // from lane i - 1; row bytes f .. f + NB - 1.
        const uint32_t qd = a.off_d + (uint32_t)y * ds + (uint32_t)(P0 * CN);
        const int al = AD ? 0 : (int)(qd & 3u);  // uniform
        uint32_t dw[CN];
        if constexpr (AD) {
#pragma unroll
            for (int k = 0; k < CN; k++) dw[k] = out[k];
        } else {
            const uint32_t left = __builtin_amdgcn_mov_dpp(out[CN - 1], 0x138, 0xf, 0xf, true);
            const uint32_t sel = 0x07060504u - (uint32_t)al * 0x01010101u;
            dw[0] = perm(out[0], left, sel);
#pragma unroll
            for (int k = 1; k < CN; k++) dw[k] = perm(out[k], out[k - 1], sel);
        }
        const int f = px0 * CN - al;
        const uint32_t at = qd + (uint32_t)(4 * (lane - 2) * CN) - (uint32_t)al;
        if (produce && f >= 0 && f + NB <= nb) {
            if constexpr (CN == 1) __builtin_amdgcn_raw_buffer_store_b32(dw[0], rd, at, 0, 0);
            else __builtin_amdgcn_raw_buffer_store_b96(vu32x3{dw[0], dw[1], dw[2]}, rd, at, 0, 0);
        } else if (produce && f + NB > 0 && f < nb) {
#pragma unroll
            for (int k = 0; k < NB; k++)
                if (f + k >= 0 && f + k < nb) __builtin_amdgcn_raw_buffer_store_b8((uint8_t)(dw[k / 4] >> (8 * (k % 4))), rd, at + k, 0, 0);
        }
    }
}

// One workgroup: 4 waves on 4 consecutive bands of one strip, tiles in XCD
// bands.
template <class OP, int CN, int TB, bool AD>
__global__ __launch_bounds__(256) void sep_stream_u8(SStreamArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    stream_wave<OP, CN, TB, AD>(a, strip, 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

}  // namespace sstream

}  // namespace visar
