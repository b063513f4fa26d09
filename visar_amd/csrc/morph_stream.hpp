// morph_stream.hpp -- one erode, dilate or gradient pass with a large
// element on one-channel u8 images, the rows streamed from memory into
// registers, for gfx950: the n-th Minkowski power of an element that is a
// union of centred rectangles, which takes n iterated passes in one.
//
// Reference semantics: src/imgproc/morph.zig:63-96 (iterations repeat the
// pass; every pass ignores cells outside the image) and :489-519 (one pass).
// For an element B that is a union of centred rectangles R_k and cells
// outside the image ignored, two passes equal one pass with B + B = the
// union of the centred rectangles R_j + R_k: output x reads z through some
// in-image y with y - x in R_j and z - y in R_k exactly when z - x lies in
// R_j + R_k (choose y coordinate by coordinate between x and z, which keeps
// it inside the image, a box), so n passes are one pass with the n-fold sum.
// Gradient (morph.zig:550-554) is the dilation minus the erosion of the
// same source, both taken here from the same loaded rows.
//
// Each wave owns a band of `tb` output rows of a column strip. Lane i holds
// bytes X0 + 8 (i - HL) .. + 7 of every row (one 8-byte load); the HL lanes
// at each end only lend their bytes to their neighbours (DPP wave shifts),
// which gives every output lane up to 12 bytes of halo on each side. Lanes
// past either end of a row hold the row's edge byte, so every lane's halo
// is the edge-replicated row (the same result as ignoring outside cells for
// these elements: each rectangle that holds an outside cell also holds the
// in-image cell it replicates). Rows outside the image are clamped likewise.
// The horizontal windows grow one byte per level in 16-bit sample pairs;
// each input row adds its level-a_k window to the rolling accumulators of
// the output rows within b_k of it, as in morph_direct.hpp. The row loop
// runs in chunks of 2B + 1 rows, so accumulator and prefetch slots are
// compile-time indices. Requires rows of a multiple of 4 bytes (at least 8),
// each starting on a dword, in every image.
#pragma once

#include <utility>

#include "morph_chain.hpp"

namespace visar {

namespace mstream {

using namespace mchain;

template <uint32_t BASE, int N>
struct Pow {
    static constexpr RectSet R = power(BASE, N);
    static constexpr int A = max_a(R), B = max_b(R);
};

constexpr int kDppShr1 = 0x138, kDppShl1 = 0x130;  // lane i takes lane i - 1's (shr) or i + 1's (shl) value
constexpr int kBufWord3 = 0x00020000;              // buffer resource word 3 for gfx9 raw (untyped) access
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));

// The lane without a source reads 0 (bound_ctrl): no old value to set up.
__device__ __attribute__((always_inline)) inline uint32_t shr1(uint32_t v) {
    return __builtin_amdgcn_mov_dpp(v, kDppShr1, 0xf, 0xf, true);
}
__device__ __attribute__((always_inline)) inline uint32_t shl1(uint32_t v) {
    return __builtin_amdgcn_mov_dpp(v, kDppShl1, 0xf, 0xf, true);
}

template <int... I, typename F>
__device__ __attribute__((always_inline)) inline void for_seq(std::integer_sequence<int, I...>, F &&f) {
    (f(std::integral_constant<int, I>{}), ...);
}

// Input row r of a chunk (pairs P: P[2 HQ + j] is pair j of the lane's own
// bytes) adds its windows to the accumulators of output rows r - B + e,
// |e| <= b_k, slot (row mod L).
template <uint32_t BASE, int N, bool MAX, int HQ, int NP>
__device__ __attribute__((always_inline)) inline void contribute(const u16x2 (&P)[NP], u16x2 (&acc)[2 * Pow<BASE, N>::B + 1][4],
                                                                 int r) {
    using PW = Pow<BASE, N>;
    constexpr int B = PW::B, L = 2 * B + 1;
    u16x2 H[4];
#pragma unroll
    for (int j = 0; j < 4; j++) H[j] = P[2 * HQ + j];
    for_seq(std::make_integer_sequence<int, PW::A + 1>{}, [&](auto wc) {
        constexpr int W = decltype(wc)::value;
        if constexpr (W >= 1) h_level<1, HQ, MAX, W>(P, H);
        for_seq(std::make_integer_sequence<int, PW::R.n>{}, [&](auto kc) {
            constexpr int K = decltype(kc)::value;
            if constexpr (PW::R.a[K] == W) {
                constexpr int b = PW::R.b[K];
#pragma unroll
                for (int e = -b; e <= b; e++) {
                    const int s = ((r - B + e) % L + L) % L;
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[s][j] = pop<MAX>(acc[s][j], H[j]);
                }
            }
        });
    });
}

// MODE 0 erode, 1 dilate, 2 gradient (dilation - erosion); POST 1: out =
// saturate(c - v), 2: out = saturate(v - c).
template <uint32_t BASE, int N, int MODE, int POST>
__device__ __attribute__((always_inline)) inline void stream_wave(const MorphStreamArgs &a, int strip, int band) {
    using PW = Pow<BASE, N>;
    constexpr int A = PW::A, B = PW::B, L = 2 * B + 1;
    static_assert(A <= kMaxA && B <= kMaxB && B >= 1, "element outside the kernel's limits");
    constexpr int HQ = A <= 4 ? 1 : A <= 8 ? 2 : 3;  // halo dwords on each side
    constexpr int HL = HQ <= 2 ? 1 : 2;              // halo lanes on each side
    constexpr int SW = 8 * (64 - 2 * HL);            // output bytes per strip
    constexpr int D = L - 1 < 8 ? L - 1 : 8;         // rows loaded ahead
    constexpr int NP = 2 * (2 + 2 * HQ);             // sample pairs per row
    const int lane = (int)threadIdx.x & 63;
    const int n = a.s.w, H = a.s.h, tb = a.tb;
    const int xc = strip * SW - 8 * HL + 8 * lane;  // the lane's first byte
    const int Y0 = band * tb;
    if (Y0 >= H) return;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, a.span_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, a.span_d, kBufWord3);
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc(a.c.p, 0, a.span_c, kBufWord3);
    const int ss = (int)a.s.stride, ds = (int)a.d.stride, cs = (int)a.c.stride;
    // Every load reads 8 bytes inside its row: lanes past an end load the
    // row's first or last 8 bytes and take its edge byte from them.
    const int lo = min(max(xc, 0), n - 8);
    const int cls = xc < 0 ? 3 : xc >= n ? 2 : xc + 8 > n ? 1 : 0;  // 1: the row's last 4 bytes
    const bool edge = __builtin_amdgcn_ballot_w64(cls != 0) != 0;    // uniform
    const bool out_lane = lane >= HL && lane < 64 - HL && xc < n;
    const int NI = tb + 2 * B;  // input rows of the band
    auto row_off = [&](int i) { return min(max(Y0 - B + i, 0), H - 1) * ss; };
    vu32x2 ring[L];
#pragma unroll
    for (int r = 0; r < D; r++)
        ring[r] = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, lo, row_off(r), 0));
    const u16x2 idmax = {0, 0}, idmin = {255, 255};
    u16x2 amax[L][4], amin[L][4];
#pragma unroll
    for (int s = 0; s < L; s++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            amax[s][j] = idmax;
            amin[s][j] = idmin;
        }
    const int chunks = (NI + L - 1) / L;
    for (int c = 0; c < chunks; c++) {
#pragma unroll
        for (int r = 0; r < L; r++) {
            const int i = c * L + r;
            if (i >= NI) break;  // uniform
            if (i + D < NI)
                ring[(r + D) % L] =
                    __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, lo, row_off(i + D), 0));
            uint32_t d0 = ring[r].x, d1 = ring[r].y;
            if (edge) {
                if (cls == 1) {
                    d0 = ring[r].y;
                    d1 = __builtin_amdgcn_perm(d0, d0, 0x03030303u);
                } else if (cls == 2) {
                    d0 = d1 = __builtin_amdgcn_perm(ring[r].y, ring[r].y, 0x03030303u);
                } else if (cls == 3) {
                    d0 = d1 = __builtin_amdgcn_perm(ring[r].x, ring[r].x, 0x00000000u);
                }
            }
            uint32_t dw[2 + 2 * HQ];
            dw[HQ] = d0;
            dw[HQ + 1] = d1;
            dw[HQ - 1] = shr1(d1);
            dw[HQ + 2] = shl1(d0);
            if constexpr (HQ >= 2) {
                dw[HQ - 2] = shr1(d0);
                dw[HQ + 3] = shl1(d1);
            }
            if constexpr (HQ >= 3) {
                dw[HQ - 3] = shr1(dw[HQ - 1]);
                dw[HQ + 4] = shl1(dw[HQ + 2]);
            }
            u16x2 P[NP];
#pragma unroll
            for (int k = 0; k < 2 + 2 * HQ; k++) {
                P[2 * k] = lo_pair(dw[k]);
                P[2 * k + 1] = hi_pair(dw[k]);
            }
            if constexpr (MODE != 0) contribute<BASE, N, true, HQ, NP>(P, amax, r);
            if constexpr (MODE != 1) contribute<BASE, N, false, HQ, NP>(P, amin, r);
            // Output row o = i - 2B has all its rows; rows o < 0 and o >= tb
            // only collect the contributions no output needs.
            const int o = i - 2 * B;
            const int so = ((r - 2 * B) % L + L) % L;
            if (o >= 0 && o < tb && Y0 + o < H && out_lane) {
                u16x2 v[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    v[j] = MODE == 2 ? __builtin_elementwise_sub_sat(amax[so][j], amin[so][j]) : MODE == 1 ? amax[so][j] : amin[so][j];
                const int yo = Y0 + o;
                if constexpr (POST != 0) {
                    const vu32x2 cv = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rc, lo, yo * cs, 0));
                    const uint32_t c0 = cls == 1 ? cv.y : cv.x, c1 = cv.y;
                    const u16x2 cp[4] = {lo_pair(c0), hi_pair(c0), lo_pair(c1), hi_pair(c1)};
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        v[j] = POST == 1 ? __builtin_elementwise_sub_sat(cp[j], v[j]) : __builtin_elementwise_sub_sat(v[j], cp[j]);
                }
                const uint32_t o0 = pack2(v[0], v[1]), o1 = pack2(v[2], v[3]);
                if (cls == 0)
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vu32x2, u32x2{o0, o1}), rd, xc, yo * ds, 0);
                else
                    __builtin_amdgcn_raw_buffer_store_b32(o0, rd, xc, yo * ds, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                amax[so][j] = idmax;
                amin[so][j] = idmin;
            }
        }
    }
}

// One workgroup: 4 waves on 4 consecutive bands of one strip, tiles in XCD
// bands (neighbouring bands and strips share their halo lines in one L2).
template <uint32_t BASE, int N, int MODE, int POST>
__global__ __launch_bounds__(256) void morph_stream_u8(MorphStreamArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    stream_wave<BASE, N, MODE, POST>(a, strip, 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

// Output bytes per strip of (BASE, N): 496, or 480 with 3 halo dwords.
template <uint32_t BASE, int N>
constexpr int strip_bytes() {
    return Pow<BASE, N>::A <= 8 ? 496 : 480;
}

template <uint32_t BASE, int N, int MODE, int POST>
void launch_one(MorphStreamArgs a, hipStream_t st) {
    const int sw = strip_bytes<BASE, N>();
    a.strips = (a.s.w + sw - 1) / sw;
    const int groups = ((a.s.h + a.tb - 1) / a.tb + 3) / 4;
    a.ntiles = a.strips * groups;
    hipLaunchKernelGGL((morph_stream_u8<BASE, N, MODE, POST>), dim3(tile_grid(a.ntiles)), dim3(256), 0, st, a);
}

// The launches of one base element: n = 1 gradient only; n = 2, 3 erode,
// dilate, gradient, dilate with post 1 (tophat's second pass), erode with
// post 1 (the last run of a gradient split over launches) and erode with
// post 2 (blackhat's), where the n-th power takes at most
// kStreamMaxRows accumulator rows. False for other combinations.
template <uint32_t BASE>
bool launch_base(const MorphStreamArgs &a, int n, int mode, int post, hipStream_t st) {
    auto pick = [&](auto nc) -> bool {
        constexpr int NN = decltype(nc)::value;
        if constexpr (acc_rows(power(BASE, NN)) <= kStreamMaxRows) {
            if (mode == 2 && post == 0) return launch_one<BASE, NN, 2, 0>(a, st), true;
            if constexpr (NN >= 2) {
                if (mode == 0 && post == 0) return launch_one<BASE, NN, 0, 0>(a, st), true;
                if (mode == 1 && post == 0) return launch_one<BASE, NN, 1, 0>(a, st), true;
                if (mode == 1 && post == 1) return launch_one<BASE, NN, 1, 1>(a, st), true;
                if (mode == 0 && post == 1) return launch_one<BASE, NN, 0, 1>(a, st), true;
                if (mode == 0 && post == 2) return launch_one<BASE, NN, 0, 2>(a, st), true;
            }
        }
        return false;
    };
    switch (n) {
    case 1: return pick(std::integral_constant<int, 1>{});
    case 2: return pick(std::integral_constant<int, 2>{});
    case 3: return pick(std::integral_constant<int, 3>{});
    default: return false;
    }
}

}  // namespace mstream

// Streamed passes for the base elements with codes of their own
// (k_morph_stream_*.hip); false when (n, mode, post) has no kernel.
#define VISAR_STREAM_EL(name) bool launch_morph_stream_##name(const MorphStreamArgs &a, int n, int mode, int post, hipStream_t st);
VISAR_STREAM_EL(rect3)
VISAR_STREAM_EL(rect5)
VISAR_STREAM_EL(rect7)
VISAR_STREAM_EL(cross3)
VISAR_STREAM_EL(cross5)
VISAR_STREAM_EL(cross7)
VISAR_STREAM_EL(ellipse5)
VISAR_STREAM_EL(ellipse7)
#undef VISAR_STREAM_EL

}  // namespace visar
