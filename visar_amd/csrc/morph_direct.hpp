// morph_direct.hpp -- one erode/dilate pass on one-channel u8 images with
// the rows streamed from memory into registers (no LDS), for gfx950.
//
// Semantics as in morph_chain.hpp (reference src/imgproc/morph.zig:489-519,
// the subtraction of morphologyEx at :532-593): the element is a union of
// centred rectangles known at compile time, out-of-image cells are read as
// the edge-replicated image (exact for such elements).
//
// Each wave is independent: its 64 lanes own 8 bytes each of a 512-byte
// column strip, and it computes TB = 16 output rows from 16 + 2K source
// rows, rolling the vertical window down them as in the fused chains and
// storing each output row as it completes. A lane needs bytes [c - 4,
// c + 12) of each row (c = its first byte): its own 8 bytes are one load,
// the neighbouring dwords come from the neighbouring lanes. Bytes past a
// row's ends are rebuilt with a byte permute (edge replication). Requires
// rows of a multiple of 4 bytes, each starting on a dword, in source,
// destination and combine image.
#pragma once

#include "morph_chain.hpp"

namespace visar {

namespace mdirect {

using namespace mchain;

constexpr int kTB = kMorphDirectRows;  // output rows per wave
constexpr int kAhead = 8;  // rows in flight ahead of the one being computed
// DPP controls: lane i takes lane i - 1's value (shr) or lane i + 1's (shl);
// the lane without a source keeps the old value.
constexpr int kDppWaveShr1 = 0x138, kDppWaveShl1 = 0x130;
constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));

// Feeds row r of a pass's input (pairs P: bytes c - 4 + 2k, +1 of the
// lane's row) into its rolling window of output rows [0, TBP); when output
// row r - 2K completes, returns true with its pairs in `out` (r is a
// constant once the caller's row loop is unrolled).
template <int K, uint32_t EL, bool MAX, int TBP>
__device__ __attribute__((always_inline)) inline bool roll_step(const u16x2 (&P)[8], u16x2 (&acc)[2 * K + 1][4], int r,
                                                                u16x2 (&out)[4]) {
    constexpr int L = 2 * K + 1;
    constexpr int HQ = 1;  // K <= 3 bytes of halo: one dword each side
    u16x2 Hh[4];
#pragma unroll
    for (int j = 0; j < 4; j++) Hh[j] = P[2 * HQ + j];
    // Input row r feeds output rows r - K + e, |e| <= v, of each rectangle.
    auto contribute = [&](auto wc) {
        constexpr int W = decltype(wc)::value;
        for_rects<EL>([&](auto kc) {
            using R = RectOf<EL, decltype(kc)::value>;
            if constexpr (R::w == W) {
#pragma unroll
                for (int e = -R::v; e <= R::v; e++) {
                    const int y = r - K + e;
                    if (y >= 0 && y < TBP) {
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[y % L][j] = pop<MAX>(acc[y % L][j], Hh[j]);
                    }
                }
            }
        });
    };
    // The widest window a rectangle needs at this row.
    int need = -1;
    for_rects<EL>([&](auto kc) {
        using R = RectOf<EL, decltype(kc)::value>;
        if (r - K + R::v >= 0 && r - K - R::v < TBP) need = max(need, R::w);
    });
    contribute(std::integral_constant<int, 0>{});
    if constexpr (K >= 1) {
        if (need >= 1) {
            h_level<1, HQ, MAX, 1>(P, Hh);
            contribute(std::integral_constant<int, 1>{});
        }
    }
    if constexpr (K >= 2) {
        if (need >= 2) {
            h_level<1, HQ, MAX, 2>(P, Hh);
            contribute(std::integral_constant<int, 2>{});
        }
    }
    if constexpr (K >= 3) {
        if (need >= 3) {
            h_level<1, HQ, MAX, 3>(P, Hh);
            contribute(std::integral_constant<int, 3>{});
        }
    }
    const int y = r - 2 * K;
    if (y < 0 || y >= TBP) return false;
    const u16x2 idp = MAX ? u16x2{0, 0} : u16x2{255, 255};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        out[j] = acc[y % L][j];
        acc[y % L][j] = idp;
    }
    return true;
}

template <int CTRL>
__device__ __attribute__((always_inline)) inline u16x2 dpp_p(u16x2 old, u16x2 v) {
    return as_p(__builtin_amdgcn_update_dpp(as_u(old), as_u(v), CTRL, 0xf, 0xf, false));
}

// NP passes (1 or 2; bit p of ops: pass p dilates) of one wave: 8 output
// rows of the strip's lanes [NP - 1, 64 - (NP - 1)). The second pass reads
// the first one's rows as they complete, its neighbouring dwords from the
// neighbouring lanes; cells of the first pass's rows outside the image are
// set to the second pass's identity.
template <int K, uint32_t EL, int NP, int POST>
__device__ __attribute__((always_inline)) inline void direct_wave(const MorphDirectArgs &a, int strip, int band) {
    constexpr int L = 2 * K + 1;
    constexpr int ROWS = kTB + 2 * K * NP;
    constexpr int TB1 = kTB + 2 * K * (NP - 1);  // rows of the first pass
    const int lane = (int)threadIdx.x & 63;
    const int n = a.s.w, H = a.s.h;
    const int X0 = strip * (512 - 16 * (NP - 1)) - 8 * (NP - 1);
    const int xc = X0 + 8 * lane;  // first byte of the lane
    const int Y0 = band * kTB;     // first output row
    if (Y0 >= H) return;
    // Buffer loads and stores: a scalar row offset plus the lane's column
    // offset, no address arithmetic per access (every image spans less
    // than 2^31 bytes; accesses outside an image's span read 0, and those
    // values are never used).
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.s.p, 0, a.span_s, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(a.d.p, 0, a.span_d, kBufWord3);
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc(a.c.p, 0, a.span_c, kBufWord3);
    const int ss = (int)a.s.stride, ds = (int)a.d.stride, cs = (int)a.c.stride;
    // Each row: the lane's own 8 bytes in one load; the dword before them is
    // the previous lane's second dword and the dword after them the next
    // lane's first (DPP wave shifts); the strip's outer two (bytes X0 - 4
    // and X0 + 512) are loads every lane makes at one address.
    uint32_t mx[ROWS], my[ROWS], hl[ROWS], hr[ROWS];
    auto load_row = [&](int r) {
        const int so = min(max(Y0 - K * NP + r, 0), H - 1) * ss;
        const vu32x2 m = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, xc, so, 0));
        mx[r] = m.x;
        my[r] = m.y;
        hl[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, X0 - 4, so, 0);
        hr[r] = __builtin_amdgcn_raw_buffer_load_b32(rs, X0 + 512, so, 0);
    };
    // Edge replication of the source: the row's first byte left of it, its
    // last byte right of it (the lane holding byte n - 1).
    const bool e_left = xc == 0, e_half = xc + 4 == n, e_full = xc + 8 == n;
    const bool edge = __builtin_amdgcn_ballot_w64(e_left || e_half || e_full) != 0;  // uniform
    // Lanes with bytes outside the image's columns (first-pass rows).
    const bool out_lo = xc < 0 || xc >= n, out_hi = out_lo || e_half;
    const bool oob = __builtin_amdgcn_ballot_w64(out_hi) != 0;  // uniform
#pragma unroll
    for (int r = 0; r < kAhead && r < ROWS; r++) load_row(r);
    const bool max1 = a.ops & 1, max2 = (a.ops >> 1) & 1;
    const u16x2 id1 = max1 ? u16x2{0, 0} : u16x2{255, 255};
    const u16x2 id2 = max2 ? u16x2{0, 0} : u16x2{255, 255};
    u16x2 acc1[L][4], acc2[L][4];
#pragma unroll
    for (int s = 0; s < L; s++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc1[s][j] = id1;
            acc2[s][j] = id2;
        }
    const bool store = lane >= NP - 1 && lane < 64 - (NP - 1) && xc >= 0 && xc < n;
    auto store_row = [&](int yo, u16x2 (&v)[4]) {
        if (Y0 + yo < H && store) {
            if constexpr (POST != 0) {
                uint32_t c0, c1 = 0;
                if (xc + 8 <= n) {
                    const vu32x2 cv = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rc, xc, (Y0 + yo) * cs, 0));
                    c0 = cv.x;
                    c1 = cv.y;
                } else {
                    c0 = __builtin_amdgcn_raw_buffer_load_b32(rc, xc, (Y0 + yo) * cs, 0);
                }
                const u16x2 cp[4] = {lo_pair(c0), hi_pair(c0), lo_pair(c1), hi_pair(c1)};
#pragma unroll
                for (int j = 0; j < 4; j++)
                    v[j] = POST == 1 ? __builtin_elementwise_sub_sat(cp[j], v[j]) : __builtin_elementwise_sub_sat(v[j], cp[j]);
            }
            const uint32_t o0 = pack2(v[0], v[1]), o1 = pack2(v[2], v[3]);
            if (xc + 8 <= n) {
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vu32x2, u32x2{o0, o1}), rd, xc, (Y0 + yo) * ds, 0);
            } else {
                // n is a multiple of 4: the row's last lane holds 4 bytes
                __builtin_amdgcn_raw_buffer_store_b32(o0, rd, xc, (Y0 + yo) * ds, 0);
            }
        }
    };

    // Rows are loaded kAhead rows before they are used; the scheduling
    // barriers keep each row's work after its own loads, so a wave computes
    // its first rows while the later ones are still on the way.
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        if (r + kAhead < ROWS) load_row(r + kAhead);
        __builtin_amdgcn_sched_barrier(0);
        uint32_t dw[4];
        dw[1] = mx[r];
        dw[2] = my[r];
        dw[0] = __builtin_amdgcn_update_dpp(hl[r], my[r], kDppWaveShr1, 0xf, 0xf, false);
        dw[3] = __builtin_amdgcn_update_dpp(hr[r], mx[r], kDppWaveShl1, 0xf, 0xf, false);
        if (edge) {
            if (e_left) dw[0] = __builtin_amdgcn_perm(dw[1], dw[1], 0x00000000u);
            if (e_half) dw[2] = dw[3] = __builtin_amdgcn_perm(dw[1], dw[1], 0x03030303u);
            if (e_full) dw[3] = __builtin_amdgcn_perm(dw[2], dw[2], 0x03030303u);
        }
        u16x2 P[8];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            P[2 * i] = lo_pair(dw[i]);
            P[2 * i + 1] = hi_pair(dw[i]);
        }
        u16x2 o1[4];
        const bool done1 = max1 ? roll_step<K, EL, true, TB1>(P, acc1, r, o1) : roll_step<K, EL, false, TB1>(P, acc1, r, o1);
        if (done1) {
            const int y1 = r - 2 * K;
            if constexpr (NP == 1) {
                store_row(y1, o1);
            } else {
                // First-pass row y1 is image row Y0 - K + y1.
                const int yi = Y0 - K + y1;
                if (yi < 0 || yi >= H) {
#pragma unroll
                    for (int j = 0; j < 4; j++) o1[j] = id2;
                } else if (oob) {
                    if (out_lo) o1[0] = o1[1] = id2;
                    if (out_hi) o1[2] = o1[3] = id2;
                }
                u16x2 P2[8];
                P2[0] = dpp_p<kDppWaveShr1>(id2, o1[2]);
                P2[1] = dpp_p<kDppWaveShr1>(id2, o1[3]);
#pragma unroll
                for (int j = 0; j < 4; j++) P2[2 + j] = o1[j];
                P2[6] = dpp_p<kDppWaveShl1>(id2, o1[0]);
                P2[7] = dpp_p<kDppWaveShl1>(id2, o1[1]);
                u16x2 o2[4];
                const bool done2 = max2 ? roll_step<K, EL, true, kTB>(P2, acc2, y1, o2)
                                        : roll_step<K, EL, false, kTB>(P2, acc2, y1, o2);
                if (done2) store_row(y1 - 2 * K, o2);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One workgroup: 4 waves on 4 consecutive bands of one strip (neighbouring
// bands share their halo rows in L2), tiles in XCD bands.
template <int K, uint32_t EL, int NP, int POST>
__global__ __launch_bounds__(256) void morph_direct_u8(MorphDirectArgs a) {
    int strip, bg;
    if (!tile_of(TileOrder{a.strips, a.ntiles, a.xcd}, strip, bg)) return;
    direct_wave<K, EL, NP, POST>(a, strip, 4 * bg + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
}

template <int K, uint32_t EL, int NP>
void launch_direct(const MorphDirectArgs &a, int post, hipStream_t st) {
    const dim3 grid(tile_grid(a.ntiles));
    switch (post) {
    case 1: hipLaunchKernelGGL((morph_direct_u8<K, EL, NP, 1>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((morph_direct_u8<K, EL, NP, 2>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((morph_direct_u8<K, EL, NP, 0>), grid, dim3(256), 0, st, a); break;
    }
}

template <int NP>
bool launch_direct_el(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st) {
    switch (el) {
    case kElRect3: launch_direct<1, kElRect3, NP>(a, post, st); return true;
    case kElRect5: launch_direct<2, kElRect5, NP>(a, post, st); return true;
    case kElRect7: launch_direct<3, kElRect7, NP>(a, post, st); return true;
    case kElCross3: launch_direct<1, kElCross3, NP>(a, post, st); return true;
    case kElCross5: launch_direct<2, kElCross5, NP>(a, post, st); return true;
    case kElCross7: launch_direct<3, kElCross7, NP>(a, post, st); return true;
    case kElEllipse5: launch_direct<2, kElEllipse5, NP>(a, post, st); return true;
    case kElEllipse7: launch_direct<3, kElEllipse7, NP>(a, post, st); return true;
    default: return false;
    }
}

}  // namespace mdirect

// One or two passes (k_morph_direct1.hip / k_morph_direct2.hip); false
// when the element has no kernel of its own.
bool launch_morph_direct1(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st);
bool launch_morph_direct2(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st);

}  // namespace visar
