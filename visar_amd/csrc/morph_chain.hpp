// morph_chain.hpp -- fused erode/dilate chains on u8 images for gfx950: the
// device code, instantiated by k_morph_chain*.hip.
//
// Reference semantics: src/imgproc/morph.zig:489-519 (one pass: out(y, x, c)
// is the min (erode) or max (dilate) of the source over the element's "on"
// cells, cells outside the image ignored), :63-96 (iterations repeat the
// pass) and :532-593 (morphologyEx: open, close, gradient, tophat and
// blackhat out of passes and a saturating subtraction).
//
// The elements these kernels take are unions of centred rectangles (every
// rect, cross and ellipse of odd size up to 7 x 7 is one): rows whose run of
// on cells is centred on the anchor, as wide in the centre row as anywhere
// and no wider further out. For such an element:
//  - the min over the element is the min over its rectangles, and the min
//    over a rectangle is the min over its rows of the min over a horizontal
//    window (separable);
//  - ignoring cells outside the image gives the same result as reading the
//    image extended by edge replication, since each rectangle that holds an
//    out-of-image cell also holds the in-image cell it replicates (both lie
//    in the rectangle, which contains the output pixel itself).
//
// One workgroup of 4 waves owns a tile of TH rows x (512 - 2 HB) bytes. It
// stages the tile with HB bytes and VB rows of halo on each side in LDS
// (replicated at the image's edges), then runs a program of passes between
// LDS buffers; the halo shrinks by one element per pass. Out-of-image cells
// of an intermediate are set to the identity of the pass that reads them
// (0 for dilate, 255 for erode), which ignores them exactly. The last pass
// stores to the destination, folding in the subtraction of the compound ops
// (saturate(C - v) or saturate(v - C) against an LDS buffer C). Rows of a
// buffer are 512 bytes (+ pads), so lane l of every wave owns bytes
// [8l, 8l + 8) of a row; the 4 waves split a pass's rows into 4 bands and
// walk down them, rolling the vertical window in registers. Samples are held
// as two 16-bit lanes per register, so one v_pk_min_u16 / v_pk_max_u16 takes
// the min or max of two samples.
#pragma once

#include <utility>

#include "tiles.hpp"

namespace visar {
namespace mchain {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef uint64_t u64_a4 __attribute__((aligned(4)));

constexpr int kThreads = 256;
constexpr int kRowDw = 128;  // dwords of tile data per LDS row

template <int CN, int K>
struct ChainGeo {
    static constexpr int HQ = (K * CN + 3) / 4;     // halo dwords a lane converts on each side
    static constexpr int PAD = (HQ + 1) & ~1;       // pad dwords at each end of a row (even)
    static constexpr int PITCH = kRowDw + 2 * PAD;  // dwords per LDS row
    static constexpr int ND = 2 + 2 * HQ;           // dwords a lane converts per row
    static constexpr int RS = HQ & 1;               // leading dword read for 8-byte alignment
    static constexpr int NR = ND + 2 * RS;          // dwords read per row (even)
    static constexpr int L = 2 * K + 1;             // rows of the rolling window
};

__device__ __attribute__((always_inline)) inline u16x2 as_p(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __attribute__((always_inline)) inline uint32_t as_u(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
// Bytes 0, 1 (lo) or 2, 3 (hi) of v as two 16-bit lanes.
__device__ __attribute__((always_inline)) inline u16x2 lo_pair(uint32_t v) {
    return as_p(__builtin_amdgcn_perm(0u, v, 0x0c010c00u));
}
__device__ __attribute__((always_inline)) inline u16x2 hi_pair(uint32_t v) {
    return as_p(__builtin_amdgcn_perm(0u, v, 0x0c030c02u));
}
// The bytes of pairs a (bytes 0, 1) and b (bytes 2, 3) as one dword.
__device__ __attribute__((always_inline)) inline uint32_t pack2(u16x2 a, u16x2 b) {
    return __builtin_amdgcn_perm(as_u(b), as_u(a), 0x06040200u);
}
template <bool MAX>
__device__ __attribute__((always_inline)) inline u16x2 pop(u16x2 a, u16x2 b) {
    if constexpr (MAX) return __builtin_elementwise_max(a, b);
    else return __builtin_elementwise_min(a, b);
}

// A compile-time element: EL = nr | (w_k | v_k << 2) << (4 + 4k) for its
// rectangles k < nr, listed from the centre row outward (so rectangle 0 is
// the widest); EL 0: the element is read from the launch's arguments.
template <uint32_t EL, int k>
struct RectOf {
    static constexpr int w = (int)((EL >> (4 + 4 * k)) & 3u), v = (int)((EL >> (6 + 4 * k)) & 3u);
};
template <typename F, int... I>
__device__ __attribute__((always_inline)) inline void for_each_k(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <uint32_t EL, typename F>
__device__ __attribute__((always_inline)) inline void for_rects(F &&f) {
    for_each_k(f, std::make_integer_sequence<int, (int)(EL & 15u)>{});
}
// True when every rectangle reaches K rows up and down: each row of a band's
// rolling window then needs every rectangle.
template <uint32_t EL, int K>
constexpr bool all_rows() {
    for (int k = 0; k < (int)(EL & 15u); k++)
        if ((int)((EL >> (6 + 4 * k)) & 3u) != K) return false;
    return true;
}

// Pair j of the lane's output bytes moved by t bytes, out of the converted
// pairs P (pair k = bytes c - 4 HQ + 2k, +1 of the lane's row, c = 8 lane).
template <int HQ, int T>
__device__ __attribute__((always_inline)) inline u16x2 tap(const u16x2 *P, int j) {
    constexpr int B = 2 * HQ;
    if constexpr ((T & 1) == 0) {
        return P[B + j + T / 2];
    } else {
        return as_p(__builtin_amdgcn_alignbyte(as_u(P[B + j + (T + 1) / 2]), as_u(P[B + j + (T - 1) / 2]), 2));
    }
}

// Horizontal level w (window [-w, w] in samples) from level w - 1.
template <int CN, int HQ, bool MAX, int W>
__device__ __attribute__((always_inline)) inline void h_level(const u16x2 *P, u16x2 *H) {
#pragma unroll
    for (int j = 0; j < 4; j++) H[j] = pop<MAX>(H[j], pop<MAX>(tap<HQ, -W * CN>(P, j), tap<HQ, W * CN>(P, j)));
}

template <int CN>
__device__ __attribute__((always_inline)) inline int rep_byte(int x, int n) {
    // Edge replication of byte x of a row of n bytes (pixels of CN bytes).
    if (x < 0) return ((x % CN) + CN) % CN;
    if (x >= n) return n - CN + (x - n) % CN;
    return x;
}

// Dword of a row at byte x (x a multiple of 4), bytes outside [0, n) replicated.
template <int CN>
__device__ __attribute__((always_inline)) inline uint32_t load_dw(gptr<const uint8_t> row, int x, int n) {
    const int a = (int)((uintptr_t)row & 3);
    if (x >= 0 && x + 4 <= n) {
        if (a == 0) return *(gptr<const uint32_t>)(row + x);
        if (x - a >= 0 && x - a + 8 <= n) {
            const uint64_t v = *(gptr<const u64_a4>)(row + x - a);
            return __builtin_amdgcn_alignbyte((uint32_t)(v >> 32), (uint32_t)v, (uint32_t)a);
        }
    }
    uint32_t v = 0;
#pragma unroll
    for (int t = 0; t < 4; t++) v |= (uint32_t)row[rep_byte<CN>(x + t, n)] << (8 * t);
    return v;
}

// Stages rows [r0, r0 + nrows) of image `im` (image row y0 + r, clamped) into
// LDS rows (lds + r * PITCH + PAD), bytes x0 .. x0 + 511 of each row. On an
// image whose rows all start on a dword, every load of a batch is issued
// before any is waited for: one dword load per lane and row at a clamped
// address, then the lanes whose dword crosses the row's ends rebuild it
// byte by byte. Other images take load_dw per dword.
template <int CN, int PITCH, int PAD>
__device__ __attribute__((always_inline)) inline void stage(const DImg &im, int x0, int y0, int r0, int nrows,
                                                            uint32_t *lds) {
    const int n = im.w * CN, H = im.h;
    // Each pair of waves stages one row per step: the row is uniform.
    const int half = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);
    const int m = (int)threadIdx.x & (kRowDw - 1);
    const int x = x0 + 4 * m;
    const bool fast = ((((uintptr_t)im.p) | (uintptr_t)im.stride) & 3) == 0 && n >= 4;
    if (fast) {
        const bool in = x >= 0 && x + 4 <= n;
        const int xc = in ? x : 0;
        for (int base = 0; base < nrows; base += 16) {
            uint32_t v[8];
            const int ys = y0 + r0 + base + half;
            if (ys >= 0 && ys + 14 < H) {
                const gptr<const uint8_t> p0 = grow<const uint8_t>(im, ys) + xc;
                const long st2 = 2 * im.stride;
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = *(gptr<const uint32_t>)(p0 + i * st2);
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int sy = min(max(ys + 2 * i, 0), H - 1);
                    v[i] = *(gptr<const uint32_t>)(grow<const uint8_t>(im, sy) + xc);
                }
            }
            if (!in) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int sy = min(max(y0 + r0 + base + 2 * i + half, 0), H - 1);
                    const gptr<const uint8_t> row = grow<const uint8_t>(im, sy);
                    uint32_t w = 0;
#pragma unroll
                    for (int t = 0; t < 4; t++) w |= (uint32_t)row[rep_byte<CN>(x + t, n)] << (8 * t);
                    v[i] = w;
                }
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int r = base + 2 * i + half;
                if (r < nrows) lds[(r0 + r) * PITCH + PAD + m] = v[i];
            }
        }
        return;
    }
    for (int base = 0; base < nrows; base += 16) {
        uint32_t v[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = base + 2 * i + half;
            const int sy = min(max(y0 + r0 + r, 0), H - 1);
            v[i] = r < nrows ? load_dw<CN>(grow<const uint8_t>(im, sy), x, n) : 0u;
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = base + 2 * i + half;
            if (r < nrows) lds[(r0 + r) * PITCH + PAD + m] = v[i];
        }
    }
}

struct PassCtx {
    int lane, wave;
    int X0, Y0;      // image position of LDS byte 0 / row 0
    bool edge;       // the tile reaches outside the image
    uint32_t km0, km1;  // per-lane keep masks: bytes inside the image's columns
};

// One pass: rows [q ev, rows - q ev) of `dst` (or, FINAL, the destination
// image rows of the tile) from `src`.
template <int CN, int K, uint32_t EL, bool MAX, bool FINAL>
__device__ __attribute__((always_inline)) inline void chain_pass(const MorphChainArgs &a, const PassCtx &cx,
                                                                 const uint32_t *src, uint32_t *dst,
                                                                 const uint32_t *cbuf, int q, uint32_t nid) {
    using G = ChainGeo<CN, K>;
    constexpr int HQ = G::HQ, PAD = G::PAD, PITCH = G::PITCH, ND = G::ND, RS = G::RS, NR = G::NR, L = G::L;
    const int lo = q * a.ev, hi = a.rows - q * a.ev;
    const int band = (hi - lo + 3) >> 2;
    const int y0w = lo + cx.wave * band, y1w = min(y0w + band, hi);
    if (y0w >= y1w) return;
    const u16x2 idp = MAX ? u16x2{0, 0} : u16x2{255, 255};
    u16x2 acc[L][4];
#pragma unroll
    for (int s = 0; s < L; s++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[s][j] = idp;
    const int n = a.s.w * CN;
    const int xb = cx.X0 + 8 * cx.lane;  // image byte of the lane's first output byte
    const bool store_lane = FINAL && cx.lane >= a.hb / 8 && cx.lane < (a.hb + a.tw) / 8 && xb < n;
    // Lanes whose 8 bytes lie inside the row on a dword of every row store
    // them with one 8-byte store.
    const bool store_dw = store_lane && xb + 8 <= n && ((((uintptr_t)a.d.p + xb) | (uintptr_t)a.d.stride) & 3) == 0;
    // Destination row of the next output row (Y0 + y0w onwards).
    gptr<uint8_t> drow = grow<uint8_t>(a.d, cx.Y0 + y0w) + xb;
    const int end = y1w + K;
    for (int r0 = y0w - K; r0 < end; r0 += L) {
#pragma unroll
        for (int qq = 0; qq < L; qq++) {
            const int rr = r0 + qq;
            if (rr >= end) break;
            // The widest horizontal window a rectangle needs at this row:
            // every rectangle's inside the band, fewer in its halo rows.
            int need = -1;
            if constexpr (EL != 0 && all_rows<EL, K>()) {
                need = RectOf<EL, 0>::w;
            } else if constexpr (EL != 0) {
                if (rr >= y0w && rr < y1w) {
                    need = RectOf<EL, 0>::w;
                } else {
                    for_rects<EL>([&](auto kc) {
                        using R = RectOf<EL, decltype(kc)::value>;
                        if (rr >= y0w - R::v && rr < y1w + R::v) need = max(need, R::w);
                    });
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k < a.nr && rr >= y0w - a.rv[k] && rr < y1w + a.rv[k]) need = max(need, a.rw[k]);
            }
            if (need >= 0) {
                // Rows a rectangle needs lie inside the buffer: rr >= q ev - ev >= 0, rr < rows.
                const u32x2 *rp = reinterpret_cast<const u32x2 *>(src + rr * PITCH + PAD + 2 * cx.lane - HQ - RS);
                uint32_t dw[NR];
#pragma unroll
                for (int i = 0; i < NR / 2; i++) {
                    const u32x2 t = rp[i];
                    dw[2 * i] = t.x;
                    dw[2 * i + 1] = t.y;
                }
                u16x2 P[2 * ND];
#pragma unroll
                for (int i = 0; i < ND; i++) {
                    P[2 * i] = lo_pair(dw[RS + i]);
                    P[2 * i + 1] = hi_pair(dw[RS + i]);
                }
                u16x2 H[4];
#pragma unroll
                for (int j = 0; j < 4; j++) H[j] = P[2 * HQ + j];
                // Level w's window into the rows it reaches: the rectangle
                // of half-width w adds H to outputs e rows away, |e| <= v.
                auto contribute = [&](auto wc) {
                    constexpr int W = decltype(wc)::value;
                    if constexpr (EL != 0) {
                        for_rects<EL>([&](auto kc) {
                            using R = RectOf<EL, decltype(kc)::value>;
                            if constexpr (R::w == W) {
#pragma unroll
                                for (int e = -R::v; e <= R::v; e++)
#pragma unroll
                                    for (int j = 0; j < 4; j++)
                                        acc[(qq + e + L) % L][j] = pop<MAX>(acc[(qq + e + L) % L][j], H[j]);
                            }
                        });
                    } else {
                        const uint32_t m = a.cmask[W];
#pragma unroll
                        for (int e = -K; e <= K; e++) {
                            if ((m >> (e + K)) & 1u) {
#pragma unroll
                                for (int j = 0; j < 4; j++)
                                    acc[(qq + e + L) % L][j] = pop<MAX>(acc[(qq + e + L) % L][j], H[j]);
                            }
                        }
                    }
                };
                contribute(std::integral_constant<int, 0>{});
                if constexpr (K >= 1) {
                    if (need >= 1) {
                        h_level<CN, HQ, MAX, 1>(P, H);
                        contribute(std::integral_constant<int, 1>{});
                    }
                }
                if constexpr (K >= 2) {
                    if (need >= 2) {
                        h_level<CN, HQ, MAX, 2>(P, H);
                        contribute(std::integral_constant<int, 2>{});
                    }
                }
                if constexpr (K >= 3) {
                    if (need >= 3) {
                        h_level<CN, HQ, MAX, 3>(P, H);
                        contribute(std::integral_constant<int, 3>{});
                    }
                }
            }
            // Output row rr - K is complete.
            const int y = rr - K;
            const int s = (qq + L - K) % L;  // a constant once the loop is unrolled
            if (y >= y0w) {
                if constexpr (!FINAL) {
                    uint32_t o0 = pack2(acc[s][0], acc[s][1]), o1 = pack2(acc[s][2], acc[s][3]);
                    if (cx.edge) {
                        const int yi = cx.Y0 + y;
                        const bool rin = yi >= 0 && yi < a.s.h;
                        const uint32_t k0 = rin ? cx.km0 : 0u, k1 = rin ? cx.km1 : 0u;
                        o0 = (o0 & k0) | (nid & ~k0);
                        o1 = (o1 & k1) | (nid & ~k1);
                    }
                    *reinterpret_cast<u32x2 *>(dst + y * PITCH + PAD + 2 * cx.lane) = u32x2{o0, o1};
                } else {
                    const int yi = cx.Y0 + y;
                    if (yi < a.s.h && store_lane) {
                        u16x2 v[4] = {acc[s][0], acc[s][1], acc[s][2], acc[s][3]};
                        if (a.post) {
                            const u32x2 c = *reinterpret_cast<const u32x2 *>(cbuf + y * PITCH + PAD + 2 * cx.lane);
                            const u16x2 cp[4] = {lo_pair(c.x), hi_pair(c.x), lo_pair(c.y), hi_pair(c.y)};
#pragma unroll
                            for (int j = 0; j < 4; j++)
                                v[j] = a.post == 1 ? __builtin_elementwise_sub_sat(cp[j], v[j])
                                                   : __builtin_elementwise_sub_sat(v[j], cp[j]);
                        }
                        const uint32_t o0 = pack2(v[0], v[1]), o1 = pack2(v[2], v[3]);
                        if (store_dw) {
                            *(gptr<u64_a4>)drow = (uint64_t)o0 | ((uint64_t)o1 << 32);
                        } else {
#pragma unroll
                            for (int t = 0; t < 8; t++)
                                if (xb + t < n) drow[t] = (uint8_t)((t < 4 ? o0 >> (8 * t) : o1 >> (8 * (t - 4))) & 255u);
                        }
                    }
                    drow += a.d.stride;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) acc[s][j] = idp;
        }
    }
}

// Tile (bx, by) of one launch (also run by the batch kernels, with the
// arguments in device memory). lds: nbuf buffers of a.rows x PITCH dwords
// (plus the combine image's for comb 3).
template <int CN, int K, uint32_t EL>
__device__ __attribute__((always_inline)) inline void chain_tile(const MorphChainArgs &a, int bx, int by, uint32_t *lds) {
    using G = ChainGeo<CN, K>;
    const int buf_dw = a.rows * G::PITCH;
    PassCtx cx;
    cx.lane = (int)threadIdx.x & 63;
    cx.wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // uniform: row bands branch on the scalar unit
    const int n = a.s.w * CN;
    cx.X0 = bx * a.tw - a.hb;
    cx.Y0 = by * a.th - a.vb;
    cx.edge = cx.X0 < 0 || cx.X0 + 4 * kRowDw > n || cx.Y0 < 0 || cx.Y0 + a.rows > a.s.h;
    cx.km0 = cx.km1 = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int x = cx.X0 + 8 * cx.lane + t;
        const uint32_t bit = (x >= 0 && x < n) ? 0xffu : 0u;
        if (t < 4) cx.km0 |= bit << (8 * t);
        else cx.km1 |= bit << (8 * (t - 4));
    }
    stage<CN, G::PITCH, G::PAD>(a.s, cx.X0, cx.Y0, 0, a.rows, lds);
    // A combine image (comb 3) is staged without halo after the pass
    // buffers; cbuf + y * PITCH then addresses its tile row y - vb.
    uint32_t *const cimg = lds + a.nbuf * buf_dw - a.vb * G::PITCH;
    if (a.comb == 3) stage<CN, G::PITCH, G::PAD>(a.c, cx.X0, cx.Y0, a.vb, a.th, cimg);
    __syncthreads();
    for (int i = 0; i < a.npass; i++) {
        const MorphChainPass ps = a.pass[i];
        const uint32_t *src = lds + ps.src * buf_dw;
        const uint32_t nid = i + 1 < a.npass && !a.pass[i + 1].op_max ? 0xffffffffu : 0u;
        if (ps.dst >= 0) {
            uint32_t *dst = lds + ps.dst * buf_dw;
            if (ps.op_max) chain_pass<CN, K, EL, true, false>(a, cx, src, dst, nullptr, ps.q, nid);
            else chain_pass<CN, K, EL, false, false>(a, cx, src, dst, nullptr, ps.q, nid);
            __syncthreads();
        } else {
            const uint32_t *cb = !a.post ? nullptr : a.comb == 3 ? cimg : lds + a.comb * buf_dw;
            if (ps.op_max) chain_pass<CN, K, EL, true, true>(a, cx, src, nullptr, cb, ps.q, 0u);
            else chain_pass<CN, K, EL, false, true>(a, cx, src, nullptr, cb, ps.q, 0u);
        }
    }
}

template <int CN, int K, uint32_t EL>
__global__ __launch_bounds__(kThreads) void morph_chain_u8(MorphChainArgs a) {
    extern __shared__ __align__(16) uint32_t lds[];
    int bx, by;
    if (!tile_of(TileOrder{a.tiles_x, a.ntiles, a.xcd}, bx, by)) return;
    chain_tile<CN, K, EL>(a, bx, by, lds);
}

template <int CN, int K, uint32_t EL>
void launch_chain(const MorphChainArgs &a, size_t lds, hipStream_t st) {
    // Dynamic LDS above 64 KiB is opted into per kernel.
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&morph_chain_u8<CN, K, EL>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((morph_chain_u8<CN, K, EL>), dim3(tile_grid(a.ntiles)), dim3(kThreads), lds, st, a);
}

// The element codes with kernels of their own: the square rect, cross and
// ellipse elements of sides 3, 5 and 7 (the 3 x 3 ellipse is the cross).
constexpr uint32_t kElRect3 = 0x51, kElRect5 = 0xA1, kElRect7 = 0xF1;
constexpr uint32_t kElCross3 = 0x412, kElCross5 = 0x822, kElCross7 = 0xC32;
constexpr uint32_t kElEllipse5 = 0x862, kElEllipse7 = 0xCA73;

template <int CN>
bool launch_chain_el(const MorphChainArgs &a, uint32_t el, size_t lds, hipStream_t st) {
    switch (el) {
    case kElRect3: launch_chain<CN, 1, kElRect3>(a, lds, st); return true;
    case kElRect5: launch_chain<CN, 2, kElRect5>(a, lds, st); return true;
    case kElRect7: launch_chain<CN, 3, kElRect7>(a, lds, st); return true;
    case kElCross3: launch_chain<CN, 1, kElCross3>(a, lds, st); return true;
    case kElCross5: launch_chain<CN, 2, kElCross5>(a, lds, st); return true;
    case kElCross7: launch_chain<CN, 3, kElCross7>(a, lds, st); return true;
    case kElEllipse5: launch_chain<CN, 2, kElEllipse5>(a, lds, st); return true;
    case kElEllipse7: launch_chain<CN, 3, kElEllipse7>(a, lds, st); return true;
    default: return false;
    }
}

}  // namespace mchain

// Kernels for the elements with codes of their own (k_morph_chain1.hip:
// one channel, k_morph_chain3.hip: three and four); false for others.
bool launch_morph_chain_el1(const MorphChainArgs &a, uint32_t el, size_t lds, hipStream_t st);
bool launch_morph_chain_el34(const MorphChainArgs &a, uint32_t el, size_t lds, hipStream_t st);

}  // namespace visar
