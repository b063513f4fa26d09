// warp_blk.hpp -- device functions of the block-layout u8 warps (k_warp.hip),
// shared with the batch kernels (batch.hip).
#pragma once

#include "tiles.hpp"

namespace visar {

// ------------------------------------------------ block-layout warps
//
// One-channel u8 bilinear warpAffine / warpPerspective with BORDER_CONSTANT.
// A wave owns 64 x 4 output pixels as four blocks of 16 x 4; lane l computes
// pixel (16 j + l % 16, l / 16) of block j, so the 64 lanes of one gather
// instruction cover a compact 16 x 4 patch and their taps a compact source
// region under any rotation (one output row per wave spreads a rotated
// instruction over many more source rows). A tap row's two samples come
// from one 8-byte load at the dword-aligned address at or below the first
// sample, aligned into place with v_alignbyte: half the load instructions of
// byte gathers, whose per-instruction address work in the texture unit is
// what bounds gather kernels. Results go through LDS so each lane stores 4
// adjacent output bytes. A wave whose footprints are not all inside the
// image (or whose 8-byte loads would pass the image's last byte) gathers
// byte by byte with the constant-border rules instead.

constexpr int kBlkW = 16, kBlkH = 4;  // pixels of one gather instruction

struct BlkCoord {
    int ix, iy;
    float ax, ay;
};

// warpAffine bilinear blend (geometry.zig:508-533): fused, in this order.
__device__ __attribute__((always_inline)) inline float blend_affine(float p00, float p01, float p10, float p11, float ax,
                                                                    float ay) {
    const float v0 = fmaf(ax, p01 - p00, p00);
    const float v1 = fmaf(ax, p11 - p10, p10);
    return fmaf(ay, v1 - v0, v0);
}

// warpPerspective bilinear blend (geometry.zig:660-668): f32 weights, the
// sum ((0 + v0 w0) + v1 w1) + ... unfused. 0 + v0 w0 is v0 w0 exactly (a
// product of non-negative values is +0 or positive), so the sum starts there.
__device__ __attribute__((always_inline)) inline float blend_persp(float p00, float p01, float p10, float p11, float ax,
                                                                   float ay) {
    const float w0 = (1 - ax) * (1 - ay), w1 = ax * (1 - ay), w2 = (1 - ax) * ay, w3 = ax * ay;
    float acc = p00 * w0;
    acc += p01 * w1;
    acc += p10 * w2;
    acc += p11 * w3;
    return acc;
}

typedef uint64_t u64_a4 __attribute__((aligned(4)));
typedef unsigned int vu32x2 __attribute__((ext_vector_type(2)));
constexpr int kWarpBufWord3 = 0x00020000;  // buffer resource word 3 for gfx9 raw (untyped) access

// Blends of the lane's 4 pixels. `lim` = (image base & 3) + (H - 1) * stride
// + W: the bytes from the dword-aligned address below the image base to the
// end of its last row (fits in 32 bits, checked by the host).
template <bool PERSP>
__device__ __attribute__((always_inline)) inline void blk_sample(const DImg &s, float cv, uint32_t lim,
                                                                 const BlkCoord (&c)[4], float (&out)[4]) {
    const int W = s.w, H = s.h;
    const uint32_t st = (uint32_t)s.stride;
    const uint32_t b3 = (uint32_t)((uintptr_t)s.p & 3);
    const uint32_t lim8 = lim - 8;  // o1 <= lim8: the 8-byte load at o1 & ~3 ends inside the image
    bool inner = true, none = true;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t o1 = (uint32_t)(c[j].iy + 1) * st + (uint32_t)c[j].ix + b3;
        inner = inner && (unsigned)c[j].ix < (unsigned)(W - 1) && (unsigned)c[j].iy < (unsigned)(H - 1) && o1 <= lim8;
        none = none && (c[j].ix < -1 || c[j].iy < -1 || c[j].ix >= W || c[j].iy >= H);
    }
    if (__all(none)) {
#pragma unroll
        for (int j = 0; j < 4; j++) out[j] = cv;
        return;
    }
    if (__all(inner)) {
        const gptr<const uint8_t> base = gp(s.p - b3);
        uint64_t r0[4], r1[4];
        uint32_t sh[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t o0 = (uint32_t)c[j].iy * st + (uint32_t)c[j].ix + b3;
            sh[j] = o0 & 3;
            r0[j] = *(const gptr<const u64_a4>)(base + (o0 & ~3u));
            r1[j] = *(const gptr<const u64_a4>)(base + ((o0 + st) & ~3u));
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            // Row y + 1 starts stride bytes on, so its first sample sits at
            // byte (o0 + stride) & 3 of its own load.
            const uint32_t s1 = (sh[j] + st) & 3;
            const uint32_t t = __builtin_amdgcn_alignbyte((uint32_t)(r0[j] >> 32), (uint32_t)r0[j], sh[j]);
            const uint32_t u = __builtin_amdgcn_alignbyte((uint32_t)(r1[j] >> 32), (uint32_t)r1[j], s1);
            const float p00 = (float)(t & 0xff), p01 = (float)((t >> 8) & 0xff);
            const float p10 = (float)(u & 0xff), p11 = (float)((u >> 8) & 0xff);
            out[j] = PERSP ? blend_persp(p00, p01, p10, p11, c[j].ax, c[j].ay)
                           : blend_affine(p00, p01, p10, p11, c[j].ax, c[j].ay);
        }
        return;
    }
    const gptr<const uint8_t> base = gp(s.p);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int x0 = c[j].ix, y0 = c[j].iy;
        const bool vx0 = (unsigned)x0 < (unsigned)W, vx1 = (unsigned)(x0 + 1) < (unsigned)W;
        const bool vy0 = (unsigned)y0 < (unsigned)H, vy1 = (unsigned)(y0 + 1) < (unsigned)H;
        const uint32_t xa = (uint32_t)min(max(x0, 0), W - 1), xb = (uint32_t)min(max(x0 + 1, 0), W - 1);
        const uint32_t ya = (uint32_t)min(max(y0, 0), H - 1) * st, yb = (uint32_t)min(max(y0 + 1, 0), H - 1) * st;
        const float p00 = vx0 && vy0 ? (float)base[ya + xa] : cv;
        const float p01 = vx1 && vy0 ? (float)base[ya + xb] : cv;
        const float p10 = vx0 && vy1 ? (float)base[yb + xa] : cv;
        const float p11 = vx1 && vy1 ? (float)base[yb + xb] : cv;
        const float v = PERSP ? blend_persp(p00, p01, p10, p11, c[j].ax, c[j].ay)
                              : blend_affine(p00, p01, p10, p11, c[j].ax, c[j].ay);
        out[j] = x0 < -1 || y0 < -1 || x0 >= W || y0 >= H ? cv : v;
    }
}

// The same blends with every pixel's loads issued as soon as its own
// coordinates are known, so they fly while the next pixel's coordinates are
// computed (the coordinates of a projective map are most of the kernel's
// arithmetic, and gating the loads on a vote over all four pixels would
// hold every load until the last coordinate). Each pixel loads its two tap
// rows with buffer loads over the lim bytes from the dword below the image
// base: the hardware's range check returns zeros for an offset outside them
// (coordinates outside the image, or offsets wrapped in 32 bits), so the
// offsets need no clamping. A pixel whose footprint is not wholly inside
// (or whose second row's 8-byte load would pass the image's last byte) is
// redone afterwards by the byte-gather rules, by the lanes that have one.
// blk_issue computes the coordinates and issues the loads of one block,
// blk_finish blends them, so a wave can issue its next block's loads before
// it blends the current one. Needs W >= 2 and H >= 2.
struct BlkLoads {
    BlkCoord c[4];
    vu32x2 r0[4], r1[4];
    uint32_t o0[4];
    bool ok[4];
};

template <typename CoordFn>
__device__ __attribute__((always_inline)) inline void blk_issue(const DImg &s, uint32_t lim, CoordFn coord, BlkLoads &L) {
    const int W = s.w, H = s.h;
    const uint32_t st = (uint32_t)s.stride;
    const uint32_t b3 = (uint32_t)((uintptr_t)s.p & 3);
    const uint32_t lim8 = lim - 8;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(s.p - b3, 0, (int)lim, kWarpBufWord3);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        L.c[j] = coord(j);
        L.o0[j] = (uint32_t)L.c[j].iy * st + (uint32_t)L.c[j].ix + b3;
        const uint32_t o1 = L.o0[j] + st;
        L.ok[j] = (unsigned)L.c[j].ix < (unsigned)(W - 1) && (unsigned)L.c[j].iy < (unsigned)(H - 1) && o1 <= lim8;
        L.r0[j] = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, L.o0[j] & ~3u, 0, 0));
        L.r1[j] = __builtin_bit_cast(vu32x2, __builtin_amdgcn_raw_buffer_load_b64(rs, o1 & ~3u, 0, 0));
    }
}

template <bool PERSP>
__device__ __attribute__((always_inline)) inline void blk_finish(const DImg &s, float cv, const BlkLoads &L,
                                                                 float (&out)[4]) {
    const int W = s.w, H = s.h;
    const uint32_t st = (uint32_t)s.stride;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t t = __builtin_amdgcn_alignbyte(L.r0[j].y, L.r0[j].x, L.o0[j] & 3);
        const uint32_t u = __builtin_amdgcn_alignbyte(L.r1[j].y, L.r1[j].x, (L.o0[j] + st) & 3);
        const float p00 = (float)(t & 0xff), p01 = (float)((t >> 8) & 0xff);
        const float p10 = (float)(u & 0xff), p11 = (float)((u >> 8) & 0xff);
        out[j] = PERSP ? blend_persp(p00, p01, p10, p11, L.c[j].ax, L.c[j].ay)
                       : blend_affine(p00, p01, p10, p11, L.c[j].ax, L.c[j].ay);
    }
    const gptr<const uint8_t> img = gp(s.p);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        if (!L.ok[j]) {
            const int x0 = L.c[j].ix, y0 = L.c[j].iy;
            if (x0 < -1 || y0 < -1 || x0 >= W || y0 >= H) {
                out[j] = cv;
            } else {
                const bool vx0 = (unsigned)x0 < (unsigned)W, vx1 = (unsigned)(x0 + 1) < (unsigned)W;
                const bool vy0 = (unsigned)y0 < (unsigned)H, vy1 = (unsigned)(y0 + 1) < (unsigned)H;
                const uint32_t xa = (uint32_t)min(max(x0, 0), W - 1), xb = (uint32_t)min(max(x0 + 1, 0), W - 1);
                const uint32_t ya = (uint32_t)min(max(y0, 0), H - 1) * st, yb = (uint32_t)min(max(y0 + 1, 0), H - 1) * st;
                const float p00 = vx0 && vy0 ? (float)img[ya + xa] : cv;
                const float p01 = vx1 && vy0 ? (float)img[ya + xb] : cv;
                const float p10 = vx0 && vy1 ? (float)img[yb + xa] : cv;
                const float p11 = vx1 && vy1 ? (float)img[yb + xb] : cv;
                out[j] = PERSP ? blend_persp(p00, p01, p10, p11, L.c[j].ax, L.c[j].ay)
                               : blend_affine(p00, p01, p10, p11, L.c[j].ax, L.c[j].ay);
            }
        }
    }
}

// Blocks (64 x 4 pixels) of a wave in the eager kernels: T consecutive
// block rows, the next block's coordinates and loads issued before the
// current block is blended, so one block's arithmetic overlaps the other's
// loads (with one block per wave every wave computes, then waits, then
// blends at about the same time). make_coord(t) gives block t's coordinate
// function; store(t, v) writes its results.
template <bool PERSP, int T, typename MakeCoord, typename Store>
__device__ __attribute__((always_inline)) inline void blk_pipeline(const DImg &s, float cv, uint32_t lim, int rows_left,
                                                                   MakeCoord make_coord, Store store) {
    BlkLoads L[2];
    blk_issue(s, lim, make_coord(0), L[0]);
#pragma unroll
    for (int t = 0; t < T; t++) {
        if (t + 1 < T && t + 1 < rows_left) blk_issue(s, lim, make_coord(t + 1), L[(t + 1) & 1]);
        float v[4];
        blk_finish<PERSP>(s, cv, L[t & 1], v);
        store(t, v);
        if (t + 1 >= rows_left) break;
    }
}

// Rounds the lane's 4 results, transposes them through the wave's 256 LDS
// bytes (4 rows x 64 pixels) and stores 4 adjacent bytes per lane.
__device__ __attribute__((always_inline)) inline void blk_store(const DImg &d, int bx0, int by0, const float (&v)[4],
                                                                uint32_t *lds) {
    const int l = (int)(threadIdx.x & 63);
    uint8_t *row = reinterpret_cast<uint8_t *>(lds) + (l / kBlkW) * 64;
#pragma unroll
    for (int j = 0; j < 4; j++) row[kBlkW * j + (l % kBlkW)] = (uint8_t)pack_u8(v[j], 0, 0);
    __syncthreads();
    const uint32_t word = lds[l];
    const int y = by0 + (l >> 4), x = bx0 + 4 * (l & 15);
    if (y < d.h && x < d.w) store4_u8(d, x, y, word);
}

// blk_store for a wave that owns its LDS words: no workgroup barrier (LDS
// operations of one wave complete in order; the fences keep the compiler
// from moving the read above the writes, or the next block's writes above
// this read).
__device__ __attribute__((always_inline)) inline void blk_store_wave(const DImg &d, int bx0, int by0, const float (&v)[4],
                                                                     uint32_t *lds) {
    const int l = (int)(threadIdx.x & 63);
    uint8_t *row = reinterpret_cast<uint8_t *>(lds) + (l / kBlkW) * 64;
#pragma unroll
    for (int j = 0; j < 4; j++) row[kBlkW * j + (l % kBlkW)] = (uint8_t)pack_u8(v[j], 0, 0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t word = lds[l];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int y = by0 + (l >> 4), x = bx0 + 4 * (l & 15);
    if (y < d.h && x < d.w) store4_u8(d, x, y, word);
}

// Affine coordinates of pixel (x, y) (geometry.zig:432-494). The host runs
// this kernel only when |sx| and |sy| stay below 1e9 for every pixel, where
// toIdx's clamp changes nothing.
__device__ __attribute__((always_inline)) inline BlkCoord blk_affine_coord(const float *m, int x, float bx, float by) {
    const float xf = (float)x;
    const float sx = fmaf(m[0], xf, bx), sy = fmaf(m[3], xf, by);
    const float fx = floorf(sx), fy = floorf(sy);
    return BlkCoord{(int)fx, (int)fy, sx - fx, sy - fy};
}

// One tile of the batch's one-channel u8 bilinear constant-border
// warpAffine: 64 x 16 pixels, wave w the 64 x 4 block row w, sampled by the
// eager block path (every pixel's loads issued with its coordinates) and
// stored through the wave's own 64 LDS words. Needs the source at least 2 x
// 2, lim = blk_limit of the source, and coordinates bounded as for
// warp_affine_blk_pipe_u8.
__device__ __attribute__((always_inline)) inline void affine_blk_tile(const AffineParams &p, uint32_t lim, int bx, int by,
                                                                      uint32_t *lds) {
    const int wv = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int bx0 = bx * 4 * kBlkW, by0 = (by * 4 + wv) * kBlkH;
    if (by0 >= p.d.h) return;
    const float yf = (float)(by0 + l / kBlkW);
    const float bxv = p.m[1] * yf + p.m[2], byv = p.m[4] * yf + p.m[5];
    BlkLoads L;
    blk_issue(p.s, lim, [&](int j) { return blk_affine_coord(p.m, bx0 + kBlkW * j + l % kBlkW, bxv, byv); }, L);
    float v[4];
    blk_finish<false>(p.s, p.cv, L, v);
    blk_store_wave(p.d, bx0, by0, v, lds + 64 * wv);
}

// The block kernels' 8-byte gathers address the source with 32-bit offsets
// from the dword below its base: 0 if those do not fit.
inline uint32_t blk_limit(const DImg &s) {
    const uint64_t lim = ((uintptr_t)s.p & 3) + (uint64_t)(s.h - 1) * (uint64_t)s.stride + (uint64_t)s.w;
    return lim >= 8 && lim + 8 < (1ull << 32) ? (uint32_t)lim : 0;
}

// Whether |m0 x + m1 y + m2| and |m3 x + m4 y + m5| stay below 1e9 over the
// destination with every entry finite, so the affine coordinates need no clamp.
inline bool affine_bounded(const AffineParams &p) {
    for (int i = 0; i < 6; i++)
        if (!std::isfinite(p.m[i])) return false;
    const double w = p.d.w, h = p.d.h;
    const double bx = std::fabs(p.m[0]) * w + std::fabs(p.m[1]) * h + std::fabs(p.m[2]);
    const double by = std::fabs(p.m[3]) * w + std::fabs(p.m[4]) * h + std::fabs(p.m[5]);
    return bx < 1e8 && by < 1e8;
}

}  // namespace visar
