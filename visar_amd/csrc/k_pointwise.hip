// k_pointwise.hip -- streaming kernels: BGR->gray, threshold, the background
// update, exact-half / integer-area / nearest / bilinear / fractional-area
// resize, saturating subtraction and the 8-bit histogram.
//
// Each kernel maps one thread to a run of consecutive samples of one row so
// loads and stores are coalesced; the 16-pixel u8 paths use 128-bit loads
// when the row segment is 16-byte aligned.
#include <utility>
#include "tiles.hpp"
#include "resize_stream.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;


typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 128-bit load of streamed input. A plain (temporal) load: measured on a
// 3840x2160 BGR frame, the nontemporal hint made the gray kernel 8.5 us
// against 6.7 us with plain loads.
__device__ inline uint4 ld_stream(const void *p) {
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4 *)(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// ---------------------------------------------------------------- gray

// color.zig:126-129, 392-396: y = (b*3735 + g*19235 + r*9798 + 2^14) >> 15
__device__ inline uint32_t gray_fix(uint32_t b, uint32_t g, uint32_t r) {
    return (b * 3735u + g * 19235u + r * 9798u + (1u << 14)) >> 15;
}

// 16 pixels per thread; a 3-channel group is three 128-bit loads. SB is the
// blue channel's index (0 BGR, 2 RGB), fixed at compile time so every byte
// index below is a constant. When both images are contiguous the grid runs
// over the flat pixel range (no per-row indexing); otherwise blockIdx.y is
// the row.
template <int SCN, int SB, bool FLAT>
__global__ __launch_bounds__(kThreads) void gray_u8(DImg s, DImg d, long npix) {
    const long g = (long)blockIdx.x * kThreads + threadIdx.x;
    long x = g * 16;
    const uint8_t *sr;
    uint8_t *dr;
    long limit;
    if constexpr (FLAT) {
        if (x >= npix) return;
        sr = s.p + x * SCN;
        dr = d.p + x;
        limit = npix;
    } else {
        const int y = blockIdx.y;
        if (x >= s.w) return;
        sr = row_ptr<const uint8_t>(s, y) + x * SCN;
        dr = row_ptr<uint8_t>(d, y) + x;
        limit = s.w;
    }
    constexpr int BI = SB, RI = 2 - SB;
    if (x + 16 <= limit && aligned16(sr) && aligned16(dr)) {
        uint4 w[SCN];
#pragma unroll
        for (int k = 0; k < SCN; k++) w[k] = ld_stream(sr + 16 * k);
        uint32_t words[4 * SCN];
#pragma unroll
        for (int k = 0; k < SCN; k++) {
            words[4 * k + 0] = w[k].x;
            words[4 * k + 1] = w[k].y;
            words[4 * k + 2] = w[k].z;
            words[4 * k + 3] = w[k].w;
        }
        uint32_t o[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int px = q * 4 + k;
                const int ib = px * SCN + BI, ig = px * SCN + 1, ir = px * SCN + RI;
                uint32_t b = (words[ib >> 2] >> ((ib & 3) * 8)) & 255u;
                uint32_t gg = (words[ig >> 2] >> ((ig & 3) * 8)) & 255u;
                uint32_t r = (words[ir >> 2] >> ((ir & 3) * 8)) & 255u;
                v |= gray_fix(b, gg, r) << (8 * k);
            }
            o[q] = v;
        }
        *reinterpret_cast<uint4 *>(dr) = make_uint4(o[0], o[1], o[2], o[3]);
        return;
    }
    const long n = min(16L, limit - x);
    for (long i = 0; i < n; i++) dr[i] = (uint8_t)gray_fix(sr[i * SCN + BI], sr[i * SCN + 1], sr[i * SCN + RI]);
}

// The same conversion with each wave reading one contiguous block of 1024
// pixels: load instruction k of the wave covers bytes [1024 k, 1024 k +
// 1024) of the block's SCN KiB (one nontemporal 16-byte load per lane, so a
// load instruction touches 8 whole 128-byte lines instead of strided pieces
// of 24 or 32), the wave regroups the bytes per lane through its own part of
// LDS, and lane i converts and stores pixels [16 i, 16 i + 16) of the block.
// On a 3840x2160 BGR frame: 6.6-6.8 us against 7.1-7.4 us for gray_u8 (the
// strided form with nontemporal loads ran 8.5 us). The last, partial block
// takes the per-pixel path. Needs contiguous images with 16-byte aligned
// bases.
constexpr int kGrayWaveThreads = 512;
template <int SCN, int SB>
__global__ __launch_bounds__(kGrayWaveThreads) void gray_u8_wave(DImg s, DImg d, long npix) {
    __shared__ u32x4 sh[kGrayWaveThreads / 64][64 * SCN];
    const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
    const long x0 = ((long)blockIdx.x * (kGrayWaveThreads / 64) + wv) * 1024;  // the wave's first pixel
    if (x0 >= npix) return;
    const uint8_t *sp = s.p + x0 * SCN;
    uint8_t *dp = d.p + x0;
    constexpr int BI = SB, RI = 2 - SB;
    if (x0 + 1024 > npix) {
        const long x = 16 * lane, n = min(16L, npix - x0 - x);
        for (long i = 0; i < n; i++)
            dp[x + i] = (uint8_t)gray_fix(sp[(x + i) * SCN + BI], sp[(x + i) * SCN + 1], sp[(x + i) * SCN + RI]);
        return;
    }
    u32x4 a[SCN];
#pragma unroll
    for (int k = 0; k < SCN; k++)
        a[k] = __builtin_nontemporal_load((const __attribute__((address_space(1))) u32x4 *)(sp + 1024 * k + 16 * lane));
#pragma unroll
    for (int k = 0; k < SCN; k++) sh[wv][64 * k + lane] = a[k];
    // The wave reads back what its other lanes wrote: LDS operations of one
    // wave complete in order; the fence keeps the compiler from moving the
    // reads above the writes.
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t words[4 * SCN];
#pragma unroll
    for (int k = 0; k < SCN; k++) {
        const u32x4 w = sh[wv][SCN * lane + k];
        words[4 * k + 0] = w.x;
        words[4 * k + 1] = w.y;
        words[4 * k + 2] = w.z;
        words[4 * k + 3] = w.w;
    }
    uint32_t o[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int px = q * 4 + k;
            const int ib = px * SCN + BI, ig = px * SCN + 1, ir = px * SCN + RI;
            const uint32_t b = (words[ib >> 2] >> ((ib & 3) * 8)) & 255u;
            const uint32_t gg = (words[ig >> 2] >> ((ig & 3) * 8)) & 255u;
            const uint32_t r = (words[ir >> 2] >> ((ir & 3) * 8)) & 255u;
            v |= gray_fix(b, gg, r) << (8 * k);
        }
        o[q] = v;
    }
    *(__attribute__((address_space(1))) u32x4 *)(dp + 16 * lane) = u32x4{o[0], o[1], o[2], o[3]};
}

// videoio/frame.zig:167-177, 292-336 grayDownsample: the k x k block sums of
// B, G and R weighted 1868 / 9617 / 4899; k = 2 rounds (l + 2^15) >> 16,
// other k take (f32)l * (f32)(1 / (k*k*2^14)) + 0.5 truncated, at most 255.
// The weighted sum is linear, so l is also the sum over the block of each
// pixel's b*1868 + g*9617 + r*4899 (exact in u32: l < 256 * 255 * 2^14 < 2^30).
__device__ __attribute__((always_inline)) inline uint8_t gray_down_out(uint32_t l, int k, float inv) {
    if (k == 2) return (uint8_t)((l + (1u << 15)) >> 16);
    const float v = (float)l * inv + 0.5f;
    return (uint8_t)(uint32_t)(v < 255.0f ? v : 255.0f);
}

// Any layout: one output pixel per thread, byte loads; 64 x 4 outputs per
// tile, tiles walked grid-stride.
template <int CN>
__global__ __launch_bounds__(kThreads) void gray_down_u8(DImg s, DImg d, int k, float inv, int tiles_x, long tiles) {
    for (long t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int ox = (int)(t % tiles_x) * 64 + (threadIdx.x & 63);
        const long oy = t / tiles_x * 4 + (threadIdx.x >> 6);
        if (ox >= d.w || oy >= d.h) continue;
        uint32_t l = 0;
        for (int dy = 0; dy < k; dy++) {
            const gptr<const uint8_t> row = grow<const uint8_t>(s, oy * k + dy) + (long)ox * k * CN;
            for (int dx = 0; dx < k; dx++)
                l += row[dx * CN] * 1868u + row[dx * CN + 1] * 9617u + row[dx * CN + 2] * 4899u;
        }
        grow<uint8_t>(d, oy)[ox] = gray_down_out(l, k, inv);
    }
}

// Rows 4-byte (BGR) or 16-byte (BGRA) aligned: a workgroup takes one output
// row and up to kGdPix of its source pixels. Each thread reads one 4-pixel
// chunk of each of the K rows with nontemporal vector loads (measured a
// little faster here than plain loads, unlike the gray kernel) and sums
// every pixel's
// weighted luma down the rows with v_dot4_u32_u8 (the weights split into
// high and low bytes: 1868 = 7*256 + 76, 9617 = 37*256 + 145,
// 4899 = 19*256 + 35); the column sums go to LDS, and each output adds K
// of them.
constexpr int kGdPix = 1020;  // <= 4 * 256 - 3: at most 256 chunks per unit
constexpr uint32_t kGdHi = 7u | 37u << 8 | 19u << 16, kGdLo = 76u | 145u << 8 | 35u << 16;

template <int CN, int K>
__global__ __launch_bounds__(kThreads) void gray_down_seg(DImg s, DImg d, int per_unit, uint32_t units_x, uint32_t units,
                                                          float inv) {
    __shared__ uint32_t col[4 * kThreads];
    const int tid = threadIdx.x;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const long oy = u / units_x;
        const int ox0 = (int)(u % units_x) * per_unit;
        const int nout = min(per_unit, d.w - ox0);
        const int p0 = ox0 * K, p1 = p0 + nout * K;  // source pixels [p0, p1)
        const int c0 = p0 >> 2, nch = ((p1 + 3) >> 2) - c0;
        const int c = c0 + tid;
        if (u != blockIdx.x) __syncthreads();  // the previous unit's reads of col
        if (tid < nch) {
            uint32_t hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0};
            if (4 * c + 4 <= s.w) {
                if constexpr (CN == 4) {
                    u32x4 v[K];
#pragma unroll
                    for (int dy = 0; dy < K; dy++)
                        v[dy] = __builtin_nontemporal_load((gptr<const u32x4>)(grow<const uint8_t>(s, oy * K + dy) + 16l * c));
#pragma unroll
                    for (int dy = 0; dy < K; dy++)
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            hi[j] = __builtin_amdgcn_udot4(v[dy][j], kGdHi, hi[j], false);
                            lo[j] = __builtin_amdgcn_udot4(v[dy][j], kGdLo, lo[j], false);
                        }
                } else {
                    // d0 = B0 G0 R0 B1, d1 = G1 R1 B2 G2, d2 = R2 B3 G3 R3 (bytes, low first)
                    uint32_t v[K][3];
#pragma unroll
                    for (int dy = 0; dy < K; dy++) {
                        const gptr<const uint32_t> q = (gptr<const uint32_t>)(grow<const uint8_t>(s, oy * K + dy) + 12l * c);
                        v[dy][0] = __builtin_nontemporal_load(q);
                        v[dy][1] = __builtin_nontemporal_load(q + 1);
                        v[dy][2] = __builtin_nontemporal_load(q + 2);
                    }
                    const uint32_t H = kGdHi, L = kGdLo;
#pragma unroll
                    for (int dy = 0; dy < K; dy++) {
                        const uint32_t d0 = v[dy][0], d1 = v[dy][1], d2 = v[dy][2];
                        hi[0] = __builtin_amdgcn_udot4(d0, H, hi[0], false);
                        lo[0] = __builtin_amdgcn_udot4(d0, L, lo[0], false);
                        hi[1] = __builtin_amdgcn_udot4(d1, H >> 8, __builtin_amdgcn_udot4(d0, H << 24, hi[1], false), false);
                        lo[1] = __builtin_amdgcn_udot4(d1, L >> 8, __builtin_amdgcn_udot4(d0, L << 24, lo[1], false), false);
                        hi[2] = __builtin_amdgcn_udot4(d2, H >> 16, __builtin_amdgcn_udot4(d1, H << 16, hi[2], false), false);
                        lo[2] = __builtin_amdgcn_udot4(d2, L >> 16, __builtin_amdgcn_udot4(d1, L << 16, lo[2], false), false);
                        hi[3] = __builtin_amdgcn_udot4(d2, H << 8, hi[3], false);
                        lo[3] = __builtin_amdgcn_udot4(d2, L << 8, lo[3], false);
                    }
                }
            } else {
                // The chunk crosses the row end: its pixels inside the row only.
                for (int dy = 0; dy < K; dy++) {
                    const gptr<const uint8_t> row = grow<const uint8_t>(s, oy * K + dy);
                    for (int j = 0; j < 4 && 4 * c + j < s.w; j++) {
                        const gptr<const uint8_t> px = row + (long)(4 * c + j) * CN;
                        lo[j] += px[0] * 1868u + px[1] * 9617u + px[2] * 4899u;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) col[4 * tid + j] = (hi[j] << 8) + lo[j];
        }
        __syncthreads();
        const int base = p0 - 4 * c0;
        const gptr<uint8_t> out = grow<uint8_t>(d, oy) + ox0;
        for (int j = tid; j < nout; j += kThreads) {
            uint32_t l = 0;
#pragma unroll
            for (int t = 0; t < K; t++) l += col[base + j * K + t];
            out[j] = gray_down_out(l, K, inv);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void gray_generic(DImg s, DImg d, int scn, int sb) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / s.w, x = gid - y * s.w;
    if (y >= s.h) return;
    const T *p = row_ptr<const T>(s, y) + x * scn;
    T *o = row_ptr<T>(d, y) + x;
    const int bi = sb, ri = 2 - sb;
    if constexpr (std::is_same<T, float>::value) {
        *o = p[bi] * 0.114f + p[1] * 0.587f + p[ri] * 0.299f;
    } else {
        *o = (T)gray_fix(p[bi], p[1], p[ri]);
    }
}

// ----------------------------------------------------------- threshold

// threshold.zig:85-119 (threshOne) and the u8 LUT (threshold.zig:44-60).
template <typename T>
__global__ __launch_bounds__(kThreads) void threshold_k(ThreshParams p, long n_per_row) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / n_per_row, i = gid - y * n_per_row;
    if (y >= p.s.h) return;
    T v = row_ptr<const T>(p.s, y)[i];
    T o;
    if constexpr (std::is_same<T, uint8_t>::value) {
        o = p.lut[v];
    } else {
        T t = (T)p.t, m = (T)p.m;
        bool gt = p.below || v > t;
        switch (p.kind) {
        case 0: o = gt ? m : (T)0; break;
        case 1: o = gt ? (T)0 : m; break;
        case 2: o = gt ? (p.below ? (T)0 : t) : v; break;
        case 3: o = gt ? v : (T)0; break;
        default: o = gt ? (T)0 : v; break;
        }
    }
    row_ptr<T>(p.d, y)[i] = o;
}

// u8 threshold on contiguous images, 16 samples per thread: the LUT of
// threshold.zig:44-60 is "v > ti" with ti = floor(thresh) as an integer, so
// the bytes are compared directly (ti may lie outside 0..254).
__global__ __launch_bounds__(kThreads) void threshold_u8_flat(const uint8_t *s, uint8_t *d, long n, int ti, int m,
                                                              int kind) {
    const long x = ((long)blockIdx.x * kThreads + threadIdx.x) * 16;
    if (x >= n) return;
    const int tc = ti < 0 ? 0 : ti > 255 ? 255 : ti;
    auto one = [&](uint32_t v) -> uint32_t {
        const bool gt = (int)v > ti;
        switch (kind) {
        case 0: return gt ? (uint32_t)m : 0u;
        case 1: return gt ? 0u : (uint32_t)m;
        case 2: return gt ? (uint32_t)tc : v;
        case 3: return gt ? v : 0u;
        default: return gt ? 0u : v;
        }
    };
    if (x + 16 <= n && aligned16(s + x) && aligned16(d + x)) {
        uint4 w = *reinterpret_cast<const uint4 *>(s + x);
        uint32_t in[4] = {w.x, w.y, w.z, w.w}, o[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t r = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) r |= one((in[q] >> (8 * b)) & 255u) << (8 * b);
            o[q] = r;
        }
        *reinterpret_cast<uint4 *>(d + x) = make_uint4(o[0], o[1], o[2], o[3]);
        return;
    }
    for (long i = x; i < n && i < x + 16; i++) d[i] = (uint8_t)one(s[i]);
}

// ---------------------------------------------------------- background

// track.zig:33-81. Model u16 in 8.8 fixed point, per pixel in i32.
__global__ __launch_bounds__(kThreads) void background_k(DImg f, DImg m, uint16_t *model, long mstride, int shift,
                                                         int thr, int prime) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / f.w, x = gid - y * f.w;
    if (y >= f.h) return;
    int32_t fv = row_ptr<const uint8_t>(f, y)[x];
    uint16_t *bp = reinterpret_cast<uint16_t *>(reinterpret_cast<uint8_t *>(model) + y * mstride) + x;
    uint8_t *mp = row_ptr<uint8_t>(m, y) + x;
    if (prime) {
        *bp = (uint16_t)(fv << 8);
        *mp = 0;
        return;
    }
    int32_t bv = *bp;
    int32_t diff = fv - (bv >> 8);
    bool fg = (diff < 0 ? -diff : diff) > thr;
    *mp = fg ? 255 : 0;
    int sh = shift + (fg ? 2 : 0);
    *bp = (uint16_t)(bv + (((fv << 8) - bv) >> sh));
}

// The same update on contiguous frame / mask / model, 16 pixels per thread
// (128-bit frame and mask accesses, 2 x 128-bit model accesses).
__global__ __launch_bounds__(kThreads) void background_flat(const uint8_t *f, uint8_t *m, uint16_t *model, long n,
                                                            int shift, int thr, int prime) {
    const long x = ((long)blockIdx.x * kThreads + threadIdx.x) * 16;
    if (x >= n) return;
    if (x + 16 <= n && aligned16(f + x) && aligned16(m + x) && aligned16(model + x)) {
        uint4 fw = *reinterpret_cast<const uint4 *>(f + x);
        uint4 b0 = reinterpret_cast<const uint4 *>(model + x)[0], b1 = reinterpret_cast<const uint4 *>(model + x)[1];
        uint32_t fv[4] = {fw.x, fw.y, fw.z, fw.w};
        uint32_t bw[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        uint32_t mo[4] = {0, 0, 0, 0}, bo[8];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int32_t fi = (int32_t)((fv[i >> 2] >> (8 * (i & 3))) & 255u);
            const int32_t bi = (int32_t)((bw[i >> 1] >> (16 * (i & 1))) & 0xffffu);
            int32_t nb;
            uint32_t mk;
            if (prime) {
                nb = fi << 8;
                mk = 0;
            } else {
                const int32_t diff = fi - (bi >> 8);
                const bool fg = (diff < 0 ? -diff : diff) > thr;
                mk = fg ? 255u : 0u;
                nb = bi + (((fi << 8) - bi) >> (shift + (fg ? 2 : 0)));
            }
            mo[i >> 2] |= mk << (8 * (i & 3));
            if (i & 1) bo[i >> 1] |= ((uint32_t)nb & 0xffffu) << 16;
            else bo[i >> 1] = (uint32_t)nb & 0xffffu;
        }
        *reinterpret_cast<uint4 *>(m + x) = make_uint4(mo[0], mo[1], mo[2], mo[3]);
        reinterpret_cast<uint4 *>(model + x)[0] = make_uint4(bo[0], bo[1], bo[2], bo[3]);
        reinterpret_cast<uint4 *>(model + x)[1] = make_uint4(bo[4], bo[5], bo[6], bo[7]);
        return;
    }
    for (long i = x; i < n && i < x + 16; i++) {
        const int32_t fi = f[i], bi = model[i];
        if (prime) {
            model[i] = (uint16_t)(fi << 8);
            m[i] = 0;
            continue;
        }
        const int32_t diff = fi - (bi >> 8);
        const bool fg = (diff < 0 ? -diff : diff) > thr;
        m[i] = fg ? 255 : 0;
        model[i] = (uint16_t)(bi + (((fi << 8) - bi) >> (shift + (fg ? 2 : 0))));
    }
}

// -------------------------------------------------------------- resize

// Resize tiles (tiles.hpp): 256 output samples x 4 rows per workgroup, the
// one-channel exact half 1024 x 4.
__global__ __launch_bounds__(kThreads) void half_u8(DImg s, DImg d) { half_tile(s, d, blockIdx.x, blockIdx.y); }
__global__ __launch_bounds__(kThreads) void half_u8_c1(DImg s, DImg d) { half_c1_tile(s, d, blockIdx.x, blockIdx.y); }
__global__ __launch_bounds__(kThreads) void area_int_u8(DImg s, DImg d, int ix, int iy, float scale) {
    area_int_tile(s, d, ix, iy, scale, blockIdx.x, blockIdx.y);
}

// resize.zig:62-106 nearest: xofs = min(floor(x * sw/dw), sw - 1) in f64.
__global__ __launch_bounds__(kThreads) void nearest_k(DImg s, DImg d, int elem, double ifx, double ify) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / d.w, x = gid - y * d.w;
    if (y >= d.h) return;
    long sy = (long)floor((double)y * ify);
    if (sy > s.h - 1) sy = s.h - 1;
    long sx = (long)floor((double)x * ifx);
    if (sx > s.w - 1) sx = s.w - 1;
    const uint8_t *sp = s.p + sy * s.stride + sx * (long)s.cn * elem;
    uint8_t *dp = d.p + y * d.stride + x * (long)d.cn * elem;
    for (long i = 0; i < (long)d.cn * elem; i++) dp[i] = sp[i];
}

__global__ __launch_bounds__(kThreads) void linear_u8(DImg s, DImg d, const LinTap *xt, const LinTap *yt) {
    linear_tile(s, d, xt, yt, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kThreads) void linear_box_u8(DImg s, DImg d, const LinTap *xt, const LinTap *yt) {
    __shared__ uint32_t lds[kLinLdsWords];
    linear_box_tile(s, d, xt, yt, blockIdx.x, blockIdx.y, lds);
}

__global__ __launch_bounds__(kThreads) void area_frac_u8(DImg s, DImg d, const AreaEntry *xt, const int32_t *xs,
                                                         const AreaEntry *yt, const int32_t *ys) {
    area_frac_tile(s, d, xt, xs, yt, ys, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kThreads) void area_box_u8(DImg s, DImg d, const AreaEntry *xt, const int32_t *xs,
                                                        const AreaEntry *yt, const int32_t *ys) {
    __shared__ uint32_t lds[kAreaLdsWords];
    area_box_tile(s, d, xt, xs, yt, ys, blockIdx.x, blockIdx.y, lds);
}

// ----------------------------------------------------------- morph helpers

template <typename T>
__global__ __launch_bounds__(kThreads) void sub_sat_k(DImg a, DImg b, DImg d, long n_per_row) {
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / n_per_row, i = gid - y * n_per_row;
    if (y >= d.h) return;
    T x = row_ptr<const T>(a, y)[i], z = row_ptr<const T>(b, y)[i];
    T o;
    if constexpr (std::is_same<T, float>::value) o = x - z;
    else o = x > z ? (T)(x - z) : (T)0;
    row_ptr<T>(d, y)[i] = o;
}

// ------------------------------------------------------------ histogram

__global__ __launch_bounds__(kThreads) void hist_u8(DImg s, uint32_t *hist, long n_per_row) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long total = n_per_row * s.h;
    for (long gid = (long)blockIdx.x * kThreads + threadIdx.x; gid < total; gid += (long)gridDim.x * kThreads) {
        long y = gid / n_per_row, i = gid - y * n_per_row;
        atomicAdd(&h[row_ptr<const uint8_t>(s, y)[i]], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

}  // namespace

int launch_gray(const DImg &s, const DImg &d, int depth, int scn, int sb, hipStream_t st) {
    if (depth == 0) {
        const bool flat = s.stride == (long)s.w * scn && d.stride == (long)d.w;
        const long npix = (long)s.w * s.h;
        if (flat && npix >= 1024 && tuning().gray_layout == 0 && (((uintptr_t)s.p | (uintptr_t)d.p) & 15) == 0) {
            const dim3 grid((unsigned)((npix + 1024L * (kGrayWaveThreads / 64) - 1) / (1024L * (kGrayWaveThreads / 64))));
            if (scn == 3 && sb == 0) hipLaunchKernelGGL((gray_u8_wave<3, 0>), grid, dim3(kGrayWaveThreads), 0, st, s, d, npix);
            else if (scn == 3) hipLaunchKernelGGL((gray_u8_wave<3, 2>), grid, dim3(kGrayWaveThreads), 0, st, s, d, npix);
            else if (sb == 0) hipLaunchKernelGGL((gray_u8_wave<4, 0>), grid, dim3(kGrayWaveThreads), 0, st, s, d, npix);
            else hipLaunchKernelGGL((gray_u8_wave<4, 2>), grid, dim3(kGrayWaveThreads), 0, st, s, d, npix);
            return launched();
        }
        dim3 grid = flat ? dim3(blocks_for((npix + 15) / 16)) : dim3((unsigned)((s.w + 16 * kThreads - 1) / (16 * kThreads)), (unsigned)s.h);
#define GRAY(SCN, SB)                                                                                       \
    if (flat) hipLaunchKernelGGL((gray_u8<SCN, SB, true>), grid, dim3(kThreads), 0, st, s, d, npix); \
    else hipLaunchKernelGGL((gray_u8<SCN, SB, false>), grid, dim3(kThreads), 0, st, s, d, npix);
        if (scn == 3 && sb == 0) { GRAY(3, 0) }
        else if (scn == 3) { GRAY(3, 2) }
        else if (sb == 0) { GRAY(4, 0) }
        else { GRAY(4, 2) }
#undef GRAY
        return launched();
    }
    long n = (long)s.w * s.h;
    if (depth == 2) hipLaunchKernelGGL(gray_generic<uint16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, d, scn, sb);
    else if (depth == 5) hipLaunchKernelGGL(gray_generic<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, d, scn, sb);
    else return UNSUPPORTED;
    return launched();
}

int launch_threshold(const ThreshParams &p, hipStream_t st) {
    long npr = (long)p.s.w * p.s.cn;
    if (p.depth == 0 && p.s.stride == npr && p.d.stride == npr) {
        const long n = npr * p.s.h;
        hipLaunchKernelGGL(threshold_u8_flat, dim3(blocks_for((n + 15) / 16)), dim3(kThreads), 0, st, p.s.p, p.d.p, n,
                           p.ti, p.mi, p.kind);
        return launched();
    }
    long n = npr * p.s.h;
    switch (p.depth) {
    case 0: hipLaunchKernelGGL(threshold_k<uint8_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p, npr); break;
    case 2: hipLaunchKernelGGL(threshold_k<uint16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p, npr); break;
    case 3: hipLaunchKernelGGL(threshold_k<int16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p, npr); break;
    case 5: hipLaunchKernelGGL(threshold_k<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st, p, npr); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_background(const DImg &frame, const DImg &mask, uint16_t *model, long model_stride, int shift, int thr,
                      int prime, hipStream_t st) {
    long n = (long)frame.w * frame.h;
    if (frame.stride == frame.w && mask.stride == mask.w && model_stride == 2L * frame.w) {
        hipLaunchKernelGGL(background_flat, dim3(blocks_for((n + 15) / 16)), dim3(kThreads), 0, st, frame.p, mask.p, model,
                           n, shift, thr, prime);
        return launched();
    }
    hipLaunchKernelGGL(background_k, dim3(blocks_for(n)), dim3(kThreads), 0, st, frame, mask, model, model_stride, shift,
                       thr, prime);
    return launched();
}

int launch_copy(const DImg &s, const DImg &d, size_t elem, hipStream_t st) {
    size_t rl = (size_t)s.w * s.cn * elem;
    return hipMemcpy2DAsync(d.p, d.stride, s.p, s.stride, rl, s.h, hipMemcpyDeviceToDevice, st) == hipSuccess ? OK
                                                                                                               : INTERNAL;
}

inline dim3 tiles256(const DImg &d) { return dim3((unsigned)(((long)d.w * d.cn + 255) / 256), (unsigned)((d.h + 3) / 4)); }

template <int CN, int K>
static void launch_gd_seg(const DImg &s, const DImg &d, float inv, hipStream_t st) {
    const int per = kGdPix / K;
    const uint32_t units_x = (uint32_t)((d.w + per - 1) / per), units = units_x * (uint32_t)d.h;
    const unsigned grid = std::min(units, 1u << 20);
    hipLaunchKernelGGL((gray_down_seg<CN, K>), dim3(grid), dim3(kThreads), 0, st, s, d, per, units_x, units, inv);
}

template <int CN, int... Ks>
static void launch_gd_k(const DImg &s, const DImg &d, int k, float inv, hipStream_t st,
                        std::integer_sequence<int, Ks...>) {
    ((k == Ks + 1 ? launch_gd_seg<CN, Ks + 1>(s, d, inv, st) : void()), ...);
}

int launch_gray_down(const DImg &s, const DImg &d, int k, hipStream_t st) {
    if (k < 1 || k > 16 || (s.cn != 3 && s.cn != 4)) return UNSUPPORTED;
    const float inv = 1.0f / (float)(k * k * (1 << 14));
    const unsigned align = s.cn == 4 ? 16 : 4;
    // Units of one output row and up to kGdPix / k of its pixels, counted in
    // 32 bits (an image that fits in memory has fewer than 2^32).
    const uint64_t units = (uint64_t)((d.w + kGdPix / k - 1) / (kGdPix / k)) * (uint64_t)d.h;
    if ((uintptr_t)s.p % align == 0 && s.stride % align == 0 && units < (1ull << 32)) {
        if (s.cn == 4) launch_gd_k<4>(s, d, k, inv, st, std::make_integer_sequence<int, 16>());
        else launch_gd_k<3>(s, d, k, inv, st, std::make_integer_sequence<int, 16>());
        return launched();
    }
    const int tiles_x = (d.w + 63) / 64;
    const long tiles = (long)tiles_x * ((d.h + 3) / 4);
    const unsigned grid = (unsigned)std::min<long>(tiles, 1l << 20);
    if (s.cn == 3) hipLaunchKernelGGL((gray_down_u8<3>), dim3(grid), dim3(kThreads), 0, st, s, d, k, inv, tiles_x, tiles);
    else hipLaunchKernelGGL((gray_down_u8<4>), dim3(grid), dim3(kThreads), 0, st, s, d, k, inv, tiles_x, tiles);
    return launched();
}

int launch_half_u8(const DImg &s, const DImg &d, hipStream_t st) {
    if (d.cn == 1) {
        hipLaunchKernelGGL(half_u8_c1, dim3((unsigned)((d.w + 1023) / 1024), (unsigned)((d.h + 3) / 4)), dim3(kThreads), 0, st,
                           s, d);
        return launched();
    }
    hipLaunchKernelGGL(half_u8, tiles256(d), dim3(kThreads), 0, st, s, d);
    return launched();
}

int launch_area_int_u8(const DImg &s, const DImg &d, int ix, int iy, hipStream_t st) {
    float scale = 1.0f / (float)(ix * iy);
    hipLaunchKernelGGL(area_int_u8, tiles256(d), dim3(kThreads), 0, st, s, d, ix, iy, scale);
    return launched();
}

int launch_nearest(const DImg &s, const DImg &d, size_t elem, hipStream_t st) {
    long n = (long)d.w * d.h;
    double ifx = (double)s.w / (double)d.w, ify = (double)s.h / (double)d.h;
    hipLaunchKernelGGL(nearest_k, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, d, (int)elem, ifx, ify);
    return launched();
}

bool linear_stream_args(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, int tb, int run, LinStreamArgs &a) {
    if (s.stride < 0 || d.stride < 0) return false;
    a.off_s = (uint32_t)((uintptr_t)s.p & 3);
    a.off_d = (uint32_t)((uintptr_t)d.p & 3);
    const long lim_s = a.off_s + (long)(s.h - 1) * s.stride + (long)s.w * s.cn;
    const long lim_d = a.off_d + (long)(d.h - 1) * d.stride + (long)d.w * d.cn;
    // Source rows of at least 12 bytes, so only the last row's 12-byte
    // windows can reach the end of the range; the dword holding the image's
    // last byte is readable whole.
    if (lim_s >= (1L << 31) - 4 || lim_d >= (1L << 31) || (long)s.w * s.cn < 12) return false;
    a.s = s;
    a.s.p = s.p - a.off_s;
    a.d = d;
    a.d.p = d.p - a.off_d;
    a.xt = xt;
    a.yt = yt;
    a.lim_s = (uint32_t)((lim_s + 3) & ~3L);
    a.lim_d = (uint32_t)lim_d;
    const long n = (long)d.w * d.cn;
    a.strips = (int)((n + kLinStreamStrip) / kLinStreamStrip);
    // 8-row bands where they still give about 4000 waves, one band per
    // wave (measured from 1024^2 to 8192^2 sources: runs of 8-row bands need
    // more scalar registers than a wave has, and runs of 4-row bands issue
    // fewer loads at a time).
    if (tb != 4 && tb != 8) tb = a.strips * (long)((d.h + 7) / 8) >= 4000 ? 8 : 4;
    a.tb = tb;
    a.bands = (d.h + tb - 1) / tb;
    a.run = run > 0 ? run : 1;
    return true;
}

template <int TB, bool RUN>
static void launch_stream_tb(const LinStreamArgs &a, bool ad, bool as, dim3 grid, hipStream_t st) {
    if (ad && as) hipLaunchKernelGGL((lstream::linear_stream_u8<TB, true, true, RUN>), grid, dim3(256), 0, st, a);
    else if (ad) hipLaunchKernelGGL((lstream::linear_stream_u8<TB, true, false, RUN>), grid, dim3(256), 0, st, a);
    else if (as) hipLaunchKernelGGL((lstream::linear_stream_u8<TB, false, true, RUN>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((lstream::linear_stream_u8<TB, false, false, RUN>), grid, dim3(256), 0, st, a);
}

static bool launch_linear_stream(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, hipStream_t st) {
    LinStreamArgs a;
    if (!linear_stream_args(s, d, xt, yt, tuning().lin_rows, tuning().lin_run, a)) return false;
    const TileOrder o = tile_order(a.strips, (a.bands + 4 * a.run - 1) / (4 * a.run), true);
    a.ntiles = o.ntiles;
    a.xcd = o.xcd;
    const dim3 grid(tile_grid(o.ntiles));
    const bool ad = a.off_d == 0 && d.stride % 4 == 0, as = a.off_s == 0 && s.stride % 4 == 0;
    if (a.tb == 8 && a.run == 1) launch_stream_tb<8, false>(a, ad, as, grid, st);
    else if (a.tb == 8) launch_stream_tb<8, true>(a, ad, as, grid, st);
    else if (a.run == 1) launch_stream_tb<4, false>(a, ad, as, grid, st);
    else launch_stream_tb<4, true>(a, ad, as, grid, st);
    return true;
}

int launch_linear_u8(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, hipStream_t st, bool box,
                     bool stream) {
    if (stream && launch_linear_stream(s, d, xt, yt, st)) return launched();
    if (box) {
        const dim3 grid((unsigned)(((long)d.w * d.cn + 255) / 256), (unsigned)((d.h + kLinTH - 1) / kLinTH));
        hipLaunchKernelGGL(linear_box_u8, grid, dim3(kThreads), 0, st, s, d, xt, yt);
    } else {
        hipLaunchKernelGGL(linear_u8, tiles256(d), dim3(kThreads), 0, st, s, d, xt, yt);
    }
    return launched();
}

int launch_area_frac_u8(const DImg &s, const DImg &d, const AreaEntry *xt, const int32_t *xstart, const AreaEntry *yt,
                        const int32_t *ystart, hipStream_t st, bool box) {
    if (box) {
        const dim3 grid((unsigned)(((long)d.w * d.cn + 255) / 256), (unsigned)((d.h + kAreaTH - 1) / kAreaTH));
        hipLaunchKernelGGL(area_box_u8, grid, dim3(kThreads), 0, st, s, d, xt, xstart, yt, ystart);
    } else {
        hipLaunchKernelGGL(area_frac_u8, tiles256(d), dim3(kThreads), 0, st, s, d, xt, xstart, yt, ystart);
    }
    return launched();
}

int launch_sub_sat(const DImg &a, const DImg &b, const DImg &d, int depth, hipStream_t st) {
    long npr = (long)d.w * d.cn;
    long n = npr * d.h;
    switch (depth) {
    case 0: hipLaunchKernelGGL(sub_sat_k<uint8_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, a, b, d, npr); break;
    case 2: hipLaunchKernelGGL(sub_sat_k<uint16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, a, b, d, npr); break;
    case 3: hipLaunchKernelGGL(sub_sat_k<int16_t>, dim3(blocks_for(n)), dim3(kThreads), 0, st, a, b, d, npr); break;
    case 5: hipLaunchKernelGGL(sub_sat_k<float>, dim3(blocks_for(n)), dim3(kThreads), 0, st, a, b, d, npr); break;
    default: return UNSUPPORTED;
    }
    return launched();
}

int launch_hist_u8(const DImg &s, uint32_t *hist, hipStream_t st) {
    long npr = (long)s.w * s.cn;
    long n = npr * s.h;
    unsigned blocks = blocks_for(n);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(hist_u8, dim3(blocks), dim3(kThreads), 0, st, s, hist, npr);
    return launched();
}

}  // namespace visar
