// k_serve.hip -- the served front end's per-frame pipeline as one kernel.
//
// One workgroup owns a 256 x 16 tile of one frame of the batch and runs, in
// LDS, the chain of plain calls the reference detector makes
// (demos/track_lib.zig:533-544), bit for bit:
//   B = GaussianBlur 5x5, sigma 0, BORDER_REFLECT_101 (filter.zig:498-547:
//       exact integer sums, (sum + 2^15) >> 16)
//   M = Background.apply(B) (track.zig:33-81), model read from one buffer
//       and written to the other
//   T = threshold(M, mthr, 255, THRESH_BINARY) (threshold.zig:25-82)
//   mask = open(T) with a 3 x 3 rect (morph.zig:532-571), whose cells
//       outside the image are ignored (morph.zig:257)
// T is {0, 255}, so the erode is a bytewise AND and the dilate a bytewise OR
// of the neighbours, the ignored cells being 255 for the erode and 0 for the
// dilate. The tile computes M on its rows and columns +-2 (the open's reach),
// B there from the frame +-4, and writes the model only for its own pixels.
#include "serve.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kSW = 256, kSH = 16;
constexpr int kFR = kSH + 8;         // staged frame rows: tile rows +-4
constexpr int kFW = 70;              // staged dwords per frame row (cols x0-8 .. x0+271, +1 for alignment)
constexpr int kMC = 66;              // mask cells per row: cols x0-4 .. x0+259, 4 per cell
constexpr int kMR = kSH + 4;         // T rows: tile rows +-2
constexpr int kER = kSH + 2;         // erode rows: tile rows +-1

__global__ __launch_bounds__(kTileThreads) void serve_k(const ServeLaunch a) {
    __shared__ uint32_t fr[kFR * kFW + kFR];
    __shared__ uint32_t tm[kMR][kMC];
    __shared__ uint32_t em[kER][kMC];
    // Workgroup b runs on XCD b % 8 and takes tile (b % 8) ceil(N / 8) + b / 8
    // of the launch's N = frames x tiles, frame-major: each XCD walks one
    // contiguous run of tiles, so neighbouring tiles' halo rows come from
    // the L2 that read them for the tile above.
    const int tiles_x = (a.w + kSW - 1) / kSW;
    const int tiles = tiles_x * ((a.h + kSH - 1) / kSH);
    const int total = tiles * a.n;
    const int t = (int)(blockIdx.x & 7) * ((total + 7) >> 3) + (int)(blockIdx.x >> 3);
    if (t >= total) return;
    const int frame = t / tiles, tile = t - frame * tiles;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int x0 = tx * kSW, y0 = ty * kSH;
    const uint32_t job = a.job[frame];
    const ServeStream &ss = a.streams[job & 0xffffu];
    const int par = (job >> 16) & 1, primed = (job >> 17) & 1;
    const int W = a.w, H = a.h;
    DImg f;
    f.p = ss.frame[par];
    f.w = W;
    f.h = H;
    f.cn = 1;
    f.stride = a.p8;
    const gptr<const uint8_t> bg_in = gp((const uint8_t *)ss.bg[par]);
    const gptr<uint8_t> bg_out = gp((uint8_t *)ss.bg[par ^ 1]);

    const auto st = stage_rows<kFR, kFW, 5>(
        f, W, x0 - 8, 0u, 0u, fr, [&](int r) { return border_map_halo(y0 - 4 + r, H, BORDER_REFLECT101); },
        [&](int c) { return mapped_pos<1>(c, W, BORDER_REFLECT101); });

    // B, M and T on rows y0-2 .. y0+17, cols x0-4 .. x0+259.
    for (int t = threadIdx.x; t < kMR * kMC; t += kTileThreads) {
        const int r = t / kMC, c = t - r * kMC;
        const int y = y0 - 2 + r, xb = x0 - 4 + 4 * c;
        if (y < 0 || y >= H) {
            tm[r][c] = 0xffffffffu;
            continue;
        }
        // Column sums over the 5 rows for the 8 columns xb-2 .. xb+5 (staged
        // byte 4c + 2 + j), then the row taps.
        int v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = 0;
#pragma unroll
        for (int dy = 0; dy < 5; dy++) {
            const uint32_t sh = st.shift[r + dy];
            const uint32_t d0 = st.tile[r + dy][c], d1 = st.tile[r + dy][c + 1], d2 = st.tile[r + dy][c + 2],
                           d3 = st.tile[r + dy][c + 3];
            const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, sh), w1 = __builtin_amdgcn_alignbyte(d2, d1, sh),
                           w2 = __builtin_amdgcn_alignbyte(d3, d2, sh);
            const uint32_t wb[3] = {w0, w1, w2};
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int o = 2 + j;
                v[j] += __mul24(a.kx[dy], (int)((wb[o >> 2] >> ((o & 3) * 8)) & 255u));
            }
        }
        const bool own_row = y >= y0 && y < y0 + kSH;
        const bool full = xb >= 0 && xb + 4 <= W;
        uint32_t bw[2] = {0u, 0u};
        if (full) {
            const u32x2 q = *(gptr<const u32x2>)(bg_in + (long)y * a.p16 + 2L * xb);
            bw[0] = q.x;
            bw[1] = q.y;
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = xb + i;
                if (x >= 0 && x < W) {
                    const uint32_t b = *(gptr<const uint16_t>)(bg_in + (long)y * a.p16 + 2L * x);
                    bw[i >> 1] |= b << (16 * (i & 1));
                }
            }
        }
        uint32_t cell = 0, nb[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int x = xb + i;
            int h = 0;
#pragma unroll
            for (int dx = 0; dx < 5; dx++) h += __mul24(a.kx[dx], v[i + dx]);
            const int32_t fv = min((h + (1 << 15)) >> 16, 255);
            const int32_t bv = (int32_t)((bw[i >> 1] >> (16 * (i & 1))) & 0xffffu);
            int32_t nbv;
            uint32_t m;
            if (!primed) {
                nbv = fv << 8;
                m = 0;
            } else {
                const int32_t diff = fv - (bv >> 8);
                const bool fg = (diff < 0 ? -diff : diff) > a.thr;
                m = fg ? 255u : 0u;
                nbv = bv + (((fv << 8) - bv) >> (a.shift + (fg ? 2 : 0)));
            }
            nb[i >> 1] |= ((uint32_t)nbv & 0xffffu) << (16 * (i & 1));
            const uint32_t tv = (int)m > a.mthr ? 255u : 0u;
            cell |= (x >= 0 && x < W ? tv : 255u) << (8 * i);
        }
        tm[r][c] = cell;
        if (own_row && c >= 1 && c <= kSW / 4) {
            if (full) {
                *(gptr<u32x2>)(bg_out + (long)y * a.p16 + 2L * xb) = u32x2{nb[0], nb[1]};
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int x = xb + i;
                    if (x >= 0 && x < W)
                        *(gptr<uint16_t>)(bg_out + (long)y * a.p16 + 2L * x) = (uint16_t)(nb[i >> 1] >> (16 * (i & 1)));
                }
            }
        }
    }
    __syncthreads();
    // Erode: rows y0-1 .. y0+16; cells outside the image become the dilate's identity.
    for (int t = threadIdx.x; t < kER * kMC; t += kTileThreads) {
        const int r = t / kMC, c = t - r * kMC;
        const int y = y0 - 1 + r, xb = x0 - 4 + 4 * c;
        uint32_t e = 0;
        if (y >= 0 && y < H) {
            const uint32_t m = tm[r][c] & tm[r + 1][c] & tm[r + 2][c];
            const uint32_t pv = c > 0 ? tm[r][c - 1] & tm[r + 1][c - 1] & tm[r + 2][c - 1] : 0xffffffffu;
            const uint32_t nx = c + 1 < kMC ? tm[r][c + 1] & tm[r + 1][c + 1] & tm[r + 2][c + 1] : 0xffffffffu;
            e = m & ((m >> 8) | (nx << 24)) & ((m << 8) | (pv >> 24));
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = xb + i;
                if (x < 0 || x >= W) e &= ~(255u << (8 * i));
            }
        }
        em[r][c] = e;
    }
    __syncthreads();
    // Dilate and store the tile's own rows and columns.
    uint8_t *mask = ss.mask[par];
    for (int t = threadIdx.x; t < kSH * (kSW / 4); t += kTileThreads) {
        const int r = t / (kSW / 4), c = 1 + (t - r * (kSW / 4));
        const int y = y0 + r, xb = x0 - 4 + 4 * c;
        if (y >= H || xb >= W) continue;
        const uint32_t m = em[r][c] | em[r + 1][c] | em[r + 2][c];
        const uint32_t pv = em[r][c - 1] | em[r + 1][c - 1] | em[r + 2][c - 1];
        const uint32_t nx = em[r][c + 1] | em[r + 1][c + 1] | em[r + 2][c + 1];
        const uint32_t d = m | (m >> 8) | (nx << 24) | (m << 8) | (pv >> 24);
        const gptr<uint8_t> o = gp(mask + (long)y * a.p8 + xb);
        if (xb + 4 <= W) {
            *(gptr<uint32_t>)o = d;
        } else {
            for (int i = 0; xb + i < W; i++) o[i] = (uint8_t)(d >> (8 * i));
        }
    }
}

}  // namespace

int launch_serve(const ServeLaunch &a, hipStream_t st) {
    if (a.n <= 0) return OK;
    const long tiles = (long)((a.w + kSW - 1) / kSW) * ((a.h + kSH - 1) / kSH) * a.n;
    hipLaunchKernelGGL(serve_k, dim3(tile_grid((int)tiles)), dim3(kTileThreads), 0, st, a);
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

}  // namespace visar
