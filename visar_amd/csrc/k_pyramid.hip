// k_pyramid.hip -- Gaussian pyramids for gfx950 (imgproc/pyramid.zig).
//
//   pyr_down_k   pyramid.zig:65-122: the 5x5 [1 4 6 4 1] kernel at every
//                other pixel, (v + 128) >> 8 (v / 256 for f32)
//   pyr_up_k     pyramid.zig:403-504: the same kernel on the image
//                upsampled with zeros and reflected about its own edges,
//                (v + 32) >> 6 (v / 64 for f32)
//
// 1-4 channels: LDS tiles of column sums (pyr_down_tile, pyr_up_tile); more
// channels: one thread per output pixel (all channels), 64 x 4 pixels per
// tile, taps read through the cache. The integer sums are exact in 32 bits (u8 at most 65280, u16 2^24, i16
// +-2^23), so the order of the taps does not matter for them; f32 sums run
// in the reference's order: each column's vertical sum (a + f) + (b + e) * 4
// + c * 6, then the same form across the five column sums.
#include "kernels.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kPyrThreads = 256;

template <typename T> struct PyrAcc { using type = int32_t; };
template <> struct PyrAcc<float> { using type = float; };

template <typename A>
__device__ __attribute__((always_inline)) inline A tap5(A a, A b, A c, A e, A f) {
    return (a + f) + (b + e) * (A)4 + c * (A)6;
}

template <typename T, int S>
__device__ __attribute__((always_inline)) inline T pyr_finish(typename PyrAcc<T>::type v) {
    if constexpr (std::is_same<T, float>::value) return v * (1.0f / (float)(1 << S));
    else return (T)((v + (1 << (S - 1))) >> S);
}

template <typename T>
__global__ __launch_bounds__(kPyrThreads) void pyr_down_k(DImg s, DImg d, int border, TileOrder o) {
    using A = typename PyrAcc<T>::type;
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int x = bx * 64 + tile_tx(), y = by * 4 + tile_ty();
    if (x >= d.w || y >= d.h) return;
    const int cn = s.cn;
    gptr<const T> r[5];
#pragma unroll
    for (int k = 0; k < 5; k++) r[k] = (gptr<const T>)grow<const uint8_t>(s, border_map(2 * y + k - 2, s.h, border));
    int cx[5];
#pragma unroll
    for (int i = 0; i < 5; i++) cx[i] = border_map(2 * x + i - 2, s.w, border) * cn;
    const gptr<T> out = (gptr<T>)grow<uint8_t>(d, y) + (long)x * cn;
    for (int c = 0; c < cn; c++) {
        A v[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const int e = cx[i] + c;
            v[i] = tap5<A>((A)r[0][e], (A)r[1][e], (A)r[2][e], (A)r[3][e], (A)r[4][e]);
        }
        out[c] = pyr_finish<T, 8>(tap5<A>(v[0], v[1], v[2], v[3], v[4]));
    }
}

template <typename T>
__global__ __launch_bounds__(kPyrThreads) void pyr_up_k(DImg s, DImg d, TileOrder o) {
    using A = typename PyrAcc<T>::type;
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int X = bx * 64 + tile_tx(), Y = by * 4 + tile_ty();
    if (X >= d.w || Y >= d.h) return;
    const int w = s.w, h = s.h, cn = s.cn;
    const int y = Y >> 1, p = X >> 1;
    const bool odd_row = Y & 1, odd_col = X & 1;
    // Row -1 is row 1, row h is row h - 1 (the upsampled image reflected
    // about its edges); columns likewise.
    const int ya = y > 0 ? y - 1 : (h > 1 ? 1 : 0), yc = y + 1 < h ? y + 1 : y;
    const int xl = p > 0 ? p - 1 : (w > 1 ? 1 : 0), xr = p + 1 < w ? p + 1 : w - 1;
    const gptr<const T> ra = (gptr<const T>)grow<const uint8_t>(s, ya), rb = (gptr<const T>)grow<const uint8_t>(s, y),
                        rc = (gptr<const T>)grow<const uint8_t>(s, yc);
    const int cols[3] = {xl * cn, p * cn, xr * cn};
    const gptr<T> out = (gptr<T>)grow<uint8_t>(d, Y) + (long)X * cn;
    for (int c = 0; c < cn; c++) {
        A q[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int e = cols[i] + c;
            const A a = (A)ra[e], b = (A)rb[e], cc = (A)rc[e];
            q[i] = odd_row ? (b + cc) * (A)4 : (a + cc) + b * (A)6;
        }
        const A v = odd_col ? (q[1] + q[2]) * (A)4 : (q[0] + q[2]) + q[1] * (A)6;
        out[c] = pyr_finish<T, 6>(v);
    }
}

// Tiled forms for 1-4 channels. A tile has kPyrTH rows of column sums: one
// thread per padded sample column loads that column's source rows once (all
// loads issued before use) and writes the vertical sums of the tile's rows
// to LDS; then each output takes its horizontal taps from LDS.
constexpr int kPyrTH = 8;

// pyrDown: tile of kPyrTH output rows and tw output pixels; padded source
// columns 2 x0 - 2 .. 2 (x0 + tw) (2 tw + 3 pixels), source rows
// 2 y0 - 2 .. 2 (y0 + kPyrTH) (2 kPyrTH + 3).
template <typename T, int CN>
__global__ __launch_bounds__(kPyrThreads) void pyr_down_tile(DImg s, DImg d, int border, int tw, TileOrder o) {
    using A = typename PyrAcc<T>::type;
    extern __shared__ uint32_t lds_words[];
    A *V = reinterpret_cast<A *>(lds_words);  // [kPyrTH][ncols]
    __shared__ long row_off[2 * kPyrTH + 3];  // source rows 2 y0 - 2 .. 2 y0 + 2 kPyrTH
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int x0 = bx * tw, y0 = by * kPyrTH;
    const int ncols = (2 * tw + 3) * CN;
    // Rows mapped once per tile: no border logic between the loads below.
    if (threadIdx.x < 2 * kPyrTH + 3) row_off[threadIdx.x] = (long)border_map_halo(2 * y0 - 2 + (int)threadIdx.x, s.h, border) * s.stride;
    __syncthreads();
    for (int col = threadIdx.x; col < ncols; col += kPyrThreads) {
        const long off = ((long)border_map_halo(2 * x0 - 2 + col / CN, s.w, border) * CN + col % CN) * (long)sizeof(T);
        const gptr<const uint8_t> base = gp((const uint8_t *)s.p) + off;
        A r[2 * kPyrTH + 3];
#pragma unroll
        for (int k = 0; k < 2 * kPyrTH + 3; k++) r[k] = (A)*(gptr<const T>)(base + row_off[k]);
#pragma unroll
        for (int t = 0; t < kPyrTH; t++)
            V[t * ncols + col] = tap5<A>(r[2 * t], r[2 * t + 1], r[2 * t + 2], r[2 * t + 3], r[2 * t + 4]);
    }
    __syncthreads();
    const int n_out = min(tw, d.w - x0) * CN, rows = min(kPyrTH, d.h - y0);
    for (int q = threadIdx.x; q < rows * n_out; q += kPyrThreads) {
        const int t = q / n_out, e = q - t * n_out, ox = e / CN, c = e - ox * CN;
        const A *row = V + t * ncols + 2 * ox * CN + c;
        ((gptr<T>)grow<uint8_t>(d, y0 + t))[(long)x0 * CN + e] =
            pyr_finish<T, 8>(tap5<A>(row[0], row[CN], row[2 * CN], row[3 * CN], row[4 * CN]));
    }
}

// pyrUp's edge rule (pyramid.zig:440-450): -1 -> 1 (0 for n = 1), n -> n - 1.
__device__ __attribute__((always_inline)) inline int up_map(int p, int n) {
    return p < 0 ? (n > 1 ? 1 : 0) : p >= n ? n - 1 : p;
}

// pyrUp: tile of kPyrUpTH source rows (2 kPyrUpTH output rows) and tw
// source pixels (tw + 2 padded columns, about 128 samples). Column pass: the
// two halves of the workgroup take the upper and lower half of the tile's
// rows, each thread one padded column, and write the even (a + c + 6 b) and
// odd (4 (b + c)) output rows' column sums to LDS. Row pass: the halves take
// the even and the odd output row of a pair; a lane reads the three column
// sums of one source pixel and stores both of its output pixels.
constexpr int kPyrUpTH = 16;
template <typename T, int CN>
__global__ __launch_bounds__(kPyrThreads) void pyr_up_tile(DImg s, DImg d, int tw, TileOrder o) {
    using A = typename PyrAcc<T>::type;
    constexpr int kHalf = kPyrThreads / 2, kRows = kPyrUpTH / 2;
    extern __shared__ uint32_t lds_words[];
    A *E = reinterpret_cast<A *>(lds_words);  // [kPyrUpTH][ncols] even output rows
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int x0 = bx * tw, y0 = by * kPyrUpTH;
    const int ncols = (tw + 2) * CN;
    A *O = E + kPyrUpTH * ncols;  // odd output rows
    const int half = threadIdx.x >= kHalf, lane = threadIdx.x & (kHalf - 1);
    const int t0 = half * kRows;
    for (int col = lane; col < ncols; col += kHalf) {
        const long off = (long)up_map(x0 - 1 + col / CN, s.w) * CN + col % CN;
        A r[kRows + 2];
#pragma unroll
        for (int k = 0; k < kRows + 2; k++)
            r[k] = (A)((gptr<const T>)grow<const uint8_t>(s, up_map(y0 - 1 + t0 + k, s.h)))[off];
#pragma unroll
        for (int t = 0; t < kRows; t++) {
            E[(t0 + t) * ncols + col] = (r[t] + r[t + 2]) + r[t + 1] * (A)6;
            O[(t0 + t) * ncols + col] = (r[t + 1] + r[t + 2]) * (A)4;
        }
    }
    __syncthreads();
    const int n_src = min(tw, s.w - x0) * CN, src_rows = min(kPyrUpTH, s.h - y0);
    const A *plane = half ? O : E;
    for (int i = 0; i < src_rows; i++) {
        const A *rowt = plane + i * ncols;
        const gptr<T> out = (gptr<T>)grow<uint8_t>(d, 2 * (y0 + i) + half) + (long)2 * x0 * CN;
        for (int e = lane; e < n_src; e += kHalf) {
            const int p = e / CN, c = e - p * CN;  // padded column p + 1 is source pixel x0 + p
            const A *row = rowt + c;
            const A a = row[p * CN], b = row[(p + 1) * CN], cc = row[(p + 2) * CN];
            out[2 * p * CN + c] = pyr_finish<T, 6>((a + cc) + b * (A)6);
            out[(2 * p + 1) * CN + c] = pyr_finish<T, 6>((b + cc) * (A)4);
        }
    }
}

inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

template <typename T, int CN>
void pyr_tiled(const DImg &s, const DImg &d, int border, bool down, hipStream_t st) {
    const dim3 b(kPyrThreads);
    if (down) {
        const int tw = std::max(8, (kPyrThreads / CN - 3) / 2);
        // XCD bands, so vertically neighbouring tiles' shared source rows are
        // read through one L2 (1080p f32 BGR: 6.7 -> 6.1 us).
        const TileOrder o = tile_order((d.w + tw - 1) / tw, (d.h + kPyrTH - 1) / kPyrTH, true);
        const size_t lds = (size_t)kPyrTH * (2 * tw + 3) * CN * 4;
        hipLaunchKernelGGL((pyr_down_tile<T, CN>), dim3(tile_grid(o.ntiles)), b, lds, st, s, d, border, tw, o);
    } else {
        const int tw = std::max(8, kPyrThreads / 2 / CN - 2);
        const TileOrder o = tile_order((s.w + tw - 1) / tw, (s.h + kPyrUpTH - 1) / kPyrUpTH, true);  // 8.35 -> 7.9 us (1080p gray)
        const size_t lds = (size_t)2 * kPyrUpTH * (tw + 2) * CN * 4;
        hipLaunchKernelGGL((pyr_up_tile<T, CN>), dim3(tile_grid(o.ntiles)), b, lds, st, s, d, tw, o);
    }
}

template <typename T>
int pyr_launch(const DImg &s, const DImg &d, int border, bool down, hipStream_t st) {
    switch (s.cn) {
    case 1: pyr_tiled<T, 1>(s, d, border, down, st); return launched();
    case 2: pyr_tiled<T, 2>(s, d, border, down, st); return launched();
    case 3: pyr_tiled<T, 3>(s, d, border, down, st); return launched();
    case 4: pyr_tiled<T, 4>(s, d, border, down, st); return launched();
    default: break;
    }
    const TileOrder o = tile_order((d.w + 63) / 64, (d.h + 3) / 4);
    const dim3 g(tile_grid(o.ntiles)), b(kPyrThreads);
    if (down) hipLaunchKernelGGL(pyr_down_k<T>, g, b, 0, st, s, d, border, o);
    else hipLaunchKernelGGL(pyr_up_k<T>, g, b, 0, st, s, d, o);
    return launched();
}

}  // namespace

int launch_pyramid(const DImg &s, const DImg &d, int depth, int border, bool down, hipStream_t st) {
    switch (depth) {
    case 0: return pyr_launch<uint8_t>(s, d, border, down, st);
    case 2: return pyr_launch<uint16_t>(s, d, border, down, st);
    case 3: return pyr_launch<int16_t>(s, d, border, down, st);
    case 5: return pyr_launch<float>(s, d, border, down, st);
    default: return UNSUPPORTED;
    }
}

}  // namespace visar
