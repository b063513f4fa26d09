// k_pyramid.hip -- Gaussian pyramids for gfx950 (imgproc/pyramid.zig).
//
//   pyr_down_k   pyramid.zig:65-122: the 5x5 [1 4 6 4 1] kernel at every
//                other pixel, (v + 128) >> 8 (v / 256 for f32)
//   pyr_up_k     pyramid.zig:403-504: the same kernel on the image
//                upsampled with zeros and reflected about its own edges,
//                (v + 32) >> 6 (v / 64 for f32)
//
// One thread per output pixel (all channels), 64 x 4 pixels per tile. The
// integer sums are exact in 32 bits (u8 at most 65280, u16 2^24, i16
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

inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

template <typename T>
int pyr_launch(const DImg &s, const DImg &d, int border, bool down, hipStream_t st) {
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
