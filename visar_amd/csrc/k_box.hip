// k_box.hip -- 8-bit box filter for gfx950 (imgproc/filter.zig:617-806
// boxFilter / blur): exact integer window sums, then f32(sum) * f32(scale)
// rounded and saturated (FinishScaled), any border, 1-4 channels, odd
// windows up to 127 x 127.
//
// A workgroup owns a tile of TH rows and TW pixels, TW chosen so that the
// tile's padded columns ((TW + kw - 1) * cn samples) fill the 256 threads.
// Vertical pass: one thread per padded sample column loads the first window
// of kh rows and then the rows that enter and leave the window, all before
// using them, and writes the column sums of the tile's rows to LDS (running
// sums: exact in u32), one plane per channel. Horizontal pass, windows up
// to 7 (11 for adaptiveThreshold on one channel): row by row, one lane per output element
// sums its kw column sums from LDS and stores its byte (consecutive lanes,
// consecutive bytes); wider windows: one lane per run of 7 outputs, a
// running sum along the run. Window lengths up to 15 rows and up to the
// direct limit in columns run unrolled (with_odd_len).
// With `adapt` on (adaptiveThreshold), each mean is compared with its source
// byte before it is stored.
#include "kernels.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kBoxThreads = 256;
constexpr int kBoxTH = 16;  // output rows per tile
constexpr int kBoxRun = 7;  // outputs per lane for wide windows (odd: conflict-free LDS reads)
// Widest window summed directly, one lane per output; wider ones in runs.
// Measured on 1080p with both paths unrolled: gray up to 7 for a plain box
// (11x11: 7.7 us direct, 7.2 runs) and up to 11 with adaptiveThreshold's
// compare, whose source reads are strided along a run (adaptive mean 11:
// 9.2 direct, 10.0 runs); BGR up to 7 (7x7: 12.7 direct, 13.5 runs; 11x11:
// 15.3 direct, 13.1 runs).
template <int CN>
constexpr int kBoxDirect = CN == 1 ? 11 : 7;  // widest compiled direct path

template <int CN>
__global__ __launch_bounds__(kBoxThreads) void box_u8(DImg s, DImg d, int kw, int kh, int border, float scale, int tw,
                                                      TileOrder o, AdaptPost adapt) {
    extern __shared__ uint32_t colsum[];  // [kBoxTH][CN][pitch]
    __shared__ long row_off[kBoxTH + 126];  // source rows y0 - ry .. y0 + kBoxTH - 1 + ry, -1 outside
    int bx, by;
    if (!tile_of(o, bx, by)) return;
    const int x0 = bx * tw, y0 = by * kBoxTH;
    const int rx = kw / 2, ry = kh / 2;
    const int pitch = tw + kw - 1, ncols = pitch * CN, plane = ncols;
    const int W = s.w, H = s.h;
    // Rows are mapped once per tile (no per-load border logic, whose call
    // path would make the compiler wait for each load before the next).
    for (int i = threadIdx.x; i < kBoxTH + kh - 1; i += kBoxThreads) {
        const int my = border_map_halo(y0 - ry + i, H, border);
        row_off[i] = my < 0 ? -1l : (long)my * s.stride;
    }
    __syncthreads();
    for (int col = threadIdx.x; col < ncols; col += kBoxThreads) {
        const int mx = border_map_halo(x0 - rx + col / CN, W, border);
        const long off = (long)mx * CN + col % CN;
        const gptr<const uint8_t> base = gp((const uint8_t *)s.p);
        auto at = [&](int i) -> uint32_t {  // window row i (source row y0 - ry + i)
            const long r = row_off[i];
            return (mx < 0 || r < 0) ? 0u : (uint32_t)base[r + off];
        };
        uint32_t *cs = colsum + (col % CN) * pitch + col / CN;
        with_odd_len<15>(kh, [&](auto len) {
            constexpr int KH = decltype(len)::value;
            if constexpr (KH > 0) {
                // The KH + TH - 1 rows the column's sums use, each loaded
                // once, all in flight together; then the window slides.
                uint32_t r[KH + kBoxTH - 1];
#pragma unroll
                for (int i = 0; i < KH + kBoxTH - 1; i++) r[i] = at(i);
                uint32_t v = 0;
#pragma unroll
                for (int j = 0; j < KH; j++) v += r[j];
                cs[0] = v;
#pragma unroll
                for (int t = 0; t < kBoxTH - 1; t++) {
                    v += r[t + KH] - r[t];
                    cs[(t + 1) * plane] = v;
                }
            } else {
                uint32_t v = 0;
                for (int j0 = 0; j0 < kh; j0 += 8) {
                    uint32_t b[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) b[j] = j0 + j < kh ? at(j0 + j) : 0u;
#pragma unroll
                    for (int j = 0; j < 8; j++) v += b[j];
                }
                uint32_t add[kBoxTH - 1], sub[kBoxTH - 1];
#pragma unroll
                for (int t = 0; t < kBoxTH - 1; t++) {
                    add[t] = at(t + kh);
                    sub[t] = at(t);
                }
                cs[0] = v;
#pragma unroll
                for (int t = 0; t < kBoxTH - 1; t++) {
                    v += add[t] - sub[t];
                    cs[(t + 1) * plane] = v;
                }
            }
        });
    }
    __syncthreads();
    const int n_row = min(tw, W - x0) * CN, rows = min(kBoxTH, H - y0);
    const int direct_max = CN == 1 && !adapt.enabled ? 7 : kBoxDirect<CN>;
    if (kw <= direct_max) {
        // One lane per output element of a row: its kw column sums from
        // LDS, the byte stored next to its neighbours' (consecutive lanes,
        // consecutive bytes).
        with_odd_len<kBoxDirect<CN>>(kw, [&](auto len) {
            constexpr int KW = decltype(len)::value;
            for (int t = 0; t < rows; t++) {
                const uint32_t *rowt = colsum + t * plane;
                const gptr<uint8_t> out = grow<uint8_t>(d, y0 + t) + (long)x0 * CN;
                for (int e = threadIdx.x; e < n_row; e += kBoxThreads) {
                    const int ox = e / CN, c = e - ox * CN;
                    const uint32_t *r = rowt + c * pitch + ox;
                    uint32_t sum = 0;
                    if constexpr (KW > 0) {
#pragma unroll
                        for (int i = 0; i < KW; i++) sum += r[i];
                    } else {
                        for (int i = 0; i < kw; i++) sum += r[i];
                    }
                    uint8_t m = sat_v<uint8_t>((float)sum * scale);
                    if (adapt.enabled) m = adapt_out(adapt.idelta, adapt.on, adapt.off, grow<const uint8_t>(s, y0 + t)[(long)x0 * CN + e], m);
                    out[e] = m;
                }
            }
        });
        return;
    }
    // Wide windows: one lane per run of kBoxRun outputs of one channel, the
    // first window summed, then one column in and one out per output. Runs
    // of neighbouring lanes start kBoxRun (odd) words apart in a plane, so
    // their LDS reads fall in different banks.
    const int tw_row = min(tw, W - x0), runs_x = (tw_row + kBoxRun - 1) / kBoxRun, per_row = runs_x * CN;
    // Windows up to 15 columns unrolled (constant LDS offsets).
    with_odd_len<15>(kw, [&](auto len) {
        constexpr int KW = decltype(len)::value;
        const int n = KW > 0 ? KW : kw;
        for (int q = threadIdx.x; q < rows * per_row; q += kBoxThreads) {
            const int t = q / per_row, r = q - t * per_row, c = r / runs_x, run = r - c * runs_x;
            const int ox0 = run * kBoxRun;
            const uint32_t *row = colsum + t * plane + c * pitch + ox0;
            uint32_t sum = 0;
            if constexpr (KW > 0) {
#pragma unroll
                for (int i = 0; i < KW; i++) sum += row[i];
            } else {
                for (int i = 0; i < n; i++) sum += row[i];
            }
            const gptr<uint8_t> out = grow<uint8_t>(d, y0 + t) + (long)(x0 + ox0) * CN + c;
#pragma unroll
            for (int k = 0; k < kBoxRun; k++) {
                if (ox0 + k >= tw_row) break;
                if (k) sum += row[k + n - 1] - row[k - 1];
                uint8_t m = sat_v<uint8_t>((float)sum * scale);
                if (adapt.enabled)
                    m = adapt_out(adapt.idelta, adapt.on, adapt.off, grow<const uint8_t>(s, y0 + t)[(long)(x0 + ox0 + k) * CN + c], m);
                out[(long)k * CN] = m;
            }
        }
    });
}

// ------------------------------------------------ wide and 16S windows
//
// Box windows with a side above 127 (up to the reference's 255) and 16-bit
// signed samples, in two passes through a temporary of column sums: one
// thread per source sample column and 8 rows for the vertical sums, one
// thread per output sample for the horizontal ones. Integer samples sum
// exactly (u32 for u8 / u16, i32 for i16: the reference's u32 accumulator
// agrees wherever its i16 samples are non-negative); f32 samples sum in the
// engine's order (filter.zig:232-259: within the 16-lane body of a row the
// symmetric taps fold pairwise around the centre, past it the taps run in
// order), products by the unit taps being exact.
constexpr int kWideRows = 8;

// Sum of the n window samples v(0) .. v(n - 1).
template <typename A, typename V>
__device__ __attribute__((always_inline)) inline A box_taps(V v, int n, bool fold) {
    if constexpr (std::is_same<A, float>::value) {
        const int r = n / 2;
        if (fold) {
            float a = v(r);
            for (int i = 0; i < r; i++) a = a + (v(i) + v(n - 1 - i));
            return a;
        }
        float a = v(0);
        for (int j = 1; j < n; j++) a = a + v(j);
        return a;
    } else {
        A a = 0;
        for (int j = 0; j < n; j++) a += v(j);
        return a;
    }
}

template <typename T, typename A>
__global__ __launch_bounds__(kBoxThreads) void box_wide_v(DImg s, A *tmp, int kh, int border, long body) {
    const long n = (long)s.w * s.cn;
    const long c = (long)blockIdx.x * kBoxThreads + threadIdx.x;
    const int y0 = (int)blockIdx.y * kWideRows;
    if (c >= n) return;
    const int ry = kh / 2, H = s.h;
    const gptr<const T> base = gp((const T *)s.p);
    const long st = s.stride / (long)sizeof(T);
    auto at = [&](int y) -> A {  // source row y (unmapped) at column c
        const int m = border_map_halo(y, H, border);
        return m < 0 ? (A)0 : (A)base[(long)m * st + c];
    };
    const int y1 = min(y0 + kWideRows, H);
    if constexpr (std::is_same<A, float>::value) {
        for (int y = y0; y < y1; y++)
            tmp[(long)y * n + c] = box_taps<float>([&](int j) { return at(y + j - ry); }, kh, c < body);
    } else {
        A sum = 0;
        for (int j = 0; j < kh; j++) sum += at(y0 + j - ry);
        tmp[(long)y0 * n + c] = sum;
        for (int y = y0 + 1; y < y1; y++) {
            sum += at(y + ry) - at(y - 1 - ry);
            tmp[(long)y * n + c] = sum;
        }
    }
}

template <typename T, typename A>
__global__ __launch_bounds__(kBoxThreads) void box_wide_h(const A *tmp, DImg d, DImg s, int kw, int border, float scale,
                                                          long body, AdaptPost adapt) {
    const int cn = d.cn, W = d.w;
    const long n = (long)W * cn;
    const long e = (long)blockIdx.x * kBoxThreads + threadIdx.x;
    const int y = (int)blockIdx.y;
    if (e >= n) return;
    const int x = (int)(e / cn), ch = (int)(e - (long)x * cn), rx = kw / 2;
    const A *row = tmp + (long)y * n;
    auto at = [&](int px) -> A {
        const int m = border_map_halo(px, W, border);
        return m < 0 ? (A)0 : row[(long)m * cn + ch];
    };
    A a;
    if constexpr (std::is_same<A, float>::value) {
        a = box_taps<float>([&](int i) { return at(x + i - rx); }, kw, e < body);
    } else {
        a = 0;
        for (int i = 0; i < kw; i++) a += at(x + i - rx);
    }
    const float f = (float)a * scale;
    T out = sat_v<T>(f);
    if constexpr (std::is_same<T, uint8_t>::value) {
        if (adapt.enabled) out = adapt_out(adapt.idelta, adapt.on, adapt.off, grow<const uint8_t>(s, y)[e], out);
    }
    grow<T>(d, y)[e] = out;
}

inline int launched() { return hipGetLastError() == hipSuccess ? OK : INTERNAL; }

}  // namespace

int launch_box_u8(const DImg &s, const DImg &d, int kw, int kh, int border, float scale, hipStream_t st,
                  const AdaptPost &adapt) {
    const int cn = s.cn;
    if (cn < 1 || cn > 4 || kw > 127 || kh > 127) return UNSUPPORTED;
    const int tw = std::max(16, kBoxThreads / cn - (kw - 1));
    const size_t lds = (size_t)kBoxTH * (tw + kw - 1) * cn * sizeof(uint32_t);
    // XCD bands: a tile and the one below it share kh - 1 source rows, and in
    // a band they run on the same XCD, whose L2 then serves the second read
    // (1080p BGR 15x15: 16.1 -> 15.5 us).
    const TileOrder o = tile_order((s.w + tw - 1) / tw, (s.h + kBoxTH - 1) / kBoxTH, true);
    const dim3 g(tile_grid(o.ntiles)), b(kBoxThreads);
    switch (cn) {
    case 1: hipLaunchKernelGGL(box_u8<1>, g, b, lds, st, s, d, kw, kh, border, scale, tw, o, adapt); break;
    case 2: hipLaunchKernelGGL(box_u8<2>, g, b, lds, st, s, d, kw, kh, border, scale, tw, o, adapt); break;
    case 3: hipLaunchKernelGGL(box_u8<3>, g, b, lds, st, s, d, kw, kh, border, scale, tw, o, adapt); break;
    default: hipLaunchKernelGGL(box_u8<4>, g, b, lds, st, s, d, kw, kh, border, scale, tw, o, adapt); break;
    }
    return launched();
}

template <typename T, typename A>
static int run_box_wide(const DImg &s, const DImg &d, int kw, int kh, int border, float scale, void *tmp,
                        hipStream_t st, const AdaptPost &adapt) {
    const long n = (long)s.w * s.cn;
    const long body = (n / 16) * 16;
    const unsigned gx = (unsigned)((n + kBoxThreads - 1) / kBoxThreads);
    hipLaunchKernelGGL((box_wide_v<T, A>), dim3(gx, (unsigned)((s.h + kWideRows - 1) / kWideRows)), dim3(kBoxThreads), 0,
                       st, s, (A *)tmp, kh, border, body);
    hipLaunchKernelGGL((box_wide_h<T, A>), dim3(gx, (unsigned)d.h), dim3(kBoxThreads), 0, st, (const A *)tmp, d, s, kw,
                       border, scale, body, adapt);
    return launched();
}

size_t box_wide_tmp_bytes(const DImg &s) { return (size_t)s.w * s.cn * s.h * 4; }

int launch_box_wide(const DImg &s, const DImg &d, int depth, int kw, int kh, int border, float scale, void *tmp,
                    hipStream_t st, const AdaptPost &adapt) {
    if (kw > 255 || kh > 255) return UNSUPPORTED;
    switch (depth) {
    case 0: return run_box_wide<uint8_t, uint32_t>(s, d, kw, kh, border, scale, tmp, st, adapt);
    case 2: return run_box_wide<uint16_t, uint32_t>(s, d, kw, kh, border, scale, tmp, st, adapt);
    case 3: return run_box_wide<int16_t, int32_t>(s, d, kw, kh, border, scale, tmp, st, adapt);
    case 5: return run_box_wide<float, float>(s, d, kw, kh, border, scale, tmp, st, adapt);
    default: return UNSUPPORTED;
    }
}

}  // namespace visar
