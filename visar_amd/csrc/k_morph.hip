// k_morph.hip -- erode/dilate passes for gfx950.
//
// Reference semantics: src/imgproc/morph.zig:489-519 (the per-tap
// definition its run-length kernels reproduce): out(y, x, c) is the min
// (erode) or max (dilate) over the element's "on" cells of the source at
// (y + ky - ay, x + kx - ax, c); cells outside the image are ignored. A
// workgroup stages the tile plus its halo in LDS with out-of-image cells set
// to the identity (255/0 for u8), which ignores them exactly, then each
// output reduces over the element's horizontal runs.
#include "kernels.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;

template <typename T, bool MAX>
__device__ inline T ident() {
    if constexpr (std::is_same<T, float>::value) return MAX ? -INFINITY : INFINITY;
    else return (T)(MAX ? Lim<T>::lo : Lim<T>::hi);
}

template <typename T, bool MAX>
__device__ inline T opv(T a, T v) {
    if constexpr (MAX) return v > a ? v : a;
    else return v < a ? v : a;
}

template <typename T, bool MAX>
__global__ __launch_bounds__(kThreads) void morph_tile(MorphParams p, int twpx, int th) {
    extern __shared__ __align__(16) unsigned char smem[];
    T *tile = reinterpret_cast<T *>(smem);
    const DImg &s = p.s, &d = p.d;
    const int cn = s.cn;
    const long x0 = (long)blockIdx.x * twpx, y0 = (long)blockIdx.y * th;
    const int tw = twpx + p.kw - 1, tr = th + p.kh - 1;
    const int ldw = tw * cn;
    const T id = ident<T, MAX>();
    for (int idx = threadIdx.x; idx < tr * ldw; idx += kThreads) {
        int r = idx / ldw, j = idx - r * ldw;
        int q = j / cn, c = j - q * cn;
        long sy = y0 - p.ay + r, sx = x0 - p.ax + q;
        T v = id;
        if (sy >= 0 && sy < s.h && sx >= 0 && sx < s.w) v = row_ptr<const T>(s, sy)[sx * cn + c];
        tile[idx] = v;
    }
    __syncthreads();
    const int rows = (int)min((long)th, (long)s.h - y0);
    const int outw = (int)min((long)twpx, (long)s.w - x0) * cn;
    for (int idx = threadIdx.x; idx < rows * outw; idx += kThreads) {
        int r = idx / outw, e = idx - r * outw;
        T acc = id;
        for (int k = 0; k < p.nruns; k++) {
            MorphRun run = p.runs[k];
            const T *tp = tile + (r + run.ky) * ldw + e + run.x0 * cn;
            for (int i = 0; i < run.len; i++) acc = opv<T, MAX>(acc, tp[i * cn]);
        }
        row_ptr<T>(d, y0 + r)[x0 * cn + e] = acc;
    }
}

// Elements whose halo does not fit LDS: reduce straight from global memory.
template <typename T, bool MAX>
__global__ __launch_bounds__(kThreads) void morph_direct(MorphParams p) {
    const DImg &s = p.s, &d = p.d;
    const int cn = s.cn;
    long n = (long)s.w * cn;
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = gid / n, e = gid - y * n;
    if (y >= s.h) return;
    long x = e / cn, c = e - x * cn;
    T acc = ident<T, MAX>();
    for (int k = 0; k < p.nruns; k++) {
        MorphRun run = p.runs[k];
        long sy = y + run.ky - p.ay;
        if (sy < 0 || sy >= s.h) continue;
        const T *sr = row_ptr<const T>(s, sy);
        for (int i = 0; i < run.len; i++) {
            long sx = x + run.x0 + i - p.ax;
            if (sx >= 0 && sx < s.w) acc = opv<T, MAX>(acc, sr[sx * cn + c]);
        }
    }
    row_ptr<T>(d, y)[e] = acc;
}

// u8 fast path for elements up to 7 x 7 with 1-4 channels: a 64 x 4
// workgroup stages (4 * TH + KH - 1) rows x (256 + halo) samples of the
// source in LDS, out-of-image cells set to the identity (which ignores them
// exactly, morph.zig:2-3, 257), then each thread reduces 4 consecutive
// samples over the element's cells. The element's on-cells are a per-row
// bit mask passed as a kernel argument, so the tap loop is fully unrolled
// and its conditions are uniform across the wave.
template <int KW, int KH, int CN, int TH, bool MAX>
__global__ __launch_bounds__(256) void morph_small_u8(DImg s, DImg d, uint32_t rowmask0, uint32_t rowmask1, int ax,
                                                       int ay) {
    constexpr int HL = ((KW / 2) * CN + 3) / 4 * 4;  // bytes of halo on the left (>= ax * CN for centred anchors)
    constexpr int HW = ((KW - 1) * CN + 3) / 4 + HL / 4;  // halo dwords in total
    constexpr int PW = 64 + HW;                       // dwords per staged row
    constexpr int TR = 4 * TH + KH - 1;
    __shared__ uint32_t tile[TR][PW];
    const int W = s.w, H = s.h, n = W * CN;
    const int x0 = blockIdx.x * 256, y0 = blockIdx.y * 4 * TH;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const uint32_t id = MAX ? 0u : 0xffffffffu;
    const bool aligned = ((s.stride | (uintptr_t)s.p) & 3) == 0;
    for (int idx = tid; idx < TR * PW; idx += 256) {
        const int r = idx / PW, k = idx - r * PW;
        const int sy = y0 - ay + r;
        const int c = x0 - HL + 4 * k;
        uint32_t v = id;
        if (sy >= 0 && sy < H) {
            const uint8_t *row = s.p + (long)sy * s.stride;
            if (aligned && c >= 0 && c + 4 <= n) {
                v = *reinterpret_cast<const uint32_t *>(row + c);
            } else {
                v = 0;
                for (int b = 0; b < 4; b++) {
                    const int cc = c + b;
                    const uint32_t byte = (cc >= 0 && cc < n) ? row[cc] : (MAX ? 0u : 255u);
                    v |= byte << (8 * b);
                }
            }
        }
        tile[r][k] = v;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x * 4;
    const int yt = threadIdx.y * TH;
    if (x >= n || y0 + yt >= H) return;
    // Columns of the element relative to the thread's first sample: tap i of
    // output j reads byte HL + j + (i - ax) * CN of the thread's window.
    uint32_t acc[TH][4];
#pragma unroll
    for (int t = 0; t < TH; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[t][j] = MAX ? 0u : 255u;
#pragma unroll
    for (int r = 0; r < TH + KH - 1; r++) {
        constexpr int NWT = (HL + 4 + (KW - 1) * CN + 3) / 4 + 1;
        uint32_t w[NWT];
#pragma unroll
        for (int k = 0; k < NWT; k++) w[k] = (threadIdx.x + k < PW) ? tile[yt + r][threadIdx.x + k] : 0u;
#pragma unroll
        for (int t = 0; t < TH; t++) {
            const int ky = r - t;
            if (ky < 0 || ky >= KH) continue;
            const uint32_t bits = ky < 4 ? (rowmask0 >> (8 * ky)) & 255u : (rowmask1 >> (8 * (ky - 4))) & 255u;
#pragma unroll
            for (int i = 0; i < KW; i++) {
                if (!((bits >> i) & 1u)) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int o = HL + j + (i - ax) * CN;
                    const uint32_t v = (w[o >> 2] >> ((o & 3) * 8)) & 255u;
                    acc[t][j] = MAX ? max(acc[t][j], v) : min(acc[t][j], v);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;
        uint8_t *out = d.p + (long)y * d.stride + x;
        if (x + 4 <= n && (((uintptr_t)out | (uintptr_t)d.stride) & 3) == 0) {
            *reinterpret_cast<uint32_t *>(out) = acc[t][0] | (acc[t][1] << 8) | (acc[t][2] << 16) | (acc[t][3] << 24);
        } else {
            for (int j = 0; j < 4 && x + j < n; j++) out[j] = (uint8_t)acc[t][j];
        }
    }
}

template <int KW, int KH, int CN, bool MAX>
void launch_morph_small(const MorphParams &p, uint32_t m0, uint32_t m1, hipStream_t st) {
    const long n = (long)p.s.w * CN;
    const int th = p.s.h >= 1024 ? 2 : 1;
    dim3 block(64, 4);
    if (th == 2) {
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)((p.s.h + 7) / 8));
        hipLaunchKernelGGL((morph_small_u8<KW, KH, CN, 2, MAX>), grid, block, 0, st, p.s, p.d, m0, m1, p.ax, p.ay);
    } else {
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)((p.s.h + 3) / 4));
        hipLaunchKernelGGL((morph_small_u8<KW, KH, CN, 1, MAX>), grid, block, 0, st, p.s, p.d, m0, m1, p.ax, p.ay);
    }
}

template <int K, bool MAX>
bool morph_small_cn(const MorphParams &p, uint32_t m0, uint32_t m1, hipStream_t st) {
    switch (p.s.cn) {
    case 1: launch_morph_small<K, K, 1, MAX>(p, m0, m1, st); return true;
    case 2: launch_morph_small<K, K, 2, MAX>(p, m0, m1, st); return true;
    case 3: launch_morph_small<K, K, 3, MAX>(p, m0, m1, st); return true;
    case 4: launch_morph_small<K, K, 4, MAX>(p, m0, m1, st); return true;
    default: return false;
    }
}

// Elements up to 7 x 7 run in a 3 x 3, 5 x 5 or 7 x 7 frame (a smaller
// element is placed in the frame's centre rows/columns, anchors centred).
template <bool MAX>
bool try_morph_small(const MorphParams &p, hipStream_t st) {
    if (p.depth != 0 || p.s.cn > 4 || p.kw > 7 || p.kh > 7) return false;
    if (p.ax != p.kw / 2 || p.ay != p.kh / 2) return false;
    const int K = (p.kw > p.kh ? p.kw : p.kh) | 1;
    const int dx = K / 2 - p.ax, dy = K / 2 - p.ay;
    uint32_t rows[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < p.nruns; k++) {
        const MorphRun &r = p.runs[k];
        for (int i = 0; i < r.len; i++) rows[r.ky + dy] |= 1u << (r.x0 + i + dx);
    }
    const uint32_t m0 = rows[0] | rows[1] << 8 | rows[2] << 16 | rows[3] << 24;
    const uint32_t m1 = rows[4] | rows[5] << 8 | rows[6] << 16 | rows[7] << 24;
    MorphParams q = p;
    q.ax = K / 2;
    q.ay = K / 2;
    switch (K) {
    case 1:
    case 3: return morph_small_cn<3, MAX>(q, m0, m1, st);
    case 5: return morph_small_cn<5, MAX>(q, m0, m1, st);
    case 7: return morph_small_cn<7, MAX>(q, m0, m1, st);
    default: return false;
    }
}

template <typename T, bool MAX>
int run_morph(const MorphParams &p, hipStream_t st) {
    if constexpr (std::is_same<T, uint8_t>::value) {
        if (try_morph_small<MAX>(p, st)) return hipGetLastError() == hipSuccess ? OK : INTERNAL;
    }
    const int cn = p.s.cn;
    const size_t cap = 60 * 1024;
    int twpx = cn == 1 ? 256 : cn == 2 ? 128 : cn <= 4 ? 64 : 16;
    int th = 16;
    auto need = [&](int tw, int t) { return (size_t)(t + p.kh - 1) * (tw + p.kw - 1) * cn * sizeof(T); };
    while (need(twpx, th) > cap && th > 4) th /= 2;
    while (need(twpx, th) > cap && twpx > 16) twpx /= 2;
    if (need(twpx, th) <= cap) {
        dim3 grid((unsigned)((p.s.w + twpx - 1) / twpx), (unsigned)((p.s.h + th - 1) / th));
        hipLaunchKernelGGL((morph_tile<T, MAX>), grid, dim3(kThreads), need(twpx, th), st, p, twpx, th);
    } else {
        long n = (long)p.s.w * cn * p.s.h;
        hipLaunchKernelGGL((morph_direct<T, MAX>), dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p);
    }
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

template <typename T>
int run_morph_op(const MorphParams &p, hipStream_t st) {
    return p.op_max ? run_morph<T, true>(p, st) : run_morph<T, false>(p, st);
}

}  // namespace

int launch_morph_pass(const MorphParams &p, hipStream_t st) {
    switch (p.depth) {
    case 0: return run_morph_op<uint8_t>(p, st);
    case 2: return run_morph_op<uint16_t>(p, st);
    case 3: return run_morph_op<int16_t>(p, st);
    case 5: return run_morph_op<float>(p, st);
    default: return UNSUPPORTED;
    }
}

}  // namespace visar
