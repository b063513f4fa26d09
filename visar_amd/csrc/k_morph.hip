// k_morph.hip -- erode/dilate passes for gfx950.
//
// Reference semantics: src/imgproc/morph.zig:489-519 (the per-tap
// definition its run-length kernels reproduce): out(y, x, c) is the min
// (erode) or max (dilate) over the element's "on" cells of the source at
// (y + ky - ay, x + kx - ax, c); cells outside the image are ignored. A
// workgroup stages the tile plus its halo in LDS with out-of-image cells set
// to the identity (255/0 for u8), which ignores them exactly, then each
// output reduces over the element's horizontal runs.
#include "tiles.hpp"

namespace visar {

int morph_small_args(const MorphParams &p, MorphSmallArgs &a);

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

template <int KW, int CN, int TH, bool MAX>
__global__ __launch_bounds__(256) void morph_small_u8(MorphSmallArgs a) {
    __shared__ uint32_t lds[MorphShape<KW, CN, TH>::LDS_WORDS];
    morph_small_tile<KW, CN, TH, MAX>(a, blockIdx.x, blockIdx.y, lds);
}

template <int K, int CN, bool MAX>
void launch_morph_small(const MorphSmallArgs &a, hipStream_t st) {
    const long n = (long)a.src.w * CN;
    if (a.src.h >= 1024) {
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)((a.src.h + 7) / 8));
        hipLaunchKernelGGL((morph_small_u8<K, CN, 2, MAX>), grid, dim3(256), 0, st, a);
    } else {
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)((a.src.h + 3) / 4));
        hipLaunchKernelGGL((morph_small_u8<K, CN, 1, MAX>), grid, dim3(256), 0, st, a);
    }
}

template <int K, bool MAX>
bool morph_small_cn(const MorphSmallArgs &a, hipStream_t st) {
    switch (a.src.cn) {
    case 1: launch_morph_small<K, 1, MAX>(a, st); return true;
    case 2: launch_morph_small<K, 2, MAX>(a, st); return true;
    case 3: launch_morph_small<K, 3, MAX>(a, st); return true;
    case 4: launch_morph_small<K, 4, MAX>(a, st); return true;
    default: return false;
    }
}

template <bool MAX>
bool try_morph_small(const MorphParams &p, hipStream_t st) {
    MorphSmallArgs a;
    int k = morph_small_args(p, a);
    if (!k) return false;
    switch (k) {
    case 3: return morph_small_cn<3, MAX>(a, st);
    case 5: return morph_small_cn<5, MAX>(a, st);
    case 7: return morph_small_cn<7, MAX>(a, st);
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

// Elements up to 7 x 7 on u8 images with 1-4 channels run in a 3 x 3, 5 x 5
// or 7 x 7 frame (the element placed so that its centred anchor is the
// frame's centre). Returns the frame size, or 0 when the fast path does not apply.
int morph_small_args(const MorphParams &p, MorphSmallArgs &a) {
    if (p.depth != 0 || p.s.cn > 4 || p.kw > 7 || p.kh > 7) return 0;
    if (p.ax != p.kw / 2 || p.ay != p.kh / 2) return 0;
    int K = (p.kw > p.kh ? p.kw : p.kh) | 1;
    if (K == 1) K = 3;  // a single cell runs in the 3 x 3 frame
    const int dx = K / 2 - p.ax, dy = K / 2 - p.ay;
    uint32_t rows[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < p.nruns; k++) {
        const MorphRun &r = p.runs[k];
        for (int i = 0; i < r.len; i++) rows[r.ky + dy] |= 1u << (r.x0 + i + dx);
    }
    a.src = p.s;
    a.dst = p.d;
    a.mask0 = rows[0] | rows[1] << 8 | rows[2] << 16 | rows[3] << 24;
    a.mask1 = rows[4] | rows[5] << 8 | rows[6] << 16 | rows[7] << 24;
    return K;
}

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
