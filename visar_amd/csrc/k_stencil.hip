// k_stencil.hip -- fast path of the separable engine for u8 sources and
// square kernels of 3, 5 or 7 taps (GaussianBlur u8 and Sobel u8 -> i16/u8,
// the shapes of every benchmark configuration).
//
// Exact integer sums make the pass order free (reference filter.zig:163-167).
// A 64 x 4 workgroup owns 256 consecutive samples of the flat row and 4*TH
// output rows. It stages the rows it needs, plus the column halo, in LDS with
// the border rules applied while loading (border.zig:22-39): rows are mapped
// once, whole in-row dwords are coalesced dword loads, only dwords that
// straddle the row ends are assembled byte by byte. Every thread then runs
// the same branch-free code: the horizontal taps of its 4 samples on each
// staged row, the vertical taps over those results, with full-rate 24-bit
// multiplies (taps and partial sums stay below 2^23 in magnitude).
#include "kernels.hpp"

namespace visar {

namespace {

template <typename D>
struct Store4;
template <>
struct Store4<uint8_t> {
    __device__ static void put(uint8_t *p, const int (&v)[4]) {
        uint32_t w = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
        *reinterpret_cast<uint32_t *>(p) = w;
    }
};
template <>
struct Store4<int16_t> {
    __device__ static void put(int16_t *p, const int (&v)[4]) {
        uint32_t a = (uint32_t)(uint16_t)v[0] | ((uint32_t)(uint16_t)v[1] << 16);
        uint32_t b = (uint32_t)(uint16_t)v[2] | ((uint32_t)(uint16_t)v[3] << 16);
        *reinterpret_cast<uint2 *>(p) = make_uint2(a, b);
    }
};

template <typename D>
__device__ inline int finish(int a, int fin, float scale) {
    if (fin == FIN_FIX8) {
        int r = (a + (1 << 15)) >> 16;
        return r > 255 ? 255 : r;
    }
    if (fin == FIN_SCALED) return (int)sat_v<D>((float)a * scale);
    int lo = (int)Lim<D>::lo, hi = (int)Lim<D>::hi;
    return a < lo ? lo : a > hi ? hi : a;
}

// Byte c of a source row with column border mapping (c may lie outside
// [0, n)): pixel floor(c / CN), channel c - CN * pixel.
template <int CN>
__device__ inline uint32_t mapped_byte(const uint8_t *row, int c, int W, int border) {
    const int px = (c + 64 * CN) / CN - 64;
    const int ch = c - px * CN;
    const int mx = border_map(px, W, border);
    return mx < 0 ? 0u : (uint32_t)row[mx * CN + ch];
}

template <int K, int CN, int TH, typename D>
__global__ __launch_bounds__(256) void sep_small_u8(SepParams p) {
    constexpr int R = K / 2;
    constexpr int HW = (R * CN + 3) / 4;  // halo dwords on each side
    constexpr int NW = 2 * HW + 1;        // dwords per row per thread
    constexpr int PW = 64 + 2 * HW;       // dwords per tile row
    constexpr int TR = 4 * TH + 2 * R;    // tile rows
    __shared__ uint32_t tile[TR][PW];
    const int W = p.src.w, H = p.src.h;
    const int n = W * CN;
    const int x0 = blockIdx.x * 256;      // first sample of the block
    const int y0 = blockIdx.y * 4 * TH;   // first output row of the block
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const bool aligned = ((p.src.stride | (uintptr_t)p.src.p) & 3) == 0;
    // Stage the tile: rows border-mapped (constant border rows are zero),
    // whole in-row dwords loaded as dwords, the rest byte by byte with
    // column mapping.
    for (int idx = tid; idx < TR * PW; idx += 256) {
        const int r = idx / PW, k = idx - r * PW;
        const int sy = border_map(y0 - R + r, H, p.border);
        const int c = x0 - 4 * HW + 4 * k;
        uint32_t v = 0;
        if (sy >= 0) {
            const uint8_t *row = p.src.p + (long)sy * p.src.stride;
            if (aligned && c >= 0 && c + 4 <= n) {
                v = *reinterpret_cast<const uint32_t *>(row + c);
            } else {
                for (int b = 0; b < 4; b++) v |= mapped_byte<CN>(row, c + b, W, p.border) << (8 * b);
            }
        }
        tile[r][k] = v;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x * 4;
    const int yt = threadIdx.y * TH;  // first output row of this thread, tile-relative
    if (x >= n || y0 + yt >= H) return;
    int kx[K], ky[K];
#pragma unroll
    for (int i = 0; i < K; i++) kx[i] = p.kxi[i], ky[i] = p.kyi[i];
    int hz[TH + 2 * R][4];
#pragma unroll
    for (int r = 0; r < TH + 2 * R; r++) {
        uint32_t w[NW];
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = tile[yt + r][threadIdx.x + k];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int a = 0;
#pragma unroll
            for (int i = 0; i < K; i++) {
                const int o = 4 * HW + j + (i - R) * CN;
                a += __mul24(kx[i], (int)((w[o >> 2] >> ((o & 3) * 8)) & 255u));
            }
            hz[r][j] = a;
        }
    }
#pragma unroll
    for (int t = 0; t < TH; t++) {
        const int y = y0 + yt + t;
        if (y >= H) break;
        int o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int a = 0;
#pragma unroll
            for (int k = 0; k < K; k++) a += __mul24(ky[k], hz[t + k][j]);
            o[j] = finish<D>(a, p.fin, p.scale);
        }
        D *out = reinterpret_cast<D *>(p.dst.p + (long)y * p.dst.stride) + x;
        if (x + 4 <= n && (((uintptr_t)out | (uintptr_t)p.dst.stride) & (4 * sizeof(D) - 1)) == 0) {
            Store4<D>::put(out, o);
        } else {
            for (int j = 0; j < 4 && x + j < n; j++) out[j] = (D)o[j];
        }
    }
}

template <int K, int CN, typename D>
bool launch_k(const SepParams &p, hipStream_t st, int th) {
    const long n = (long)p.src.w * CN;
    dim3 block(64, 4);
    auto grid = [&](int t) { return dim3((unsigned)((n + 255) / 256), (unsigned)((p.src.h + 4 * t - 1) / (4 * t))); };
    switch (th) {
    case 1: hipLaunchKernelGGL((sep_small_u8<K, CN, 1, D>), grid(1), block, 0, st, p); break;
    case 2: hipLaunchKernelGGL((sep_small_u8<K, CN, 2, D>), grid(2), block, 0, st, p); break;
    default: hipLaunchKernelGGL((sep_small_u8<K, CN, 4, D>), grid(4), block, 0, st, p); break;
    }
    return true;
}

template <int K, typename D>
bool launch_cn(const SepParams &p, hipStream_t st, int th) {
    switch (p.src.cn) {
    case 1: return launch_k<K, 1, D>(p, st, th);
    case 2: return launch_k<K, 2, D>(p, st, th);
    case 3: return launch_k<K, 3, D>(p, st, th);
    case 4: return launch_k<K, 4, D>(p, st, th);
    default: return false;
    }
}

template <typename D>
bool launch_d(const SepParams &p, hipStream_t st, int th) {
    switch (p.nkx) {
    case 3: return launch_cn<3, D>(p, st, th);
    case 5: return launch_cn<5, D>(p, st, th);
    case 7: return launch_cn<7, D>(p, st, th);
    default: return false;
    }
}

}  // namespace

// Rows per thread for an image of n samples per row and h rows: enough
// workgroups to cover the 256 CUs several times over.
int stencil_rows_per_thread(long n, int h) {
    long groups = (n + 255) / 256;
    if (groups * ((h + 15) / 16) >= 2048) return 4;
    if (groups * ((h + 7) / 8) >= 1024) return 2;
    return 1;
}

// Returns false when the call is not a shape this kernel covers.
bool launch_sep_small_u8(const SepParams &p, hipStream_t st, int th) {
    if (p.f32mode || p.sdepth != 0 || p.nkx != p.nky || p.src.cn > 4) return false;
    if (p.nkx != 3 && p.nkx != 5 && p.nkx != 7) return false;
    if (p.ddepth == 0 && p.fin == FIN_FIX8) return launch_d<uint8_t>(p, st, th) && hipGetLastError() == hipSuccess;
    if (p.ddepth == 3 && (p.fin == FIN_CAST || p.fin == FIN_SCALED))
        return launch_d<int16_t>(p, st, th) && hipGetLastError() == hipSuccess;
    return false;
}

}  // namespace visar
