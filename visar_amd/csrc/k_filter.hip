// k_filter.hip -- the separable filter engine (GaussianBlur, Sobel,
// sepFilter2D) for gfx950.
//
// Reference semantics: src/imgproc/filter.zig:120-205 (separable/chunk) and
// :232-328 (taps). A workgroup owns a tile of TH output rows x TWPX pixels.
// Pass 1 writes the vertically filtered values of the tile's rows, including
// the horizontal halo (border-mapped columns: the reference replicates
// vertically filtered columns, filter.zig:193-199), into LDS. Pass 2 runs the
// horizontal taps out of LDS and finishes. Integer modes are exact, so pass
// order is free; the f32 mode keeps the reference's per-lane order.
#include "kernels.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kThreads = 256;
constexpr int kStageRows = 8;  // source rows whose loads are in flight together

// Calls that can carry adaptiveThreshold's compare (SepParams::adapt).
template <typename S, typename D>
constexpr bool kAdaptable = std::is_same<S, uint8_t>::value && std::is_same<D, uint8_t>::value;

template <typename S>
__device__ inline float to_f(S v) {
    return (float)v;
}

// Integer finish (filter.zig:61-116) of an exact accumulator.
template <typename D, typename A>
__device__ inline D finish_int(A a, int fin, float scale) {
    if (fin == FIN_FIX8) {
        uint32_t r = ((uint32_t)a + (1u << 15)) >> 16;
        return (D)(r > 255u ? 255u : r);
    }
    if (fin == FIN_FIX16) {
        uint64_t r = ((uint64_t)a + (1ull << 31)) >> 32;
        return (D)(r > 65535ull ? 65535ull : r);
    }
    if (fin == FIN_SCALED) return sat_v<D>((float)a * scale);
    if constexpr (std::is_same<D, float>::value) {
        return (float)a;
    } else {
        double lo = Lim<D>::lo, hi = Lim<D>::hi;
        long long v = (long long)a;
        return (D)(v < (long long)lo ? (long long)lo : v > (long long)hi ? (long long)hi : v);
    }
}

template <typename D>
__device__ inline D finish_f(float a, int fin, float scale) {
    if (fin == FIN_SCALED) a = a * scale;
    return sat_v<D>(a);
}

// tapsV (filter.zig:232-252) over values v(j), folded by symmetry.
template <typename F>
__device__ inline float taps_fold(const float *k, int n, int sym, F v) {
    int r = n / 2;
    if (sym == SYM_NONE) {
        float a = 0.0f;
        for (int j = 0; j < n; j++) a = a + v(j) * k[j];
        return a;
    }
    float a = sym == SYM_EVEN ? v(r) * k[r] : 0.0f;
    for (int i = 0; i < r; i++) {
        float t = v(i), b = v(n - 1 - i);
        float pair = sym == SYM_EVEN ? t + b : b - t;
        a = a + pair * (sym == SYM_EVEN ? k[i] : k[n - 1 - i]);
    }
    return a;
}

// tapsS (filter.zig:255-259) and the compile-time-kernel tail (:297-301).
template <typename F>
__device__ inline float taps_seq(const float *k, int n, int from_zero, F v) {
    float a;
    int j0;
    if (from_zero) {
        a = 0.0f;
        j0 = 0;
    } else {
        a = k[0] * v(0);
        j0 = 1;
    }
    for (int j = j0; j < n; j++) a = a + k[j] * v(j);
    return a;
}

// taps_fold / taps_seq / integer taps for a window length N fixed at compile
// time (with_odd_len): fully unrolled, so the N reads issue together and the
// weights sit in scalar registers. The operations and their order are those
// of the runtime versions.
template <int N, typename F>
__device__ inline float taps_fold_n(const float *k, int sym, F v) {
    float t[N];
#pragma unroll
    for (int j = 0; j < N; j++) t[j] = v(j);
    if (sym == SYM_NONE) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < N; j++) a = a + t[j] * k[j];
        return a;
    }
    constexpr int r = N / 2;
    const bool even = sym == SYM_EVEN;
    float a = even ? t[r] * k[r] : 0.0f;
#pragma unroll
    for (int i = 0; i < r; i++) a = a + (even ? t[i] + t[N - 1 - i] : t[N - 1 - i] - t[i]) * (even ? k[i] : k[N - 1 - i]);
    return a;
}

template <int N, typename F>
__device__ inline float taps_seq_n(const float *k, int from_zero, F v) {
    float t[N];
#pragma unroll
    for (int j = 0; j < N; j++) t[j] = v(j);
    float a = k[0] * t[0];
    if (from_zero) a = 0.0f + a;  // 0 + k0 v0, as the loop from zero computes it
#pragma unroll
    for (int j = 1; j < N; j++) a = a + k[j] * t[j];
    return a;
}

template <int N, typename A, typename F>
__device__ inline A taps_int_n(const int32_t *k, F v) {
    A t[N];
#pragma unroll
    for (int j = 0; j < N; j++) t[j] = (A)v(j);
    A a = 0;
#pragma unroll
    for (int j = 0; j < N; j++) a += (A)k[j] * t[j];
    return a;
}

// Vertical value of column `col` (flat sample index) for output row y, with
// the tile's mapped source rows in `rmap` (-1 = constant border row). The
// loads are issued in chunks (4 symmetric pairs or 8 taps) before they are
// accumulated, so a column's taps are in flight together; the accumulation
// order is that of taps_fold / taps_seq above.
template <typename S, typename VT, bool F32>
__device__ inline VT vert(const SepParams &p, const int *rmap, long col) {
    auto ld = [&](int j) -> VT {
        const int sy = rmap[j];
        if constexpr (F32) return sy < 0 ? 0.0f : to_f(row_ptr<const S>(p.src, sy)[col]);
        else return sy < 0 ? (VT)0 : (VT)row_ptr<const S>(p.src, sy)[col];
    };
    const int n = p.nky;
    if constexpr (F32) {
        const float *k = p.kyf;
        if (col < p.body && p.vsym != SYM_NONE) {
            const int r = n / 2;
            const bool even = p.vsym == SYM_EVEN;
            float a = even ? ld(r) * k[r] : 0.0f;
            for (int i0 = 0; i0 < r; i0 += 4) {
                float t[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    t[u] = i0 + u < r ? ld(i0 + u) : 0.0f;
                    b[u] = i0 + u < r ? ld(n - 1 - i0 - u) : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (i0 + u < r) a = a + (even ? t[u] + b[u] : b[u] - t[u]) * (even ? k[i0 + u] : k[n - 1 - i0 - u]);
            }
            return a;
        }
        // Taps in order: the unfolded body (0 + v k), the compile-time
        // kernel's tail (0 + k v) or the scalar tail (k0 v0 + k v).
        int j0 = 0;
        float a = 0.0f;
        if (col >= p.body && !p.const_kernel) {
            a = k[0] * ld(0);
            j0 = 1;
        }
        for (int c0 = j0; c0 < n; c0 += 8) {
            float v8[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v8[u] = c0 + u < n ? ld(c0 + u) : 0.0f;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (c0 + u < n) a = a + k[c0 + u] * v8[u];
        }
        return a;
    } else {
        VT a = 0;
        for (int c0 = 0; c0 < n; c0 += 8) {
            VT v8[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v8[u] = c0 + u < n ? ld(c0 + u) : (VT)0;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (c0 + u < n) a += (VT)p.kyi[c0 + u] * v8[u];
        }
        return a;
    }
}

// The same vertical value from the tile's staged source rows: tap j of the
// column is stg[j * ldw] (0 already stored for constant-border samples).
template <int N, typename S, typename VT, bool F32>
__device__ inline VT vert_staged(const SepParams &p, const S *stg, int ldw, long col) {
    const int n = p.nky;
    if constexpr (F32) {
        auto v = [&](int j) -> float { return to_f(stg[j * ldw]); };
        if constexpr (N > 0) return col < p.body ? taps_fold_n<N>(p.kyf, p.vsym, v) : taps_seq_n<N>(p.kyf, p.const_kernel, v);
        else return col < p.body ? taps_fold(p.kyf, n, p.vsym, v) : taps_seq(p.kyf, n, p.const_kernel, v);
    } else if constexpr (N > 0) {
        return taps_int_n<N, VT>(p.kyi, [&](int j) { return stg[j * ldw]; });
    } else {
        VT a = 0;
        for (int j = 0; j < n; j++) a += (VT)p.kyi[j] * (VT)stg[j * ldw];
        return a;
    }
}

// ONE: one channel whose padded tile row fits kThreads samples. The LDS row
// pitch is then the constant kThreads and the channel stride 1, so every
// tap's LDS offset is an immediate of its read instead of an address the
// vector unit computes; f32-mode calls also stage the samples converted to
// f32, once, instead of converting each at every tap that reads it.
template <typename S, typename D, typename VT, typename AT, bool F32, bool ONE>
__global__ __launch_bounds__(kThreads) void sep_tile(SepParams p, int twpx, int th) {
    using ST = typename std::conditional<F32 && ONE, float, S>::type;  // staged sample
    extern __shared__ __align__(16) unsigned char smem[];
    const int cn = ONE ? 1 : p.src.cn, W = p.src.w, H = p.src.h;
    const int rx = p.nkx / 2, ry = p.nky / 2;
    const int ncols = (twpx + 2 * rx) * cn;  // padded tile row, samples
    const int ldw = ONE ? kThreads : ncols;  // LDS row pitch
    const long x0 = (long)blockIdx.x * twpx;
    const long y0 = (long)blockIdx.y * th;
    const int rows = (int)min((long)th, (long)H - y0);
    // adaptiveThreshold's compare, read here once: read inside the lambdas
    // below, the fields make the compiler copy the parameter block to
    // scratch memory.
    const int a_en = p.adapt.enabled, a_d = p.adapt.idelta, a_on = p.adapt.on, a_off = p.adapt.off;
    int *rmap = reinterpret_cast<int *>(smem);
    VT *V = reinterpret_cast<VT *>(smem + (((th + 2 * ry) * sizeof(int) + 15) & ~(size_t)15));
    ST *stg = reinterpret_cast<ST *>(V + (size_t)th * ldw);  // [rows + 2 ry][ldw] source samples
    for (int i = threadIdx.x; i < rows + 2 * ry; i += kThreads) rmap[i] = (int)border_map(y0 + i - ry, H, p.border);
    __syncthreads();
    // Pass 0: the tile's source rows (halo columns border-mapped, constant
    // border 0) into LDS; one column per thread, its mapping computed once,
    // eight rows' loads in flight before they are stored.
    const int srows = rows + 2 * ry;
    for (int j = threadIdx.x; j < ncols; j += kThreads) {
        const int q = j / cn, ch = j - q * cn;
        const int mx = border_map_halo((int)(x0 - rx + q), W, p.border);
        const long off = (long)mx * cn + ch;
        for (int r0 = 0; r0 < srows; r0 += kStageRows) {
            S v[kStageRows];
#pragma unroll
            for (int u = 0; u < kStageRows; u++) {
                const int sy = r0 + u < srows ? rmap[r0 + u] : -1;
                v[u] = (mx >= 0 && sy >= 0) ? row_ptr<const S>(p.src, sy)[off] : (S)0;
            }
#pragma unroll
            for (int u = 0; u < kStageRows; u++)
                if (r0 + u < srows) stg[(r0 + u) * ldw + j] = (ST)v[u];
        }
    }
    __syncthreads();
    // Pass 1: vertical taps from the staged rows into LDS, one column per
    // thread.
    with_odd_len<15>(p.nky, [&](auto len) {
        constexpr int N = decltype(len)::value;
        for (int j = threadIdx.x; j < ncols; j += kThreads) {
            const int q = j / cn, ch = j - q * cn;
            const long mx = border_map_halo((int)(x0 - rx + q), W, p.border);
            for (int r = 0; r < rows; r++)
                V[r * ldw + j] = mx >= 0 ? vert_staged<N, ST, VT, F32>(p, stg + r * ldw + j, ldw, mx * cn + ch) : (VT)0;
        }
    });
    __syncthreads();
    // Pass 2: horizontal taps and finish.
    const int outw = (int)min((long)twpx, (long)W - x0) * cn;
    with_odd_len<15>(p.nkx, [&](auto len) {
        constexpr int N = decltype(len)::value;
        for (int r = 0; r < rows; r++)
        for (int e = threadIdx.x; e < outw; e += kThreads) {
            const VT *vr = V + r * ldw + e;
            long ge = x0 * cn + e;
            D o;
            if constexpr (F32) {
                auto v = [&](int i) -> float { return vr[i * cn]; };
                float a;
                if constexpr (N > 0) a = ge < p.body ? taps_fold_n<N>(p.kxf, p.hsym, v) : taps_seq_n<N>(p.kxf, p.const_kernel, v);
                else a = ge < p.body ? taps_fold(p.kxf, p.nkx, p.hsym, v) : taps_seq(p.kxf, p.nkx, p.const_kernel, v);
                o = finish_f<D>(a, p.fin, p.scale);
            } else {
                AT a;
                if constexpr (N > 0) {
                    a = taps_int_n<N, AT>(p.kxi, [&](int i) { return vr[i * cn]; });
                } else {
                    a = 0;
                    for (int i = 0; i < p.nkx; i++) a += (AT)p.kxi[i] * (AT)vr[i * cn];
                }
                o = finish_int<D, AT>(a, p.fin, p.scale);
            }
            // adaptiveThreshold: the source sample is staged row r + ry,
            // column e + rx cn.
            if constexpr (kAdaptable<S, D>)
                if (a_en) o = adapt_out(a_d, a_on, a_off, (int)stg[(r + ry) * ldw + e + rx * cn], o);
            row_ptr<D>(p.dst, y0 + r)[ge] = o;
        }
    });
}

// Fallback when a tile's halo does not fit LDS (very wide kernels on many
// channels): every output recomputes the vertical values it needs.
template <typename S, typename D, typename VT, typename AT, bool F32>
__global__ __launch_bounds__(kThreads) void sep_direct(SepParams p) {
    const int cn = p.src.cn, W = p.src.w, H = p.src.h;
    const int rx = p.nkx / 2, ry = p.nky / 2;
    const long n = (long)W * cn;
    long gid = (long)blockIdx.x * kThreads + threadIdx.x;
    long y = blockIdx.y;
    if (gid >= n) return;
    int rmap[kMaxTaps];
    for (int j = 0; j < p.nky; j++) rmap[j] = (int)border_map(y + j - ry, H, p.border);
    long x = gid / cn, ch = gid - x * cn;
    auto vcol = [&](int i) -> VT {
        long mx = border_map_halo((int)(x + i - rx), W, p.border);
        return mx < 0 ? (VT)0 : vert<S, VT, F32>(p, rmap, mx * cn + ch);
    };
    D o;
    if constexpr (F32) {
        auto v = [&](int i) -> float { return (float)vcol(i); };
        float a = gid < p.body ? taps_fold(p.kxf, p.nkx, p.hsym, v) : taps_seq(p.kxf, p.nkx, p.const_kernel, v);
        o = finish_f<D>(a, p.fin, p.scale);
    } else {
        AT a = 0;
        for (int i = 0; i < p.nkx; i++) a += (AT)p.kxi[i] * (AT)vcol(i);
        o = finish_int<D, AT>(a, p.fin, p.scale);
    }
    if constexpr (kAdaptable<S, D>)
        if (p.adapt.enabled) o = adapt_out(p.adapt.idelta, p.adapt.on, p.adapt.off, row_ptr<const S>(p.src, y)[gid], o);
    row_ptr<D>(p.dst, y)[gid] = o;
}

template <typename S, typename D, typename VT, typename AT, bool F32>
int run_sep(const SepParams &p, hipStream_t st) {
    const int cn = p.src.cn;
    const int rx = p.nkx / 2, ry = p.nky / 2;
    const size_t lds_cap = 60 * 1024;
    // The padded tile row ((twpx + 2 rx) * cn samples) fits the 256 threads,
    // so passes 0 and 1 take one column per thread and no wave does a second
    // one while the others wait at the barrier.
    int twpx = std::max(16, kThreads / std::min(cn, 16) - 2 * rx);
    // Rows per tile, measured on 1080p: 6 for one channel (an 11x11 f32
    // blur: 4 rows 17.4 us, 6 rows 15.2, 8 rows 16.2; adaptive Gaussian 11:
    // 18.1, 16.0, 17.1), 10 for three or more (an 11x11 u8 BGR blur: 6 rows
    // 31.7 us, 8 rows 30.0, 10 rows 28.1, 12 rows 29.1).
    int th = tuning().sep_rows ? tuning().sep_rows : cn >= 3 ? 10 : 6;
    // LDS bytes of a tile (the row table, the vertical values, the staged
    // rows), as sep_tile lays them out: `one` when its ONE form applies.
    auto need_as = [&](int tw, int t, bool one) {
        const size_t pitch = one ? (size_t)kThreads : (size_t)(tw + 2 * rx) * cn;
        const size_t st_size = one && F32 ? sizeof(float) : sizeof(S);
        return (((size_t)(t + 2 * ry) * sizeof(int) + 15) & ~(size_t)15) + (size_t)t * pitch * sizeof(VT) +
               (size_t)(t + 2 * ry) * pitch * st_size;
    };
    auto need = [&](int tw, int t) { return need_as(tw, t, false); };
    while (need(twpx, th) > lds_cap && th > 1) th /= 2;
    while (need(twpx, th) > lds_cap && twpx > 1) twpx /= 2;
    if (need(twpx, th) <= lds_cap) {
        dim3 grid((unsigned)((p.src.w + twpx - 1) / twpx), (unsigned)((p.src.h + th - 1) / th));
        const bool one = cn == 1 && twpx + 2 * rx <= kThreads && need_as(twpx, th, true) <= lds_cap;
        if (one)
            hipLaunchKernelGGL((sep_tile<S, D, VT, AT, F32, true>), grid, dim3(kThreads), need_as(twpx, th, true), st, p, twpx, th);
        else
            hipLaunchKernelGGL((sep_tile<S, D, VT, AT, F32, false>), grid, dim3(kThreads), need(twpx, th), st, p, twpx, th);
    } else {
        long n = (long)p.src.w * cn;
        dim3 grid((unsigned)((n + kThreads - 1) / kThreads), (unsigned)p.src.h);
        hipLaunchKernelGGL((sep_direct<S, D, VT, AT, F32>), grid, dim3(kThreads), 0, st, p);
    }
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

template <typename S>
int run_sep_f32_dst(const SepParams &p, hipStream_t st) {
    switch (p.ddepth) {
    case 0: return run_sep<S, uint8_t, float, float, true>(p, st);
    case 2: return run_sep<S, uint16_t, float, float, true>(p, st);
    case 3: return run_sep<S, int16_t, float, float, true>(p, st);
    case 5: return run_sep<S, float, float, float, true>(p, st);
    default: return UNSUPPORTED;
    }
}

}  // namespace

int launch_sep(const SepParams &p, hipStream_t st) {
    const int th = tuning().stencil_rows ? tuning().stencil_rows : stencil_rows_per_thread((long)p.src.w * p.src.cn, p.src.h);
    if (!p.adapt.enabled && launch_sep_small_u8(p, st, th)) return OK;
    if (p.f32mode) {
        switch (p.sdepth) {
        case 0: return run_sep_f32_dst<uint8_t>(p, st);
        case 2: return run_sep_f32_dst<uint16_t>(p, st);
        case 3: return run_sep_f32_dst<int16_t>(p, st);
        case 5: return run_sep_f32_dst<float>(p, st);
        default: return UNSUPPORTED;
        }
    }
    // Integer modes: u8 Gaussian (FIX8), u16 Gaussian (FIX16), u8 -> i16 Sobel.
    if (p.sdepth == 0 && p.ddepth == 0) return run_sep<uint8_t, uint8_t, int32_t, int32_t, false>(p, st);
    if (p.sdepth == 2 && p.ddepth == 2) return run_sep<uint16_t, uint16_t, uint32_t, uint64_t, false>(p, st);
    if (p.sdepth == 0 && p.ddepth == 3) return run_sep<uint8_t, int16_t, int32_t, int32_t, false>(p, st);
    return UNSUPPORTED;
}

}  // namespace visar
