// k_stencil.hip -- fast path of the separable engine: u8 sources with square
// 3/5/7 kernels (GaussianBlur u8, Sobel u8 -> i16), the shapes of every
// benchmark configuration. The per-tile work is stencil_tile (tiles.hpp),
// shared with the persistent batch kernel.
#include <initializer_list>

#include "stencil_stream.hpp"

namespace visar {

namespace {

template <int K, int CN, int TH, typename D>
__global__ __launch_bounds__(256) void sep_small_u8(StencilArgs p, TileOrder o) {
    __shared__ uint32_t lds[StencilShape<K, CN, TH>::LDS_WORDS];
    int bx, by;
    if (tile_of(o, bx, by)) stencil_tile<K, CN, TH, D>(p, bx, by, lds);
}

template <int MODE, int CN, int TH, typename D>
__global__ __launch_bounds__(256) void sep_swar(StencilArgs p, TileOrder o) {
    __shared__ uint32_t lds[StencilShape<SwarTaps<MODE>::K, CN, TH>::LDS_WORDS];
    int bx, by;
    if (tile_of(o, bx, by)) stencil_swar_tile<MODE, CN, TH, D>(p, bx, by, lds);
}

// Tiles of 256 bytes x 4 th rows.
inline TileOrder stencil_order(const StencilArgs &p, int cn, int th) {
    return tile_order((int)(((long)p.src.w * cn + 255) / 256), (p.src.h + 4 * th - 1) / (4 * th));
}

template <int MODE, int CN, typename D>
void launch_swar_th(const StencilArgs &p, hipStream_t st, int th) {
    const TileOrder o = stencil_order(p, CN, th == 1 || th == 2 ? th : 4);
    const dim3 g(tile_grid(o.ntiles));
    switch (th) {
    case 1: hipLaunchKernelGGL((sep_swar<MODE, CN, 1, D>), g, dim3(256), 0, st, p, o); break;
    case 2: hipLaunchKernelGGL((sep_swar<MODE, CN, 2, D>), g, dim3(256), 0, st, p, o); break;
    default: hipLaunchKernelGGL((sep_swar<MODE, CN, 4, D>), g, dim3(256), 0, st, p, o); break;
    }
}

template <int MODE, typename D>
bool launch_swar(const StencilArgs &p, hipStream_t st, int th) {
    if (p.src.cn == 1) launch_swar_th<MODE, 1, D>(p, st, th);
    else if (p.src.cn == 3) launch_swar_th<MODE, 3, D>(p, st, th);
    else return false;
    return true;
}

template <int K, int CN, typename D>
bool launch_k(const StencilArgs &p, hipStream_t st, int th) {
    const TileOrder o = stencil_order(p, CN, th == 1 || th == 2 ? th : 4);
    const dim3 g(tile_grid(o.ntiles));
    switch (th) {
    case 1: hipLaunchKernelGGL((sep_small_u8<K, CN, 1, D>), g, dim3(256), 0, st, p, o); break;
    case 2: hipLaunchKernelGGL((sep_small_u8<K, CN, 2, D>), g, dim3(256), 0, st, p, o); break;
    default: hipLaunchKernelGGL((sep_small_u8<K, CN, 4, D>), g, dim3(256), 0, st, p, o); break;
    }
    return true;
}

template <int K, typename D>
bool launch_cn(const StencilArgs &p, hipStream_t st, int th) {
    switch (p.src.cn) {
    case 1: return launch_k<K, 1, D>(p, st, th);
    case 2: return launch_k<K, 2, D>(p, st, th);
    case 3: return launch_k<K, 3, D>(p, st, th);
    case 4: return launch_k<K, 4, D>(p, st, th);
    default: return false;
    }
}

template <typename D>
bool launch_d(const StencilArgs &p, int k, hipStream_t st, int th) {
    switch (k) {
    case 3: return launch_cn<3, D>(p, st, th);
    case 5: return launch_cn<5, D>(p, st, th);
    case 7: return launch_cn<7, D>(p, st, th);
    default: return false;
    }
}

template <int MODE>
void launch_direct_tb(const StencilDirectArgs &a, int tb, hipStream_t st) {
    const dim3 g(tile_grid(a.ntiles)), b(64 * a.wpg);
    switch (tb) {
    case 4: hipLaunchKernelGGL((sdirect::sep_direct_u8<MODE, 4>), g, b, 0, st, a); break;
    case 2: hipLaunchKernelGGL((sdirect::sep_direct_u8<MODE, 2>), g, b, 0, st, a); break;
    case 8: hipLaunchKernelGGL((sdirect::sep_direct_u8<MODE, 8>), g, b, 0, st, a); break;
    case 16: hipLaunchKernelGGL((sdirect::sep_direct_u8<MODE, 16>), g, b, 0, st, a); break;
    default: hipLaunchKernelGGL((sdirect::sep_direct_u8<MODE, 1>), g, b, 0, st, a); break;
    }
}

// The streamed one-channel kernels (stencil_direct.hpp) where the layout
// allows them: rows of a multiple of 4 bytes (at least 8) starting on
// dwords, images spanning less than 2^31 bytes, a border other than WRAP.
bool launch_direct(const StencilArgs &p, int mode, hipStream_t st) {
    if (tuning().stencil_direct == 1 || p.src.cn != 1) return false;
    if (p.border != BORDER_CONSTANT && p.border != BORDER_REPLICATE && p.border != BORDER_REFLECT &&
        p.border != BORDER_REFLECT101)
        return false;
    const DImg &s = p.src, &d = p.dst;
    const long n = s.w;
    const long ob = mode == SWAR_SOBEL_DX || mode == SWAR_SOBEL_DY ? 2 : 1;
    if (n < 8 || n % 4 != 0 || s.h < 1 || d.w != s.w || d.h != s.h) return false;
    if ((((uintptr_t)s.p | (uintptr_t)d.p) & 3) != 0 || s.stride % 4 != 0 || d.stride % 4 != 0) return false;
    const long span_s = (long)(s.h - 1) * s.stride + n, span_d = (long)(d.h - 1) * d.stride + n * ob;
    if (span_s >= (1L << 31) || span_d >= (1L << 31)) return false;
    const int wg = tuning().stencil_wg;
    const int strips = (int)((n + sdirect::kStrip - 1) / sdirect::kStrip);
    // Band height: the tallest of 4, 2, 1 rows that still gives about 4000
    // waves (measured at 640x480 to 4096x4096: shorter bands issue every
    // row's loads sooner and keep more waves in flight; taller ones reread
    // fewer halo rows).
    int tb = tuning().stencil_band;
    if (tb != 1 && tb != 2 && tb != 4 && tb != 8 && tb != 16) {
        tb = 4;
        while (tb > 1 && (long)strips * ((s.h + tb - 1) / tb) < 4000) tb /= 2;
    }
    StencilDirectArgs a;
    a.s = p;
    a.span_s = (uint32_t)span_s;
    a.span_d = (uint32_t)span_d;
    a.strips = strips;
    a.bands = (s.h + tb - 1) / tb;
    a.wpg = wg == 1 || wg == 2 ? wg : 4;
    const TileOrder o = tile_order(a.strips, (a.bands + a.wpg - 1) / a.wpg, true);
    a.ntiles = o.ntiles;
    a.xcd = o.xcd;
    switch (mode) {
    case SWAR_G3: launch_direct_tb<SWAR_G3>(a, tb, st); break;
    case SWAR_G5: launch_direct_tb<SWAR_G5>(a, tb, st); break;
    case SWAR_G7: launch_direct_tb<SWAR_G7>(a, tb, st); break;
    case SWAR_SOBEL_DX: launch_direct_tb<SWAR_SOBEL_DX>(a, tb, st); break;
    case SWAR_SOBEL_DY: launch_direct_tb<SWAR_SOBEL_DY>(a, tb, st); break;
    default: return false;
    }
    return true;
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

bool stencil_args(const SepParams &p, StencilArgs &a) {
    if (p.f32mode || p.sdepth != 0 || p.nkx != p.nky || p.src.cn > 4) return false;
    if (p.nkx != 3 && p.nkx != 5 && p.nkx != 7) return false;
    if (!(p.ddepth == 0 && p.fin == FIN_FIX8) && !(p.ddepth == 3 && (p.fin == FIN_CAST || p.fin == FIN_SCALED)))
        return false;
    a.src = p.src;
    a.dst = p.dst;
    a.border = p.border;
    a.fin = p.fin;
    a.scale = p.scale;
    for (int i = 0; i < 7; i++) {
        a.kx[i] = i < p.nkx ? p.kxi[i] : 0;
        a.ky[i] = i < p.nky ? p.kyi[i] : 0;
    }
    return true;
}

// The packed 16-bit mode whose taps the call has exactly, or -1.
int swar_mode(const StencilArgs &a, int k, int ddepth) {
    auto is = [&](const int *t, std::initializer_list<int> v) {
        int i = 0;
        for (int x : v)
            if (t[i++] != x) return false;
        return true;
    };
    if (a.src.cn != 1 && a.src.cn != 3) return -1;
    if (ddepth == 0 && a.fin == FIN_FIX8) {
        if (k == 3 && is(a.kx, {64, 128, 64}) && is(a.ky, {64, 128, 64})) return SWAR_G3;
        if (k == 5 && is(a.kx, {16, 64, 96, 64, 16}) && is(a.ky, {16, 64, 96, 64, 16})) return SWAR_G5;
        if (k == 7 && is(a.kx, {8, 28, 56, 72, 56, 28, 8}) && is(a.ky, {8, 28, 56, 72, 56, 28, 8})) return SWAR_G7;
    }
    if (ddepth == 3 && a.fin == FIN_CAST && k == 3) {
        if (is(a.kx, {-1, 0, 1}) && is(a.ky, {1, 2, 1})) return SWAR_SOBEL_DX;
        if (is(a.kx, {1, 2, 1}) && is(a.ky, {-1, 0, 1})) return SWAR_SOBEL_DY;
    }
    return -1;
}

// Returns false when the call is not a shape this kernel covers.
template <int MODE, int CN, int TB>
static void launch_stream_ad(const SStreamArgs &a, bool ad, hipStream_t st) {
    const dim3 g(tile_grid(a.ntiles)), b(256);
    if (ad) hipLaunchKernelGGL((sstream::sep_stream_u8<sstream::GaussOp<MODE>, CN, TB, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((sstream::sep_stream_u8<sstream::GaussOp<MODE>, CN, TB, false>), g, b, 0, st, a);
}

template <int MODE>
static void launch_stream_mode(const SStreamArgs &a, int cn, int tb, bool ad, hipStream_t st) {
    if (cn == 1) {
        if (tb == 2) launch_stream_ad<MODE, 1, 2>(a, ad, st);
        else launch_stream_ad<MODE, 1, 4>(a, ad, st);
    } else {
        if (tb == 4) launch_stream_ad<MODE, 3, 4>(a, ad, st);
        else launch_stream_ad<MODE, 3, 2>(a, ad, st);
    }
}

bool stencil_stream_args(const StencilArgs &p, int tb, SStreamArgs &a) {
    const DImg &s = p.src, &d = p.dst;
    if (s.cn != 1 && s.cn != 3) return false;
    if (s.w < 1 || s.h < 1 || d.w != s.w || d.h != s.h || d.cn != s.cn || s.stride < 0 || d.stride < 0) return false;
    a.s = p;
    a.off_s = (uint32_t)((uintptr_t)s.p & 3);
    a.off_d = (uint32_t)((uintptr_t)d.p & 3);
    const long lim_s = a.off_s + (long)(s.h - 1) * s.stride + (long)s.w * s.cn;
    const long lim_d = a.off_d + (long)(d.h - 1) * d.stride + (long)d.w * d.cn;
    if (lim_s >= (1L << 31) || lim_d >= (1L << 31)) return false;
    a.s.src.p = s.p - a.off_s;
    a.s.dst.p = d.p - a.off_d;
    a.lim_s = (uint32_t)lim_s;
    a.lim_d = (uint32_t)lim_d;
    a.strips = (s.w + 6 + sstream::kStripPx - 1) / sstream::kStripPx;
    a.bands = (s.h + tb - 1) / tb;
    return true;
}

// The general streamed kernels (stencil_stream.hpp): GaussianBlur 3x3, 5x5,
// 7x7 on one or three channels, any alignment, width and border. Bands of
// 4 rows on one channel, 2 on three (the tuning's stencil_band 2 or 4
// overrides).
static bool launch_stream(const StencilArgs &p, int mode, hipStream_t st) {
    if (tuning().stencil_direct == 1 || (mode != SWAR_G3 && mode != SWAR_G5 && mode != SWAR_G7)) return false;
    const int k = tuning().stencil_band;
    const int tb = k == 2 || k == 4 ? k : p.src.cn == 1 ? 4 : 2;
    SStreamArgs a;
    if (!stencil_stream_args(p, tb, a)) return false;
    const TileOrder o = tile_order(a.strips, (a.bands + 3) / 4, true);
    a.ntiles = o.ntiles;
    a.xcd = o.xcd;
    const bool ad = a.off_d == 0 && p.dst.stride % 4 == 0;
    switch (mode) {
    case SWAR_G3: launch_stream_mode<SWAR_G3>(a, p.src.cn, tb, ad, st); break;
    case SWAR_G5: launch_stream_mode<SWAR_G5>(a, p.src.cn, tb, ad, st); break;
    default: launch_stream_mode<SWAR_G7>(a, p.src.cn, tb, ad, st); break;
    }
    return true;
}

bool launch_sep_small_u8(const SepParams &p, hipStream_t st, int th) {
    StencilArgs a;
    if (!stencil_args(p, a)) return false;
    const int mode = swar_mode(a, p.nkx, p.ddepth);
    if (mode >= 0 && launch_direct(a, mode, st)) return hipGetLastError() == hipSuccess;
    if (mode >= 0 && launch_stream(a, mode, st)) return hipGetLastError() == hipSuccess;
    bool ok;
    switch (mode) {
    case SWAR_G3: ok = launch_swar<SWAR_G3, uint8_t>(a, st, th); break;
    case SWAR_G5: ok = launch_swar<SWAR_G5, uint8_t>(a, st, th); break;
    case SWAR_G7: ok = launch_swar<SWAR_G7, uint8_t>(a, st, th); break;
    case SWAR_SOBEL_DX: ok = launch_swar<SWAR_SOBEL_DX, int16_t>(a, st, th); break;
    case SWAR_SOBEL_DY: ok = launch_swar<SWAR_SOBEL_DY, int16_t>(a, st, th); break;
    default: ok = p.ddepth == 0 ? launch_d<uint8_t>(a, p.nkx, st, th) : launch_d<int16_t>(a, p.nkx, st, th);
    }
    return ok && hipGetLastError() == hipSuccess;
}

}  // namespace visar
