// k_stencil.hip -- fast path of the separable engine: u8 sources with square
// 3/5/7 kernels (GaussianBlur u8, Sobel u8 -> i16), the shapes of every
// benchmark configuration. The per-tile work is stencil_tile (tiles.hpp),
// shared with the persistent batch kernel.
#include <initializer_list>

#include "tiles.hpp"

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
bool launch_sep_small_u8(const SepParams &p, hipStream_t st, int th) {
    StencilArgs a;
    if (!stencil_args(p, a)) return false;
    bool ok;
    switch (swar_mode(a, p.nkx, p.ddepth)) {
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
