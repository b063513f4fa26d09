// k_morph_pow.hip -- the n-fold single passes of morph_pow.hpp.
#include "morph_pow.hpp"

namespace visar {

namespace {

template <uint32_t BASE, int N, bool MAX, int POST>
void launch_pow(MorphDirectArgs a, hipStream_t st) {
    constexpr int SW = 4 * (64 - 2 * mpow::halo_lanes(mpow::Pow<BASE, N>::A));
    // 8 output rows per wave by default: on the 63 morphology cases 8 rows
    // measured 1.701-1.706 ms, 4 rows 1.705-1.725 ms.
    const int t = tuning().morph_pow_rows;
    const int tb = t == 4 ? 4 : 8;
    a.strips = (a.s.w + SW - 1) / SW;
    a.bands4 = ((a.s.h + tb - 1) / tb + 3) / 4;
    const TileOrder o = tile_order(a.strips, a.bands4, true);
    a.ntiles = o.ntiles;
    a.xcd = o.xcd;
    const dim3 g(tile_grid(a.ntiles));
    if (tb == 8) hipLaunchKernelGGL((mpow::morph_pow_u8<BASE, N, MAX, POST, 8>), g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((mpow::morph_pow_u8<BASE, N, MAX, POST, 4>), g, dim3(256), 0, st, a);
}

template <uint32_t BASE, int N>
bool launch_n(const MorphDirectArgs &a, int op_max, int post, hipStream_t st) {
    if constexpr (mpow::fits<BASE, N>()) {
        if (post == 0) op_max ? launch_pow<BASE, N, true, 0>(a, st) : launch_pow<BASE, N, false, 0>(a, st);
        else if (post == 1) op_max ? launch_pow<BASE, N, true, 1>(a, st) : launch_pow<BASE, N, false, 1>(a, st);
        else op_max ? launch_pow<BASE, N, true, 2>(a, st) : launch_pow<BASE, N, false, 2>(a, st);
        return true;
    } else {
        return false;
    }
}

// a == nullptr: only whether a launch exists.
template <uint32_t BASE>
bool launch_base(const MorphDirectArgs *a, int n, int op_max, int post, hipStream_t st) {
    if (!a) return n == 2 ? mpow::fits<BASE, 2>() : n == 3 ? mpow::fits<BASE, 3>() : false;
    return n == 2 ? launch_n<BASE, 2>(*a, op_max, post, st) : n == 3 ? launch_n<BASE, 3>(*a, op_max, post, st) : false;
}

bool pow_switch(const MorphDirectArgs *a, uint32_t el, int n, int op_max, int post, hipStream_t st) {
    using namespace mchain;
    switch (el) {
    case kElRect3: return launch_base<kElRect3>(a, n, op_max, post, st);
    case kElRect5: return launch_base<kElRect5>(a, n, op_max, post, st);
    case kElRect7: return launch_base<kElRect7>(a, n, op_max, post, st);
    case kElCross3: return launch_base<kElCross3>(a, n, op_max, post, st);
    case kElCross5: return launch_base<kElCross5>(a, n, op_max, post, st);
    case kElCross7: return launch_base<kElCross7>(a, n, op_max, post, st);
    case kElEllipse5: return launch_base<kElEllipse5>(a, n, op_max, post, st);
    case kElEllipse7: return launch_base<kElEllipse7>(a, n, op_max, post, st);
    default: return false;
    }
}

}  // namespace

bool morph_pow_has(uint32_t el, int n) { return pow_switch(nullptr, el, n, 0, 0, nullptr); }

bool launch_morph_pow(const MorphDirectArgs &a, uint32_t el, int n, int op_max, int post, hipStream_t st) {
    return pow_switch(&a, el, n, op_max, post, st);
}

}  // namespace visar
