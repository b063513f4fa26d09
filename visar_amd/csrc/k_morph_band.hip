// k_morph_band.hip -- the single-pass kernels of morph_band.hpp.
#include "morph_band.hpp"

namespace visar {

namespace {

template <uint32_t EL, bool MAX, int POST>
void launch_tb(MorphDirectArgs a, int tb, hipStream_t st) {
    a.strips = (a.s.w + mband::kStrip - 1) / mband::kStrip;
    a.bands4 = ((a.s.h + tb - 1) / tb + 3) / 4;
    a.ntiles = a.strips * a.bands4;
    const dim3 g(tile_grid(a.ntiles));
    switch (tb) {
    case 8: hipLaunchKernelGGL((mband::morph_band_u8<EL, MAX, POST, 8>), g, dim3(256), 0, st, a); break;
    case 16: hipLaunchKernelGGL((mband::morph_band_u8<EL, MAX, POST, 16>), g, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((mband::morph_band_u8<EL, MAX, POST, 4>), g, dim3(256), 0, st, a); break;
    }
}

template <uint32_t EL>
bool launch_el(const MorphDirectArgs &a, int op_max, int post, int tb, hipStream_t st) {
    if (op_max && post == 0) return launch_tb<EL, true, 0>(a, tb, st), true;
    if (!op_max && post == 0) return launch_tb<EL, false, 0>(a, tb, st), true;
    if (op_max && post == 1) return launch_tb<EL, true, 1>(a, tb, st), true;
    if (!op_max && post == 1) return launch_tb<EL, false, 1>(a, tb, st), true;
    if (!op_max && post == 2) return launch_tb<EL, false, 2>(a, tb, st), true;
    return false;
}

}  // namespace

bool launch_morph_band(const MorphDirectArgs &a, uint32_t el, int op_max, int post, hipStream_t st) {
    using namespace mchain;
    // 4 output rows per wave by default: on the 63 morphology cases 4, 8 and
    // 16 rows measured 1.95, 1.97 and 2.06 ms (morph_direct.hpp's passes 2.00).
    const int t = tuning().morph_pass_rows;
    const int tb = t == 8 || t == 16 ? t : 4;
    switch (el) {
    case kElRect3: return launch_el<kElRect3>(a, op_max, post, tb, st);
    case kElRect5: return launch_el<kElRect5>(a, op_max, post, tb, st);
    case kElRect7: return launch_el<kElRect7>(a, op_max, post, tb, st);
    case kElCross3: return launch_el<kElCross3>(a, op_max, post, tb, st);
    case kElCross5: return launch_el<kElCross5>(a, op_max, post, tb, st);
    case kElCross7: return launch_el<kElCross7>(a, op_max, post, tb, st);
    case kElEllipse5: return launch_el<kElEllipse5>(a, op_max, post, tb, st);
    case kElEllipse7: return launch_el<kElEllipse7>(a, op_max, post, tb, st);
    default: return false;
    }
}

}  // namespace visar
