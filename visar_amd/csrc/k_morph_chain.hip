// k_morph_chain.hip -- fused erode/dilate chains (morph_chain.hpp): the
// kernels that read the element from their arguments, and the launcher.
#include "morph_chain.hpp"

namespace visar {

using namespace mchain;

namespace {

template <int CN>
int launch_chain_cn(const MorphChainArgs &a, int K, size_t lds, hipStream_t st) {
    switch (K) {
    case 1: launch_chain<CN, 1, 0>(a, lds, st); break;
    case 2: launch_chain<CN, 2, 0>(a, lds, st); break;
    case 3: launch_chain<CN, 3, 0>(a, lds, st); break;
    default: return UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

}  // namespace

size_t morph_chain_lds_limit() {
    static size_t cache[16] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (dev < 0 || dev >= 16) return 65536;
    if (!cache[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
        cache[dev] = (size_t)v;
    }
    return cache[dev];
}

int morph_chain_pitch_dw(int cn, int K) {
    const int hq = (K * cn + 3) / 4;
    return kRowDw + 2 * ((hq + 1) & ~1);
}

int launch_morph_chain(const MorphChainArgs &a, int K, size_t lds, hipStream_t st) {
    // The element's code: its rectangles packed as in RectOf.
    uint32_t el = (uint32_t)a.nr;
    for (int k = 0; k < a.nr; k++) el |= (uint32_t)(a.rw[k] | a.rv[k] << 2) << (4 + 4 * k);
    const int cn = a.s.cn;
    if ((cn == 1 && launch_morph_chain_el1(a, el, lds, st)) || ((cn == 3 || cn == 4) && launch_morph_chain_el34(a, el, lds, st)))
        return hipGetLastError() == hipSuccess ? OK : INTERNAL;
    switch (cn) {
    case 1: return launch_chain_cn<1>(a, K, lds, st);
    case 2: return launch_chain_cn<2>(a, K, lds, st);
    case 3: return launch_chain_cn<3>(a, K, lds, st);
    case 4: return launch_chain_cn<4>(a, K, lds, st);
    default: return UNSUPPORTED;
    }
}

}  // namespace visar
