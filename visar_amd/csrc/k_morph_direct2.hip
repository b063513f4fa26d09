// k_morph_direct2.hip -- the two-pass kernels of morph_direct.hpp.
#include "morph_direct.hpp"

namespace visar {

bool launch_morph_direct2(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st) {
    return mdirect::launch_direct_el<2>(a, el, post, st);
}

}  // namespace visar
