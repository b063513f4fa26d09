// k_morph_direct_max.hip -- the dilate kernels of morph_direct.hpp.
#include "morph_direct.hpp"

namespace visar {

bool launch_morph_direct_max(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st) {
    return mdirect::launch_direct_el<true>(a, el, post, st);
}

}  // namespace visar
