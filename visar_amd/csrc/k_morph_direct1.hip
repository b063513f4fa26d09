// k_morph_direct1.hip -- the one-pass kernels of morph_direct.hpp.
#include "morph_direct.hpp"

namespace visar {

bool launch_morph_direct1(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st) {
    return mdirect::launch_direct_el<1>(a, el, post, st);
}

}  // namespace visar
