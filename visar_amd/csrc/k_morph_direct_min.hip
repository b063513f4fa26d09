// k_morph_direct_min.hip -- the erode kernels of morph_direct.hpp.
#include "morph_direct.hpp"

namespace visar {

bool launch_morph_direct_min(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st) {
    return mdirect::launch_direct_el<false>(a, el, post, st);
}

}  // namespace visar
