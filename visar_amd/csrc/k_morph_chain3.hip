// k_morph_chain3.hip -- fused erode/dilate chains (morph_chain.hpp) for
// three- and four-channel images and the elements with codes of their own.
#include "morph_chain.hpp"

namespace visar {

bool launch_morph_chain_el34(const MorphChainArgs &a, uint32_t el, size_t lds, hipStream_t st) {
    return a.s.cn == 3 ? mchain::launch_chain_el<3>(a, el, lds, st) : mchain::launch_chain_el<4>(a, el, lds, st);
}

}  // namespace visar
