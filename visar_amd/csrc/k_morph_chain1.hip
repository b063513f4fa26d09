// k_morph_chain1.hip -- fused erode/dilate chains (morph_chain.hpp) for
// one-channel images and the elements with codes of their own.
#include "morph_chain.hpp"

namespace visar {

bool launch_morph_chain_el1(const MorphChainArgs &a, uint32_t el, size_t lds, hipStream_t st) {
    return mchain::launch_chain_el<1>(a, el, lds, st);
}

}  // namespace visar
