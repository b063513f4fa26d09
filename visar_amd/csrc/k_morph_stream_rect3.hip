// k_morph_stream_rect3.hip -- the streamed passes of morph_stream.hpp for the
// base element rect3 (one file per element, so they compile in parallel),
// and the dispatch over the elements.
#include "morph_stream.hpp"

namespace visar {

bool launch_morph_stream_rect3(const MorphStreamArgs &a, int n, int mode, int post, hipStream_t st) {
    return mstream::launch_base<mchain::kElRect3>(a, n, mode, post, st);
}

bool launch_morph_stream(const MorphStreamArgs &a, uint32_t el, int n, int mode, int post, hipStream_t st) {
    switch (el) {
    case mchain::kElRect3: return launch_morph_stream_rect3(a, n, mode, post, st);
    case mchain::kElRect5: return launch_morph_stream_rect5(a, n, mode, post, st);
    case mchain::kElRect7: return launch_morph_stream_rect7(a, n, mode, post, st);
    case mchain::kElCross3: return launch_morph_stream_cross3(a, n, mode, post, st);
    case mchain::kElCross5: return launch_morph_stream_cross5(a, n, mode, post, st);
    case mchain::kElCross7: return launch_morph_stream_cross7(a, n, mode, post, st);
    case mchain::kElEllipse5: return launch_morph_stream_ellipse5(a, n, mode, post, st);
    case mchain::kElEllipse7: return launch_morph_stream_ellipse7(a, n, mode, post, st);
    default: return false;
    }
}

}  // namespace visar
