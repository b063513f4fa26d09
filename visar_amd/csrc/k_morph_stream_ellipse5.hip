// k_morph_stream_ellipse5.hip -- the streamed passes of morph_stream.hpp for the
// base element ellipse5 (one file per element, so they compile in parallel).
#include "morph_stream.hpp"

namespace visar {

bool launch_morph_stream_ellipse5(const MorphStreamArgs &a, int n, int mode, int post, hipStream_t st) {
    return mstream::launch_base<mchain::kElEllipse5>(a, n, mode, post, st);
}

}  // namespace visar
