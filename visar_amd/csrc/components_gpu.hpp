// components_gpu.hpp -- connectedComponentsWithStats on the GPU (k_components.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace visar {

struct CcArgs {
    const uint8_t *src;  // device mask, nonzero = foreground
    size_t sstride;
    uint32_t w, h;
    int eight;           // 8-connectivity
    int32_t *labels;     // device label image, or null
    size_t lstride;
    // filled by cc_label
    uint32_t n, kw;
    uint32_t *L, *key, *lab, *mark, *bsum;
    uint64_t *acc;  // per label: area, twice the x sum, y sum, unused
    uint32_t *box;  // per label: min x, min y, max x, max y
    uint64_t *rows; // per row: zeros, first zero, last zero, twice the nonzero x sum
    uint32_t *host_total;  // pinned host word for the label count (set by the caller)
};

// Device scratch one cc_label call needs for a w x h mask.
size_t cc_scratch_bytes(uint32_t w, uint32_t h);

// Labels the mask on `st` into `scratch`; waits once for the label count
// (including the background). The statistics are in c.acc / c.box (labels
// 1 .. count - 1) and c.rows when the stream has drained.
int cc_label(CcArgs &c, void *scratch, hipStream_t st, uint32_t *count);

}  // namespace visar
