// components.hpp -- host connected-component labelling with statistics (the
// served front end's CPU workers) and the C ABI's statistics rows.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace visar {

// Integer accumulators of one label (components.zig:23-31, 178-210).
struct CompStats {
    uint64_t x0, y0, x1, y1, area, sx2, sy;
};

// Labels the nonzero pixels of a w x h u8 mask (row stride `sstride`);
// writes i32 labels when `labels` is not null; returns the number of labels
// including the background, with per-label statistics in `out`.
size_t label_components(const uint8_t *src, size_t sstride, int32_t *labels, size_t lstride, uint32_t w, uint32_t h,
                        bool eight, std::vector<CompStats> &out);

// The C ABI's stats rows (left, top, width, height, area) and centroids.
void write_stats(const std::vector<CompStats> &st, int32_t *stats, double *centroids, size_t max);

}  // namespace visar
