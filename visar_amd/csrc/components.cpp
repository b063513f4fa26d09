// components.cpp -- connectedComponentsWithStats: the C ABI entry point
// (labelled on the GPU by k_components.hip, finished here) and a host
// labeller used by the served front end's CPU workers.
//
// Reference: src/imgproc/components.zig:408-535, C ABI capi.zig:533-550.
// Labels follow OpenCV: 4-connected components are numbered in raster order
// of their first pixel, 8-connected ones by their first 2x2 block in
// block-raster order. Works on runs: each row's nonzero runs are joined to
// the overlapping runs of the row above through a union-find whose roots are
// the earliest run in raster order; statistics are integer sums (twice the
// x sum) so they do not depend on the order runs are visited.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/visar.h"
#include "components.hpp"
#include "components_gpu.hpp"
#include "runtime.hpp"

namespace visar {

namespace {

struct Run {
    uint32_t y, x0, x1;
};

inline uint32_t find(std::vector<uint32_t> &p, uint32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

inline void unite(std::vector<uint32_t> &p, uint32_t a, uint32_t b) {
    a = find(p, a);
    b = find(p, b);
    if (a == b) return;
    if (a < b) p[b] = a;
    else p[a] = b;
}

}  // namespace

size_t label_components(const uint8_t *src, size_t sstride, int32_t *labels, size_t lstride, uint32_t w, uint32_t h,
                        bool eight, std::vector<CompStats> &out) {
    std::vector<Run> runs;
    std::vector<uint32_t> row_start(h + 1, 0);
    runs.reserve(1024);
    for (uint32_t y = 0; y < h; y++) {
        row_start[y] = (uint32_t)runs.size();
        const uint8_t *r = src + (size_t)y * sstride;
        uint32_t x = 0;
        while (x < w) {
            while (x < w && !r[x]) x++;
            if (x >= w) break;
            uint32_t x0 = x;
            while (x < w && r[x]) x++;
            runs.push_back({y, x0, x - 1});
        }
    }
    row_start[h] = (uint32_t)runs.size();
    const uint32_t n = (uint32_t)runs.size();
    std::vector<uint32_t> parent(n);
    for (uint32_t i = 0; i < n; i++) parent[i] = i;
    const long slack = eight ? 1 : 0;
    for (uint32_t y = 1; y < h; y++) {
        uint32_t p = row_start[y - 1], pe = row_start[y];
        for (uint32_t i = row_start[y]; i < row_start[y + 1]; i++) {
            const Run &r = runs[i];
            while (p < pe && (long)runs[p].x1 + slack < (long)r.x0) p++;
            for (uint32_t q = p; q < pe && (long)runs[q].x0 <= (long)r.x1 + slack; q++) unite(parent, i, q);
        }
    }
    std::vector<uint32_t> root(n);
    for (uint32_t i = 0; i < n; i++) root[i] = find(parent, i);
    // Final label of each root run.
    std::vector<uint32_t> lab(n, 0);
    uint32_t next = 1;
    if (!eight) {
        for (uint32_t i = 0; i < n; i++)
            if (root[i] == i) lab[i] = next++;
    } else {
        // Per block row: roots whose first run lies in it, ordered by the
        // smallest block column of their runs inside the block row.
        std::vector<uint32_t> key(n, 0), roots;
        for (uint32_t by = 0; 2 * by < h; by++) {
            uint32_t a = row_start[2 * by], b = row_start[std::min(h, 2 * by + 2)];
            roots.clear();
            for (uint32_t i = a; i < b; i++) {
                uint32_t rt = root[i], k = runs[i].x0 >> 1;
                if (rt == i) {
                    key[i] = k;
                    roots.push_back(i);
                } else if (rt >= a && k < key[rt]) {
                    key[rt] = k;
                }
            }
            std::sort(roots.begin(), roots.end(), [&](uint32_t u, uint32_t v) { return key[u] < key[v]; });
            for (uint32_t r : roots) lab[r] = next++;
        }
    }
    out.assign(next, CompStats());
    for (auto &s : out) {
        s.x0 = s.y0 = std::numeric_limits<uint64_t>::max();
        s.x1 = s.y1 = s.area = s.sx2 = s.sy = 0;
    }
    CompStats &bg = out[0];
    for (uint32_t y = 0; y < h; y++) {
        int32_t *lr = labels ? (int32_t *)((uint8_t *)labels + (size_t)y * lstride) : nullptr;
        uint32_t x = 0;
        uint64_t fg = 0, fgx2 = 0;
        const uint32_t a = row_start[y], b = row_start[y + 1];
        for (uint32_t i = a; i < b; i++) {
            const Run &r = runs[i];
            uint32_t l = lab[root[i]];
            if (lr) {
                for (; x < r.x0; x++) lr[x] = 0;
                for (; x <= r.x1; x++) lr[x] = (int32_t)l;
            }
            uint64_t len = (uint64_t)r.x1 - r.x0 + 1;
            CompStats &s = out[l];
            s.x0 = std::min<uint64_t>(s.x0, r.x0);
            s.x1 = std::max<uint64_t>(s.x1, r.x1);
            s.y0 = std::min<uint64_t>(s.y0, y);
            s.y1 = std::max<uint64_t>(s.y1, y);
            s.area += len;
            s.sx2 += len * ((uint64_t)r.x0 + r.x1);
            s.sy += len * y;
            fg += len;
            fgx2 += len * ((uint64_t)r.x0 + r.x1);
        }
        if (lr)
            for (; x < w; x++) lr[x] = 0;
        uint64_t zeros = (uint64_t)w - fg;
        if (!zeros) continue;
        uint64_t first = (a < b && runs[a].x0 == 0) ? runs[a].x1 + 1 : 0;
        uint64_t last = (a < b && runs[b - 1].x1 == w - 1) ? runs[b - 1].x0 - 1 : (uint64_t)w - 1;
        bg.x0 = std::min(bg.x0, first);
        bg.x1 = std::max(bg.x1, last);
        bg.y0 = std::min<uint64_t>(bg.y0, y);
        bg.y1 = std::max<uint64_t>(bg.y1, y);
        bg.area += zeros;
        bg.sx2 += (uint64_t)w * (w - 1) - fgx2;
        bg.sy += zeros * y;
    }
    return next;
}

void write_stats(const std::vector<CompStats> &st, int32_t *stats, double *centroids, size_t max) {
    for (size_t i = 0; i < st.size() && i < max; i++) {
        const CompStats &s = st[i];
        if (stats) {
            int32_t *o = stats + i * 5;
            if (s.area == 0) {
                o[0] = -1, o[1] = std::numeric_limits<int32_t>::max(), o[2] = 0, o[3] = 0, o[4] = 0;
            } else {
                o[0] = (int32_t)s.x0, o[1] = (int32_t)s.y0, o[2] = (int32_t)(s.x1 - s.x0 + 1),
                o[3] = (int32_t)(s.y1 - s.y0 + 1), o[4] = (int32_t)s.area;
            }
        }
        if (centroids) {
            if (s.area == 0) {
                centroids[i * 2] = centroids[i * 2 + 1] = std::nan("");
            } else {
                double a = (double)s.area;
                centroids[i * 2] = (double)s.sx2 / 2.0 / a;
                centroids[i * 2 + 1] = (double)s.sy / a;
            }
        }
    }
}

}  // namespace visar

using namespace visar;

extern "C" int visar_connected_components(const visar_image *src, const visar_image *labels, int connectivity,
                                          int32_t *stats, double *centroids, size_t max, size_t *count) {
    int rc;
    if ((rc = check_image(src, 0)) || (rc = check_image(labels, 4))) return rc;
    if (!count) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (src->channels != 1 || labels->channels != 1 || src->width != labels->width || src->height != labels->height)
        return fail(SIZE_MISMATCH, "SizeMismatch");
    const uint32_t w = src->width, h = src->height;
    if ((uint64_t)w * h >= (1ull << 31)) return fail(UNSUPPORTED, "Unsupported");
    // Labelled on the GPU (k_components.hip); the host turns the integer
    // sums into the statistics rows and centroids.
    Call call;
    DImg s, lv;
    if ((rc = call.begin())) return rc;
    if ((rc = call.input(src, s)) || (rc = call.output(labels, lv))) return rc;
    CcArgs c{};
    c.src = s.p;
    c.sstride = (size_t)s.stride;
    c.w = w;
    c.h = h;
    c.eight = connectivity != 4;
    c.labels = reinterpret_cast<int32_t *>(lv.p);
    c.lstride = (size_t)lv.stride;
    void *scratch = call.scratch(cc_scratch_bytes(w, h));
    c.host_total = (uint32_t *)call.pinned(sizeof(uint32_t));
    if (!scratch || !c.host_total) return fail(OUT_OF_MEMORY, "OutOfMemory");
    uint32_t n = 0;
    if (cc_label(c, scratch, call.stream(), &n) != OK) return call.finish(INTERNAL);
    // The statistics come back through pinned memory.
    uint64_t *acc = (uint64_t *)call.pinned(4 * (size_t)n * 8), *rows = (uint64_t *)call.pinned(4 * (size_t)h * 8);
    uint32_t *box = (uint32_t *)call.pinned(4 * (size_t)n * 4);
    if (!acc || !rows || !box) return fail(OUT_OF_MEMORY, "OutOfMemory");
    hipStream_t st = call.stream();
    if (hipMemcpyAsync(acc, c.acc, 4 * (size_t)n * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(box, c.box, 4 * (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(rows, c.rows, 4 * (size_t)h * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return call.finish(INTERNAL);
    std::vector<CompStats> out(n);
    CompStats &bg = out[0];
    bg.x0 = bg.y0 = std::numeric_limits<uint64_t>::max();
    bg.x1 = bg.y1 = bg.area = bg.sx2 = bg.sy = 0;
    for (uint32_t y = 0; y < h; y++) {
        const uint64_t zeros = rows[4 * (size_t)y];
        if (!zeros) continue;
        bg.x0 = std::min(bg.x0, rows[4 * (size_t)y + 1]);
        bg.x1 = std::max(bg.x1, rows[4 * (size_t)y + 2]);
        bg.y0 = std::min<uint64_t>(bg.y0, y);
        bg.y1 = std::max<uint64_t>(bg.y1, y);
        bg.area += zeros;
        bg.sx2 += (uint64_t)w * (w - 1) - rows[4 * (size_t)y + 3];
        bg.sy += zeros * y;
    }
    for (uint32_t l = 1; l < n; l++) {
        CompStats &o = out[l];
        o.x0 = box[4 * (size_t)l];
        o.y0 = box[4 * (size_t)l + 1];
        o.x1 = box[4 * (size_t)l + 2];
        o.y1 = box[4 * (size_t)l + 3];
        o.area = acc[4 * (size_t)l];
        o.sx2 = acc[4 * (size_t)l + 1];
        o.sy = acc[4 * (size_t)l + 2];
    }
    *count = n;
    write_stats(out, stats, centroids, max);
    return call.finish(OK);
}
