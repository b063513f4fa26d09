// batch.hip -- many images, mixed operations, one persistent launch.
//
// A plan turns each item into a job: a tile function of tiles.hpp (the same
// code the plain entry points launch, so results are identical), the job's
// tile grid and its arguments. Jobs are grouped by tile variant (kind,
// kernel size, channels, rows per thread); each group gets its own
// persistent kernel, specialised for that one variant so its registers and
// code are only what the variant needs, and the groups' kernels run back to
// back in one launch (one HIP graph). Each job's tiles are cut into work
// units of about the same number of bytes (a run of consecutive tiles of one
// job), and a group's unit list is cut into 8 contiguous segments of equal
// bytes, one per XCD: workgroup w is dispatched to XCD w % 8 and first drains
// that XCD's segment (neighbouring tiles of an image, and their halo rows,
// stay in one L2), then steals units from the other segments until all are
// done. Every queue is a counter bumped with atomicAdd; a workgroup claims its
// next unit before it works through the current one, so the atomic's latency
// hides behind the work. A small kernel clears the counters before each run
// (no hipMemsetAsync, so the run can be captured in a HIP graph).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/visar_batch.h"
#include "runtime.hpp"
#include "tables.hpp"
#include "tiles.hpp"

namespace visar {

namespace {

constexpr int kXcds = 8;

enum JobKind : int {
    JK_STENCIL = 0,
    JK_MORPH = 1,
    JK_AFFINE_LIN4 = 2,
    JK_AFFINE_PX = 3,
    JK_HALF_C1 = 4,
    JK_HALF = 5,
    JK_LINEAR = 6,
    JK_AREA_INT = 7,
    JK_AREA_FRAC = 8,
    JK_STENCIL_SWAR = 9,
    JK_AFFINE_BOX = 10,
};

struct ResizeArgs {
    DImg s, d;
    const LinTap *xt, *yt;
    const AreaEntry *xe, *ye;
    const int32_t *xs, *ys;
    int ix, iy;
    float scale;
};

struct BatchJob {
    int kind;
    int variant;  // template selector (kernel size, channels, rows per thread, min/max)
    int tiles_x;
    int ntiles;
    union {
        StencilArgs st;
        MorphSmallArgs mo;
        AffineParams af;
        ResizeArgs rs;
    };
};

// A run of consecutive tiles [first, first + count) of job `job`.
struct BatchUnit {
    int job, first, count, pad;
};

// Drains one group's queues: seg[0..8] are the group's 8 segment bounds in
// the unit list, counters[0..7] its queue counters.
template <typename Tile>
__device__ __attribute__((always_inline)) inline void drain(const BatchJob *__restrict__ jobs, const BatchUnit *__restrict__ units,
                                                            const int *__restrict__ seg, int *counters, int *next, Tile tile) {
    const int home = blockIdx.x % kXcds;
    for (int k = 0; k < kXcds; k++) {
        const int q = (home + k) % kXcds;
        const int base = seg[q], size = seg[q + 1] - base;
        if (size <= 0) continue;
        if (threadIdx.x == 0) *next = atomicAdd(&counters[q], 1);
        __syncthreads();
        int i = __builtin_amdgcn_readfirstlane(*next);
        __syncthreads();
        while (i < size) {
            int claim = 0;
            if (threadIdx.x == 0) claim = atomicAdd(&counters[q], 1);  // the next unit, claimed early
            const BatchUnit u = units[base + i];
            const BatchJob &j = jobs[u.job];
            const int tiles_x = j.tiles_x;
            int by = u.first / tiles_x, bx = u.first - by * tiles_x;
            for (int c = 0; c < u.count; c++) {
                tile(j, bx, by);
                __syncthreads();
                if (++bx == tiles_x) {
                    bx = 0;
                    by++;
                }
            }
            if (threadIdx.x == 0) *next = claim;
            __syncthreads();
            i = __builtin_amdgcn_readfirstlane(*next);
            __syncthreads();
        }
    }
}

#define BATCH_ARGS const BatchJob *__restrict__ jobs, const BatchUnit *__restrict__ units, const int *__restrict__ seg, int *counters

template <int K, int CN, int TH>
__global__ __launch_bounds__(kTileThreads) void batch_stencil(BATCH_ARGS) {
    __shared__ uint32_t lds[StencilShape<K, CN, TH>::LDS_WORDS];
    __shared__ int next;
    drain(jobs, units, seg, counters, &next,
          [&](const BatchJob &j, int bx, int by) { stencil_tile<K, CN, TH, uint8_t>(j.st, bx, by, lds); });
}

template <int MODE, int CN, int TH>
__global__ __launch_bounds__(kTileThreads) void batch_stencil_swar(BATCH_ARGS) {
    __shared__ uint32_t lds[StencilShape<SwarTaps<MODE>::K, CN, TH>::LDS_WORDS];
    __shared__ int next;
    drain(jobs, units, seg, counters, &next,
          [&](const BatchJob &j, int bx, int by) { stencil_swar_tile<MODE, CN, TH, uint8_t>(j.st, bx, by, lds); });
}

__global__ __launch_bounds__(kTileThreads) void batch_affine_box(BATCH_ARGS) {
    __shared__ uint32_t lds[kWarpLdsWords];
    __shared__ int next;
    drain(jobs, units, seg, counters, &next,
          [&](const BatchJob &j, int bx, int by) { affine_box_tile(j.af, bx, by, lds); });
}

template <int K, int CN, bool MAX>
__global__ __launch_bounds__(kTileThreads) void batch_morph(BATCH_ARGS) {
    __shared__ uint32_t lds[MorphShape<K, CN, 2>::LDS_WORDS];
    __shared__ int next;
    drain(jobs, units, seg, counters, &next,
          [&](const BatchJob &j, int bx, int by) { morph_small_tile<K, CN, 2, MAX>(j.mo, bx, by, lds); });
}

template <int KIND>
__global__ __launch_bounds__(kTileThreads) void batch_other(BATCH_ARGS) {
    __shared__ int next;
    drain(jobs, units, seg, counters, &next, [&](const BatchJob &j, int bx, int by) {
        if constexpr (KIND == JK_AFFINE_LIN4) affine_lin4_tile(j.af, bx, by);
        else if constexpr (KIND == JK_AFFINE_PX) affine_px_tile<uint8_t>(j.af, bx, by);
        else if constexpr (KIND == JK_HALF_C1) half_c1_tile(j.rs.s, j.rs.d, bx, by);
        else if constexpr (KIND == JK_HALF) half_tile(j.rs.s, j.rs.d, bx, by);
        else if constexpr (KIND == JK_LINEAR) linear_tile(j.rs.s, j.rs.d, j.rs.xt, j.rs.yt, bx, by);
        else if constexpr (KIND == JK_AREA_INT) area_int_tile(j.rs.s, j.rs.d, j.rs.ix, j.rs.iy, j.rs.scale, bx, by);
        else area_frac_tile(j.rs.s, j.rs.d, j.rs.xe, j.rs.xs, j.rs.ye, j.rs.ys, bx, by);
    });
}

__global__ void batch_reset(int *counters, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

using GroupKernel = void (*)(BATCH_ARGS);

template <int K, int CN>
GroupKernel stencil_kernel(int th) {
    return th == 2 ? batch_stencil<K, CN, 2> : batch_stencil<K, CN, 1>;
}

template <int MODE>
GroupKernel swar_kernel(int cn, int th) {
    if (cn == 1) return th == 2 ? batch_stencil_swar<MODE, 1, 2> : batch_stencil_swar<MODE, 1, 1>;
    return th == 2 ? batch_stencil_swar<MODE, 3, 2> : batch_stencil_swar<MODE, 3, 1>;
}

template <int K, int CN>
GroupKernel morph_kernel(int mx) {
    return mx ? batch_morph<K, CN, true> : batch_morph<K, CN, false>;
}

// The kernel of a (kind, variant) group; variant = k * 100 + cn * 10 + th
// (stencils) or + max (morphology).
GroupKernel group_kernel(int kind, int variant) {
    const int k = variant / 100, cn = (variant / 10) % 10, last = variant % 10;
    switch (kind) {
    case JK_STENCIL:
        if (cn == 1) return k == 3 ? stencil_kernel<3, 1>(last) : k == 5 ? stencil_kernel<5, 1>(last) : stencil_kernel<7, 1>(last);
        return k == 3 ? stencil_kernel<3, 3>(last) : k == 5 ? stencil_kernel<5, 3>(last) : stencil_kernel<7, 3>(last);
    case JK_MORPH:
        if (cn == 1) return k == 3 ? morph_kernel<3, 1>(last) : k == 5 ? morph_kernel<5, 1>(last) : morph_kernel<7, 1>(last);
        return k == 3 ? morph_kernel<3, 3>(last) : k == 5 ? morph_kernel<5, 3>(last) : morph_kernel<7, 3>(last);
    case JK_STENCIL_SWAR: return k == SWAR_G3 ? swar_kernel<SWAR_G3>(cn, last) : swar_kernel<SWAR_G5>(cn, last);
    case JK_AFFINE_BOX: return batch_affine_box;
    case JK_AFFINE_LIN4: return batch_other<JK_AFFINE_LIN4>;
    case JK_AFFINE_PX: return batch_other<JK_AFFINE_PX>;
    case JK_HALF_C1: return batch_other<JK_HALF_C1>;
    case JK_HALF: return batch_other<JK_HALF>;
    case JK_LINEAR: return batch_other<JK_LINEAR>;
    case JK_AREA_INT: return batch_other<JK_AREA_INT>;
    default: return batch_other<JK_AREA_FRAC>;
    }
}

struct Group {
    int kind, variant;
    int workgroups;
};

}  // namespace

}  // namespace visar

using namespace visar;

struct visar_batch {
    std::vector<BatchJob> jobs;
    std::vector<BatchUnit> units;
    std::vector<Group> groups;
    std::vector<int> seg;  // 9 bounds per group
    size_t tiles = 0;
    int workgroups = 0;
    void *dev = nullptr;  // jobs | units | segments | counters | tables
    size_t units_off = 0, seg_off = 0, cnt_off = 0;
    size_t dev_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

void drop_graph(visar_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

DImg dview(const visar_image &c) {
    DImg d;
    d.p = static_cast<uint8_t *>(c.data);
    d.w = (int)c.width;
    d.h = (int)c.height;
    d.cn = (int)c.channels;
    d.stride = (long)c.stride;
    return d;
}

struct Planned {
    BatchJob job;
    std::vector<uint8_t> table;  // resize tables, appended to the device arena
    int table_kind = 0;          // 1 linear, 2 area
    size_t nx = 0, ny = 0, nxs = 0;
};

int tiles256(const DImg &d, int rows) {
    return (int)(((long)d.w * d.cn + 255) / 256) * ((d.h + rows - 1) / rows);
}

// Validates one item and fills its job; returns a visar status.
int plan_item(const visar_batch_item &it, Planned &pl) {
    BatchJob &j = pl.job;
    std::memset(&j, 0, sizeof j);
    int rc;
    if ((rc = check_image(&it.src, 0)) || (rc = check_image(&it.dst, 0))) return rc;
    if (!is_device_ptr(it.src.data) || !is_device_ptr(it.dst.data)) return fail(UNSUPPORTED, "HostMemoryInBatch");
    if (it.src.channels != it.dst.channels || (it.src.channels != 1 && it.src.channels != 3))
        return fail(UNSUPPORTED, "UnsupportedChannels");
    const DImg s = dview(it.src), d = dview(it.dst);
    const bool same = it.src.width == it.dst.width && it.src.height == it.dst.height;
    switch (it.op) {
    case VISAR_BATCH_GAUSSIAN_BLUR: {
        if (it.border < 0 || it.border > 4) return fail(UNSUPPORTED, "Unsupported");
        if (!same) return fail(SIZE_MISMATCH, "SizeMismatch");
        const int k = it.ksize;
        if (k != 3 && k != 5 && k != 7) return fail(UNSUPPORTED, "Unsupported");
        // A one-pixel-wide (or -high) image maps every tap of that axis to
        // the same sample, which equals the reference's k = 1 there.
        uint32_t taps[kMaxTaps];
        gaussian_kernel_fixed(taps, k, it.sigma, 8);
        j.kind = JK_STENCIL;
        j.st.src = s;
        j.st.dst = d;
        j.st.border = it.border;
        j.st.fin = FIN_FIX8;
        j.st.scale = 1.0f;
        for (int i = 0; i < 7; i++) j.st.kx[i] = j.st.ky[i] = i < k ? (int)taps[i] : 0;
        int th = tuning().batch_stencil_rows ? tuning().batch_stencil_rows : 1;
        th = th >= 2 ? 2 : 1;
        j.variant = k * 100 + (int)it.src.channels * 10 + th;
        const int mode = swar_mode(j.st, k, 0);
        if (mode == SWAR_G3 || mode == SWAR_G5) {
            j.kind = JK_STENCIL_SWAR;
            j.variant = mode * 100 + (int)it.src.channels * 10 + th;
        }
        j.tiles_x = (int)(((long)s.w * s.cn + 255) / 256);
        j.ntiles = tiles256(s, 4 * th);
        return OK;
    }
    case VISAR_BATCH_MORPHOLOGY: {
        if (it.morph_op != 0 && it.morph_op != 1) return fail(UNSUPPORTED, "Unsupported");
        if (it.shape < 0 || it.shape > 2) return fail(UNSUPPORTED, "Unsupported");
        if (it.ksize <= 0 || it.ksize > 255 || it.iterations < 0) return fail(INVALID_ARGUMENT, "InvalidArgument");
        if (it.iterations != 1 || it.ksize > 7) return fail(UNSUPPORTED, "Unsupported");
        if (!same) return fail(SIZE_MISMATCH, "SizeMismatch");
        std::vector<uint8_t> mask;
        structuring_element(it.shape, it.ksize, it.ksize, mask);
        MorphParams mp;
        std::memset(&mp, 0, sizeof mp);
        mp.s = s;
        mp.d = d;
        mp.depth = 0;
        mp.kw = mp.kh = it.ksize;
        mp.ax = mp.ay = it.ksize / 2;
        for (int ky = 0; ky < it.ksize; ky++) {
            int x = 0;
            while (x < it.ksize) {
                if (!mask[(size_t)ky * it.ksize + x]) {
                    x++;
                    continue;
                }
                int x0 = x;
                while (x < it.ksize && mask[(size_t)ky * it.ksize + x]) x++;
                mp.runs[mp.nruns++] = MorphRun{(int16_t)ky, (int16_t)x0, (int16_t)(x - x0), 0};
            }
        }
        const int kf = morph_small_args(mp, j.mo);
        if (!kf) return fail(UNSUPPORTED, "Unsupported");
        j.kind = JK_MORPH;
        j.variant = kf * 100 + (int)it.src.channels * 10 + it.morph_op;
        j.tiles_x = (int)(((long)s.w * s.cn + 255) / 256);
        j.ntiles = tiles256(s, 8);
        return OK;
    }
    case VISAR_BATCH_WARP_AFFINE: {
        if (it.border < 0 || it.border > 4) return fail(UNSUPPORTED, "Unsupported");
        double inv[6];
        if (!invert_affine(it.m, inv)) return fail(UNSUPPORTED, "Unsupported");
        AffineParams &p = j.af;
        p.s = s;
        p.d = d;
        p.depth = 0;
        p.interp = it.interpolation == 0 ? 0 : 1;
        p.border = it.border;
        for (int i = 0; i < 6; i++) p.m[i] = (float)inv[i];
        p.cvd = sat_cast<uint8_t>(it.value);
        p.cv = (float)p.cvd;
        if (it.src.channels == 1 && p.interp == 1 && p.border == BORDER_CONSTANT) {
            j.kind = JK_AFFINE_BOX;
            j.tiles_x = (d.w + kWarpTW - 1) / kWarpTW;
            j.ntiles = j.tiles_x * ((d.h + kWarpTH - 1) / kWarpTH);
            return OK;
        }
        if (it.src.channels == 1 && p.interp == 1) {
            j.kind = JK_AFFINE_LIN4;
            j.tiles_x = (d.w + 255) / 256;
        } else {
            j.kind = JK_AFFINE_PX;
            j.tiles_x = (d.w + 63) / 64;
        }
        j.ntiles = j.tiles_x * ((d.h + 3) / 4);
        return OK;
    }
    case VISAR_BATCH_RESIZE: {
        if (it.interpolation != 1 && it.interpolation != 3) return fail(UNSUPPORTED, "Unsupported");
        if (same) return fail(UNSUPPORTED, "Unsupported");
        ResizeArgs &r = j.rs;
        r.s = s;
        r.d = d;
        const bool exact_half = it.src.width == 2 * it.dst.width && it.src.height == 2 * it.dst.height;
        const bool area_up = it.interpolation == 3 && (it.dst.width > it.src.width || it.dst.height > it.src.height);
        if (exact_half && !area_up) {
            if (d.cn == 1) {
                j.kind = JK_HALF_C1;
                j.tiles_x = (d.w + 1023) / 1024;
            } else {
                j.kind = JK_HALF;
                j.tiles_x = (d.w * d.cn + 255) / 256;
            }
            j.ntiles = j.tiles_x * ((d.h + 3) / 4);
            return OK;
        }
        j.tiles_x = (d.w * d.cn + 255) / 256;
        j.ntiles = j.tiles_x * ((d.h + 3) / 4);
        if (it.interpolation == 1 || area_up) {
            j.kind = JK_LINEAR;
            std::vector<LinTap> t(it.dst.width + it.dst.height);
            for (uint32_t x = 0; x < it.dst.width; x++) {
                LinTap q = linear_tap_u8(x, it.src.width, it.dst.width, area_up, true);
                q.o0 *= (int32_t)it.src.channels;
                q.o1 *= (int32_t)it.src.channels;
                t[x] = q;
            }
            for (uint32_t y = 0; y < it.dst.height; y++)
                t[it.dst.width + y] = linear_tap_u8(y, it.src.height, it.dst.height, area_up, false);
            pl.table.resize(t.size() * sizeof(LinTap));
            std::memcpy(pl.table.data(), t.data(), pl.table.size());
            pl.table_kind = 1;
            pl.nx = it.dst.width;
            return OK;
        }
        const uint32_t ix = it.src.width / it.dst.width, iy = it.src.height / it.dst.height;
        if (ix * it.dst.width == it.src.width && iy * it.dst.height == it.src.height) {
            j.kind = JK_AREA_INT;
            r.ix = (int)ix;
            r.iy = (int)iy;
            r.scale = 1.0f / (float)(ix * iy);
            return OK;
        }
        j.kind = JK_AREA_FRAC;
        std::vector<AreaEntry> xe, ye;
        std::vector<int32_t> xs, ys;
        area_tab(it.src.width, it.dst.width, (double)it.src.width / (double)it.dst.width, xe, xs);
        area_tab(it.src.height, it.dst.height, (double)it.src.height / (double)it.dst.height, ye, ys);
        size_t bytes = (xe.size() + ye.size()) * sizeof(AreaEntry) + (xs.size() + ys.size()) * sizeof(int32_t);
        pl.table.resize(bytes);
        uint8_t *o = pl.table.data();
        std::memcpy(o, xe.data(), xe.size() * sizeof(AreaEntry));
        o += xe.size() * sizeof(AreaEntry);
        std::memcpy(o, ye.data(), ye.size() * sizeof(AreaEntry));
        o += ye.size() * sizeof(AreaEntry);
        std::memcpy(o, xs.data(), xs.size() * sizeof(int32_t));
        o += xs.size() * sizeof(int32_t);
        std::memcpy(o, ys.data(), ys.size() * sizeof(int32_t));
        pl.table_kind = 2;
        pl.nx = xe.size();
        pl.ny = ye.size();
        pl.nxs = xs.size();
        return OK;
    }
    default: return fail(UNSUPPORTED, "Unsupported");
    }
}

}  // namespace

extern "C" {

int visar_batch_create(visar_batch **out) {
    if (!out) return fail(INVALID_ARGUMENT, "InvalidArgument");
    int rc = ensure_runtime();
    if (rc) return rc;
    *out = new (std::nothrow) visar_batch();
    return *out ? VISAR_OK : fail(OUT_OF_MEMORY, "OutOfMemory");
}

int visar_batch_plan(visar_batch *b, const visar_batch_item *items, size_t n, size_t *bad_item) {
    if (!b || (!items && n)) return fail(INVALID_ARGUMENT, "InvalidArgument");
    int rc = ensure_runtime();
    if (rc) return rc;
    std::vector<Planned> pl(n);
    for (size_t i = 0; i < n; i++) {
        if ((rc = plan_item(items[i], pl[i]))) {
            if (bad_item) *bad_item = i;
            return rc;
        }
    }
    // Group jobs by kind and variant (stable: the caller's order within a group).
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) {
        if (pl[a].job.kind != pl[c].job.kind) return pl[a].job.kind < pl[c].job.kind;
        return pl[a].job.variant < pl[c].job.variant;
    });
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    const int per_cu = tuning().batch_wg_per_cu ? tuning().batch_wg_per_cu : 4;
    const long wg_max = (long)cus * per_cu;
    // Work units: runs of consecutive tiles of one job, each about `target`
    // bytes of input + output, small enough that every workgroup gets many
    // (the tail balances) and large enough that a claim's atomic is rare.
    double total_bytes = 0;
    std::vector<double> job_bytes(n);
    for (size_t k = 0; k < n; k++) {
        const visar_batch_item &it = items[order[k]];
        job_bytes[k] = (double)it.src.width * it.src.height * it.src.channels +
                       (double)it.dst.width * it.dst.height * it.dst.channels;
        total_bytes += job_bytes[k];
    }
    const double target = std::min(256.0 * 1024, std::max(16.0 * 1024, total_bytes / ((double)wg_max * 16)));
    b->units.clear();
    std::vector<double> unit_bytes;
    long tiles = 0;
    for (size_t k = 0; k < n; k++) {
        const int nt = pl[order[k]].job.ntiles;
        const double tb = job_bytes[k] / std::max(1, nt);
        const int per = (int)std::max(1.0, std::min((double)nt, std::floor(target / std::max(tb, 1.0))));
        for (int f = 0; f < nt; f += per) {
            const int c = std::min(per, nt - f);
            b->units.push_back(BatchUnit{(int)k, f, c, 0});
            unit_bytes.push_back(tb * c);
        }
        tiles += nt;
    }
    b->tiles = (size_t)tiles;
    const size_t nu = b->units.size();
    // Groups: runs of jobs with one (kind, variant); each gets 8 segments of
    // equal bytes over its units, and a launch width that fits its units.
    b->groups.clear();
    b->seg.clear();
    b->workgroups = 0;
    {
        size_t u = 0;
        while (u < nu) {
            const BatchJob &j0 = pl[order[b->units[u].job]].job;
            size_t e = u;
            double gbytes = 0;
            while (e < nu) {
                const BatchJob &je = pl[order[b->units[e].job]].job;
                if (je.kind != j0.kind || je.variant != j0.variant) break;
                gbytes += unit_bytes[e];
                e++;
            }
            std::vector<int> sg(kXcds + 1, (int)e);
            sg[0] = (int)u;
            double acc = 0;
            int q = 1;
            for (size_t x = u; x < e && q < kXcds; x++) {
                acc += unit_bytes[x];
                while (q < kXcds && acc >= gbytes * q / kXcds) sg[q++] = (int)x + 1;
            }
            b->seg.insert(b->seg.end(), sg.begin(), sg.end());
            int wg = (int)std::max(1L, std::min<long>(wg_max, (long)(e - u)));
            if (wg > kXcds) wg -= wg % kXcds;  // every XCD's home segment is served
            b->groups.push_back(Group{j0.kind, j0.variant, wg});
            b->workgroups += wg;
            u = e;
        }
    }
    const size_t ng = b->groups.size();
    // Device arena: jobs | units | segment bounds | counters | tables.
    auto up256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t jobs_bytes = up256((n ? n : 1) * sizeof(BatchJob));
    const size_t units_off = jobs_bytes;
    const size_t seg_off = units_off + up256((nu ? nu : 1) * sizeof(BatchUnit));
    const size_t cnt_off = seg_off + up256((ng ? ng : 1) * (kXcds + 1) * sizeof(int));
    size_t total = cnt_off + up256((ng ? ng : 1) * kXcds * sizeof(int));
    std::vector<size_t> tab_at(n, 0);
    for (size_t i = 0; i < n; i++) {
        if (!pl[i].table.empty()) {
            tab_at[i] = total;
            total += up256(pl[i].table.size());
        }
    }
    drop_graph(b);
    if (b->dev_bytes < total) {
        if (b->dev) (void)hipFree(b->dev);
        b->dev = nullptr;
        b->dev_bytes = 0;
        if (hipMalloc(&b->dev, total) != hipSuccess) {
            (void)hipGetLastError();
            b->jobs.clear();
            b->units.clear();
            b->groups.clear();
            return fail(OUT_OF_MEMORY, "OutOfMemory");
        }
        b->dev_bytes = total;
    }
    b->units_off = units_off;
    b->seg_off = seg_off;
    b->cnt_off = cnt_off;
    uint8_t *base = static_cast<uint8_t *>(b->dev);
    std::vector<uint8_t> host(total, 0);
    b->jobs.clear();
    for (size_t k = 0; k < n; k++) {
        size_t i = order[k];
        BatchJob j = pl[i].job;
        if (pl[i].table_kind == 1) {
            const LinTap *t = reinterpret_cast<const LinTap *>(base + tab_at[i]);
            j.rs.xt = t;
            j.rs.yt = t + pl[i].nx;
        } else if (pl[i].table_kind == 2) {
            const uint8_t *p = base + tab_at[i];
            j.rs.xe = reinterpret_cast<const AreaEntry *>(p);
            j.rs.ye = j.rs.xe + pl[i].nx;
            j.rs.xs = reinterpret_cast<const int32_t *>(j.rs.ye + pl[i].ny);
            j.rs.ys = j.rs.xs + pl[i].nxs;
        }
        if (!pl[i].table.empty()) std::memcpy(host.data() + tab_at[i], pl[i].table.data(), pl[i].table.size());
        b->jobs.push_back(j);
    }
    if (n) std::memcpy(host.data(), b->jobs.data(), n * sizeof(BatchJob));
    if (nu) std::memcpy(host.data() + units_off, b->units.data(), nu * sizeof(BatchUnit));
    std::memcpy(host.data() + seg_off, b->seg.data(), b->seg.size() * sizeof(int));
    hipStream_t st = current_stream();
    if (hipMemcpyAsync(b->dev, host.data(), total, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(INTERNAL, "CopyFailed");
    return VISAR_OK;
}

static int launch_batch(visar_batch *b, hipStream_t st) {
    uint8_t *base = static_cast<uint8_t *>(b->dev);
    const int *seg = reinterpret_cast<const int *>(base + b->seg_off);
    int *counters = reinterpret_cast<int *>(base + b->cnt_off);
    const int ng = (int)b->groups.size();
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(256), 0, st, counters, ng * kXcds);
    const BatchJob *jobs = reinterpret_cast<const BatchJob *>(base);
    const BatchUnit *units = reinterpret_cast<const BatchUnit *>(base + b->units_off);
    for (int g = 0; g < ng; g++) {
        const Group &gr = b->groups[g];
        hipLaunchKernelGGL(group_kernel(gr.kind, gr.variant), dim3((unsigned)gr.workgroups), dim3(kTileThreads), 0, st,
                           jobs, units, seg + g * (kXcds + 1), counters + g * kXcds);
    }
    return hipGetLastError() == hipSuccess ? OK : INTERNAL;
}

int visar_batch_run(visar_batch *b, int use_graph) {
    if (!b) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (b->units.empty()) return VISAR_OK;
    hipStream_t st = current_stream();
    if (use_graph) {
        if (!b->exec || b->graph_stream != st) {
            drop_graph(b);
            hipStream_t cap = nullptr;
            if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) return fail(INTERNAL, "StreamFailed");
            bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
            ok = ok && launch_batch(b, cap) == OK;
            ok = hipStreamEndCapture(cap, &b->graph) == hipSuccess && ok;
            ok = ok && hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) == hipSuccess;
            (void)hipStreamDestroy(cap);
            if (!ok) {
                drop_graph(b);
                return fail(INTERNAL, "GraphCaptureFailed");
            }
            b->graph_stream = st;
        }
        if (hipGraphLaunch(b->exec, st) != hipSuccess) return fail(INTERNAL, "LaunchFailed");
    } else if (launch_batch(b, st) != OK) {
        return fail(INTERNAL, "LaunchFailed");
    }
    if (sync_mode() && hipStreamSynchronize(st) != hipSuccess) return fail(INTERNAL, "KernelFailed");
    return VISAR_OK;
}

int visar_batch_info(const visar_batch *b, size_t *tiles, size_t *workgroups) {
    if (!b) return fail(INVALID_ARGUMENT, "InvalidArgument");
    if (tiles) *tiles = b->tiles;
    if (workgroups) *workgroups = (size_t)b->workgroups;
    return VISAR_OK;
}

void visar_batch_free(visar_batch *b) {
    if (!b) return;
    drop_graph(b);
    if (b->dev) (void)hipFree(b->dev);
    delete b;
}

}  // extern "C"
