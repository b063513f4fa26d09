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
// job), and every unit is one workgroup of its group's launch, so the
// hardware dispatcher balances the load and no workgroup waits on a shared
// counter. Workgroup b is dispatched to XCD b % 8; it takes unit
// (b % 8) * per + b / 8 of its group, so each XCD works in order through one
// contiguous eighth of the group's units (neighbouring tiles of an image,
// and the halo rows they share, stay in that XCD's L2).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/visar_batch.h"
#include "runtime.hpp"
#include "tables.hpp"
#include "morph_chain.hpp"
#include "warp_blk.hpp"
#include "resize_stream.hpp"
#include "stencil_stream.hpp"

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
    JK_LINEAR_BOX = 11,
    JK_AREA_BOX = 12,
    JK_AFFINE_C3 = 13,
    JK_AFFINE_BLK = 14,
    JK_MORPH_CHAIN = 15,
    JK_LINEAR_STREAM = 16,
    JK_STENCIL_STREAM = 17,
    JK_MORPH_STREAM = 18,
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
    uint32_t lim; // JK_AFFINE_C3: the source's tap_limit; JK_AFFINE_BLK: its blk_limit
    int tiles_x;
    int ntiles;
    union {
        StencilArgs st;
        MorphSmallArgs mo;
        AffineParams af;
        ResizeArgs rs;
        MorphChainArgs mc;
        LinStreamArgs ls;
        SStreamArgs ss;
    };
};

// A run of consecutive tiles [first, first + count) of job `job`.
struct BatchUnit {
    int job, first, count, pad;
};

// Runs the unit of this workgroup: `count` units start at `units`.
template <typename Tile>
__device__ __attribute__((always_inline)) inline void run_unit(const BatchJob *__restrict__ jobs, const BatchUnit *__restrict__ units,
                                                               int count, Tile tile) {
    const int per = (count + kXcds - 1) / kXcds;
    const int i = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if (i >= count) return;
    const BatchUnit u = units[i];
    const BatchJob &j = jobs[u.job];
    const int tiles_x = j.tiles_x;
    int by = u.first / tiles_x, bx = u.first - by * tiles_x;
    for (int c = 0; c < u.count; c++) {
        if (c) __syncthreads();  // the previous tile's LDS reads are done
        tile(j, bx, by);
        if (++bx == tiles_x) {
            bx = 0;
            by++;
        }
    }
}

#define BATCH_ARGS const BatchJob *__restrict__ jobs, const BatchUnit *__restrict__ units, int count

template <int K, int CN, int TH>
__global__ __launch_bounds__(kTileThreads) void batch_stencil(BATCH_ARGS) {
    __shared__ uint32_t lds[StencilShape<K, CN, TH>::LDS_WORDS];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { stencil_tile<K, CN, TH, uint8_t>(j.st, bx, by, lds); });
}

template <int MODE, int CN, int TH>
__global__ __launch_bounds__(kTileThreads) void batch_stencil_swar(BATCH_ARGS) {
    __shared__ uint32_t lds[StencilShape<SwarTaps<MODE>::K, CN, TH>::LDS_WORDS];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { stencil_swar_tile<MODE, CN, TH, uint8_t>(j.st, bx, by, lds); });
}

__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(5))) void batch_affine_box(BATCH_ARGS) {
    __shared__ uint32_t lds[kWarpLdsWords];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { affine_box_tile(j.af, bx, by, lds); });
}

__global__ __launch_bounds__(kTileThreads) void batch_affine_blk(BATCH_ARGS) {
    __shared__ uint32_t lds[4 * 64];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { affine_blk_tile(j.af, j.lim, bx, by, lds); });
}

using GroupKernel = void (*)(BATCH_ARGS);

// One erode / dilate pass of the fused chain kernel (morph_chain.hpp) on
// tiles of kBatchChainRows rows.
constexpr int kBatchChainRows = 16;
template <int CN, int K, uint32_t EL>
__global__ __launch_bounds__(kTileThreads) void batch_morph_chain(BATCH_ARGS) {
    __shared__ __align__(16) uint32_t lds[(kBatchChainRows + 2 * K) * mchain::ChainGeo<CN, K>::PITCH];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { mchain::chain_tile<CN, K, EL>(j.mc, bx, by, lds); });
}

template <int CN>
GroupKernel chain_kernel(uint32_t el) {
    using namespace mchain;
    switch (el) {
    case kElRect3: return batch_morph_chain<CN, 1, kElRect3>;
    case kElRect5: return batch_morph_chain<CN, 2, kElRect5>;
    case kElRect7: return batch_morph_chain<CN, 3, kElRect7>;
    case kElCross3: return batch_morph_chain<CN, 1, kElCross3>;
    case kElCross5: return batch_morph_chain<CN, 2, kElCross5>;
    case kElCross7: return batch_morph_chain<CN, 3, kElCross7>;
    case kElEllipse5: return batch_morph_chain<CN, 2, kElEllipse5>;
    default: return batch_morph_chain<CN, 3, kElEllipse7>;
    }
}

// Whether element code el has a chain kernel in the batch.
bool batch_chain_el(uint32_t el) {
    using namespace mchain;
    return el == kElRect3 || el == kElRect5 || el == kElRect7 || el == kElCross3 || el == kElCross5 || el == kElCross7 ||
           el == kElEllipse5 || el == kElEllipse7;
}

template <int K, int CN, bool MAX>
__global__ __launch_bounds__(kTileThreads) void batch_morph(BATCH_ARGS) {
    __shared__ uint32_t lds[MorphShape<K, CN, 2>::LDS_WORDS];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) { morph_small_tile<K, CN, 2, MAX>(j.mo, bx, by, lds); });
}

__global__ __launch_bounds__(kTileThreads) void batch_linear(BATCH_ARGS) {
    __shared__ uint32_t lds[kLinLdsWords];
    run_unit(jobs, units, count,
             [&](const BatchJob &j, int bx, int by) { linear_box_tile(j.rs.s, j.rs.d, j.rs.xt, j.rs.yt, bx, by, lds); });
}

// GaussianBlur 5x5 and 7x7 through the general streamed waves
// (stencil_stream.hpp): tile (strip, group of 4 bands of 4 rows), one band
// per wave, destinations of any alignment in one group (3x3 stays on the
// LDS tiles, measured faster in batches).
template <int MODE, int CN>
__global__ __launch_bounds__(kTileThreads) void batch_stencil_stream(BATCH_ARGS) {
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) {
        sstream::stream_wave<sstream::GaussOp<MODE>, CN, 4, false>(j.ss, bx, 4 * by + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
    });
}
template <int MODE>
GroupKernel stencil_stream_kernel(int cn) {
    return cn == 1 ? batch_stencil_stream<MODE, 1> : batch_stencil_stream<MODE, 3>;
}

// erode / dilate with a K x K rectangle through the same streamed waves
// (tile as above).
template <int K, bool MAX, int CN>
__global__ __launch_bounds__(kTileThreads) void batch_morph_stream(BATCH_ARGS) {
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) {
        sstream::stream_wave<sstream::RectOp<K, MAX>, CN, 4, false>(j.ss, bx, 4 * by + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6));
    });
}
template <int K>
GroupKernel morph_stream_kernel(bool mx, int cn) {
    if (cn == 1) return mx ? batch_morph_stream<K, true, 1> : batch_morph_stream<K, false, 1>;
    return mx ? batch_morph_stream<K, true, 3> : batch_morph_stream<K, false, 3>;
}

// LINEAR through the register-streamed waves (resize_stream.hpp): tile
// (strip, 4 runs of j.ls.run bands of 4 rows), one run per wave. One kernel
// for every alignment (one group measured faster than a group per aligned
// case: 1.570 -> 1.543 ms for the 512-image batch).
__global__ __launch_bounds__(kTileThreads) void batch_linear_stream(BATCH_ARGS) {
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) {
        const int w = 4 * by + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
        lstream::lin_stream_wave<4, false, false, true>(j.ls, bx, w * j.ls.run, (w + 1) * j.ls.run);
    });
}

__global__ __launch_bounds__(kTileThreads) void batch_area(BATCH_ARGS) {
    __shared__ uint32_t lds[kAreaLdsWords];
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) {
        area_box_tile(j.rs.s, j.rs.d, j.rs.xe, j.rs.xs, j.rs.ye, j.rs.ys, bx, by, lds);
    });
}

template <int KIND>
__global__ __launch_bounds__(kTileThreads) void batch_other(BATCH_ARGS) {
    run_unit(jobs, units, count, [&](const BatchJob &j, int bx, int by) {
        if constexpr (KIND == JK_AFFINE_LIN4) affine_lin4_tile(j.af, bx, by);
        else if constexpr (KIND == JK_AFFINE_C3) affine_c3_u8_tile(j.af, j.lim, bx, by);
        else if constexpr (KIND == JK_AFFINE_PX) affine_px_tile<uint8_t>(j.af, bx, by);
        else if constexpr (KIND == JK_HALF_C1) half_c1_tile(j.rs.s, j.rs.d, bx, by);
        else if constexpr (KIND == JK_HALF) half_tile(j.rs.s, j.rs.d, bx, by);
        else if constexpr (KIND == JK_LINEAR) linear_tile(j.rs.s, j.rs.d, j.rs.xt, j.rs.yt, bx, by);
        else if constexpr (KIND == JK_AREA_INT) area_int_tile(j.rs.s, j.rs.d, j.rs.ix, j.rs.iy, j.rs.scale, bx, by);
        else area_frac_tile(j.rs.s, j.rs.d, j.rs.xe, j.rs.xs, j.rs.ye, j.rs.ys, bx, by);
    });
}


template <int K, int CN>
GroupKernel stencil_kernel(int th) {
    return th == 4 ? batch_stencil<K, CN, 4> : th == 2 ? batch_stencil<K, CN, 2> : batch_stencil<K, CN, 1>;
}

template <int MODE>
GroupKernel swar_kernel(int cn, int th) {
    if (cn == 1)
        return th == 4 ? batch_stencil_swar<MODE, 1, 4> : th == 2 ? batch_stencil_swar<MODE, 1, 2> : batch_stencil_swar<MODE, 1, 1>;
    return th == 4 ? batch_stencil_swar<MODE, 3, 4> : th == 2 ? batch_stencil_swar<MODE, 3, 2> : batch_stencil_swar<MODE, 3, 1>;
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
    case JK_STENCIL_SWAR:
        return k == SWAR_G3 ? swar_kernel<SWAR_G3>(cn, last) : k == SWAR_G5 ? swar_kernel<SWAR_G5>(cn, last) : swar_kernel<SWAR_G7>(cn, last);
    case JK_AFFINE_BOX: return batch_affine_box;
    case JK_AFFINE_BLK: return batch_affine_blk;
    case JK_MORPH_CHAIN: return variant % 10 == 1 ? chain_kernel<1>((uint32_t)(variant / 10)) : chain_kernel<3>((uint32_t)(variant / 10));
    case JK_AFFINE_LIN4: return batch_other<JK_AFFINE_LIN4>;
    case JK_AFFINE_PX: return batch_other<JK_AFFINE_PX>;
    case JK_AFFINE_C3: return batch_other<JK_AFFINE_C3>;
    case JK_HALF_C1: return batch_other<JK_HALF_C1>;
    case JK_HALF: return batch_other<JK_HALF>;
    case JK_LINEAR: return batch_other<JK_LINEAR>;
    case JK_LINEAR_BOX: return batch_linear;
    case JK_MORPH_STREAM: {
        const int k = variant / 100, cn = variant % 10;
        const bool mx = (variant / 10) % 10 != 0;
        return k == 3 ? morph_stream_kernel<3>(mx, cn) : k == 5 ? morph_stream_kernel<5>(mx, cn) : morph_stream_kernel<7>(mx, cn);
    }
    case JK_STENCIL_STREAM:
        return variant / 100 == SWAR_G5 ? stencil_stream_kernel<SWAR_G5>((variant / 10) % 10)
                                        : stencil_stream_kernel<SWAR_G7>((variant / 10) % 10);
    case JK_LINEAR_STREAM: return batch_linear_stream;
    case JK_AREA_BOX: return batch_area;
    case JK_AREA_INT: return batch_other<JK_AREA_INT>;
    default: return batch_other<JK_AREA_FRAC>;
    }
}

struct Group {
    int kind, variant;
    int first, count;  // the group's units
};

int group_workgroups(const Group &g) { return (g.count + kXcds - 1) / kXcds * kXcds; }

}  // namespace

}  // namespace visar

using namespace visar;

struct visar_batch {
    std::vector<BatchJob> jobs;
    std::vector<BatchUnit> units;
    std::vector<Group> groups;
    size_t tiles = 0;
    int workgroups = 0;
    void *dev = nullptr;  // jobs | units | tables
    size_t units_off = 0;
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
        th = th >= 4 ? 4 : th >= 2 ? 2 : 1;
        j.variant = k * 100 + (int)it.src.channels * 10 + th;
        const int mode = swar_mode(j.st, k, 0);
        SStreamArgs ss;
        if ((mode == SWAR_G5 || mode == SWAR_G7) && tuning().stencil_direct != 1 && stencil_stream_args(j.st, 4, ss)) {
            j.kind = JK_STENCIL_STREAM;
            j.variant = mode * 100 + (int)it.src.channels * 10;
            j.ss = ss;
            j.tiles_x = ss.strips;
            j.ntiles = ss.strips * ((ss.bands + 3) / 4);
            return OK;
        }
        if (mode == SWAR_G3 || mode == SWAR_G5 || mode == SWAR_G7) {
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
        MorphRects mr;
        const bool rects = morph_rects(mp, mr);
        const uint32_t el = rects ? morph_el_code(mr) : 0u;
        SStreamArgs ss;
        StencilArgs sa;
        std::memset(&sa, 0, sizeof sa);
        sa.src = s;
        sa.dst = d;
        sa.border = BORDER_REPLICATE;
        if (tuning().morph_chain == 0 && (el == mchain::kElRect3 || el == mchain::kElRect5 || el == mchain::kElRect7) &&
            stencil_stream_args(sa, 4, ss)) {
            // A full rectangle: the streamed waves of stencil_stream.hpp.
            j.kind = JK_MORPH_STREAM;
            j.variant = it.ksize * 100 + it.morph_op * 10 + (int)it.src.channels;
            j.ss = ss;
            j.tiles_x = ss.strips;
            j.ntiles = ss.strips * ((ss.bands + 3) / 4);
            return OK;
        }
        if (tuning().morph_chain != 1 && rects && batch_chain_el(el)) {
            // One pass of the fused chain kernel, as the plain call's
            // morph_chain_launch sets it up for one chain of one pass.
            MorphChainArgs &a = j.mc;
            std::memset(&a, 0, sizeof a);
            const int cn = (int)it.src.channels;
            const int K = std::max(1, std::max(mr.eh, mr.ev));
            a.s = s;
            a.d = d;
            a.c = s;
            a.ev = mr.ev;
            a.vb = mr.ev;
            a.hb = (mr.eh * cn + 7) & ~7;
            a.tw = 512 - 2 * a.hb;
            a.nr = mr.nr;
            for (int k = 0; k < mr.nr; k++) {
                a.rw[k] = mr.rw[k];
                a.rv[k] = mr.rv[k];
                for (int e = -mr.rv[k]; e <= mr.rv[k]; e++) a.cmask[mr.rw[k]] |= 1u << (e + K);
            }
            a.pass[0] = MorphChainPass{0, -1, (int8_t)it.morph_op, 1};
            a.npass = 1;
            a.nbuf = 1;
            a.th = kBatchChainRows;
            a.rows = a.th + 2 * a.vb;
            a.tiles_x = (s.w * cn + a.tw - 1) / a.tw;
            a.ntiles = a.tiles_x * ((s.h + a.th - 1) / a.th);
            j.kind = JK_MORPH_CHAIN;
            j.variant = (int)morph_el_code(mr) * 10 + cn;
            j.tiles_x = a.tiles_x;
            j.ntiles = a.ntiles;
            return OK;
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
        if (it.src.channels == 1 && p.interp == 1 && p.border == BORDER_CONSTANT && s.w >= 2 && s.h >= 2 &&
            blk_limit(s) && affine_bounded(p) && tuning().warp_layout != 1) {
            j.kind = JK_AFFINE_BLK;
            j.lim = blk_limit(s);
            j.tiles_x = (d.w + 63) / 64;
            j.ntiles = j.tiles_x * ((d.h + 15) / 16);
            return OK;
        }
        if (it.src.channels == 1 && p.interp == 1 && p.border == BORDER_CONSTANT) {
            j.kind = JK_AFFINE_BOX;
            j.tiles_x = (d.w + kWarpTW - 1) / kWarpTW;
            j.ntiles = j.tiles_x * ((d.h + kWarpTH - 1) / kWarpTH);
            return OK;
        }
        if (it.src.channels == 1 && p.interp == 1) {
            j.kind = JK_AFFINE_LIN4;
            j.tiles_x = (d.w + 255) / 256;
        } else if (it.src.channels == 3 && p.interp == 1 && p.border == BORDER_CONSTANT && tap_limit(s)) {
            j.kind = JK_AFFINE_C3;
            j.lim = tap_limit(s);
            j.tiles_x = (d.w + 63) / 64;
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
            LinStreamArgs ls;
            // Bands of 4 rows in runs of 4 (tuning "lin_run" overrides).
            const int run = tuning().lin_run ? tuning().lin_run : 4;
            if (linear_stream_fits(t.data(), (int)it.dst.width, (int)it.src.channels, t.data() + it.dst.width, (int)it.dst.height) &&
                linear_stream_args(s, d, nullptr, nullptr, 4, run, ls)) {
                j.kind = JK_LINEAR_STREAM;
                j.ls = ls;
                j.tiles_x = ls.strips;
                j.ntiles = ls.strips * ((ls.bands + 4 * ls.run - 1) / (4 * ls.run));
            } else if (linear_box_fits(t.data(), (int)it.dst.width, (int)it.src.channels, t.data() + it.dst.width,
                                       (int)it.dst.height)) {
                j.kind = JK_LINEAR_BOX;
                j.ntiles = j.tiles_x * ((d.h + kLinTH - 1) / kLinTH);
            }
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
        if (area_box_fits(xe.data(), xs.data(), (int)it.dst.width, (int)it.src.channels, ye.data(), ys.data(), (int)it.dst.height)) {
            j.kind = JK_AREA_BOX;
            j.ntiles = j.tiles_x * ((d.h + kAreaTH - 1) / kAreaTH);
        }
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
    // Work units: runs of consecutive tiles of one job, each about `target`
    // bytes of input + output: large enough to spread a workgroup's start-up
    // over several tiles, small enough that the last units are short.
    const double target = 1024.0 * (tuning().batch_unit_kb ? tuning().batch_unit_kb : 16);
    b->units.clear();
    b->groups.clear();
    long tiles = 0;
    for (size_t k = 0; k < n; k++) {
        const visar_batch_item &it = items[order[k]];
        const BatchJob &j = pl[order[k]].job;
        if (b->groups.empty() || b->groups.back().kind != j.kind || b->groups.back().variant != j.variant)
            b->groups.push_back(Group{j.kind, j.variant, (int)b->units.size(), 0});
        const double bytes = (double)it.src.width * it.src.height * it.src.channels +
                             (double)it.dst.width * it.dst.height * it.dst.channels;
        const int nt = j.ntiles;
        const double tb = bytes / std::max(1, nt);
        const int per = (int)std::max(1.0, std::min((double)nt, std::floor(target / std::max(tb, 1.0))));
        for (int f = 0; f < nt; f += per) b->units.push_back(BatchUnit{(int)k, f, std::min(per, nt - f), 0});
        b->groups.back().count = (int)b->units.size() - b->groups.back().first;
        tiles += nt;
    }
    b->tiles = (size_t)tiles;
    const size_t nu = b->units.size();
    b->workgroups = 0;
    for (const Group &g : b->groups) b->workgroups += group_workgroups(g);
    // Device arena: jobs | units | tables.
    auto up256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t jobs_bytes = up256((n ? n : 1) * sizeof(BatchJob));
    const size_t units_off = jobs_bytes;
    size_t total = units_off + up256((nu ? nu : 1) * sizeof(BatchUnit));
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
    uint8_t *base = static_cast<uint8_t *>(b->dev);
    std::vector<uint8_t> host(total, 0);
    b->jobs.clear();
    for (size_t k = 0; k < n; k++) {
        size_t i = order[k];
        BatchJob j = pl[i].job;
        if (pl[i].table_kind == 1 && j.kind == JK_LINEAR_STREAM) {
            const LinTap *t = reinterpret_cast<const LinTap *>(base + tab_at[i]);
            j.ls.xt = t;
            j.ls.yt = t + pl[i].nx;
        } else if (pl[i].table_kind == 1) {
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
    hipStream_t st = current_stream();
    if (hipMemcpyAsync(b->dev, host.data(), total, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(INTERNAL, "CopyFailed");
    return VISAR_OK;
}

static int launch_batch(visar_batch *b, hipStream_t st) {
    uint8_t *base = static_cast<uint8_t *>(b->dev);
    const BatchJob *jobs = reinterpret_cast<const BatchJob *>(base);
    const BatchUnit *units = reinterpret_cast<const BatchUnit *>(base + b->units_off);
    for (const Group &g : b->groups)
        hipLaunchKernelGGL(group_kernel(g.kind, g.variant), dim3((unsigned)group_workgroups(g)), dim3(kTileThreads), 0, st,
                           jobs, units + g.first, g.count);
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
