// kernels.hpp -- launch interface between the C-ABI layer and the HIP kernels.
#pragma once

#include "common.hpp"

namespace visar {

// Launch-shape knobs (see visar_gpu_set_tuning); 0 means automatic.
struct Tuning {
    int stencil_rows = 0;        // plain stencils: output rows per thread (1, 2, 4)
    int batch_stencil_rows = 0;  // batched stencils: output rows per thread (1, 2, 4)
    int batch_unit_kb = 0;       // batched calls: KiB of input + output per workgroup
    int tile_order = 0;          // plain tile kernels: 0 each kernel's default, 1 row-major, 2 one band of tiles per XCD
    int sep_rows = 0;            // separable engine: output rows per tile (1-16)
    int warp_layout = 0;         // u8 bilinear constant-border warps: 0 block layout, loads per pixel; 1 row / LDS-box kernels; 2 block layout, loads after a wave vote
    int morph_chain = 0;         // u8 morphology: 0 automatic, 1 one kernel per pass and subtraction, 2 fused LDS chains
    int morph_rows = 0;          // fused morphology: output rows per tile (8-128)
    int stencil_direct = 0;      // one-channel packed stencils: 0 rows streamed through registers where the layout allows, 1 LDS tiles
    int stencil_band = 0;        // streamed stencils: output rows per wave (2, 4, 8, 16)
    int stencil_wg = 0;          // streamed stencils: waves per workgroup (1, 2, 4)
    int warp_blocks = 0;         // eager u8 block warps: 64 x 4 pixel block rows per wave (1, 2; 0 = 1)
    int morph_stream = 0;        // one-channel u8 morphology with iterations or gradient: 0 streamed n-fold passes, 1 pass by pass
    int morph_band = 0;          // streamed n-fold passes: output rows per band (0 = 16)
    int morph_pass = 0;          // one-channel u8 single passes: 0 column window first (morph_band.hpp), 1 morph_direct.hpp
    int morph_pass_rows = 0;     // column-window-first passes: output rows per wave (4, 8, 16; 0 = 4)
    int morph_pow_rows = 0;      // n-fold column-window passes: output rows per wave (4, 8; 0 = 8)
    int gray_layout = 0;         // u8 cvtColor to gray on contiguous images: 0 one contiguous 1024-pixel block per wave, 1 16 pixels per thread
    int lin_stream = 0;          // u8 LINEAR resize: 0 rows streamed through registers where the tables allow, 1 LDS tiles
    int lin_rows = 0;            // streamed LINEAR resize: output rows per band (4, 8; 0 automatic)
    int lin_run = 0;             // streamed LINEAR resize: bands per wave (1-16; 0 automatic)
};
Tuning &tuning();
bool set_tuning(const char *key, int value);
bool get_tuning(const char *key, int *value);

// ---------------------------------------------------------------- filters

// Finish kinds of reference src/imgproc/filter.zig:61-116.
enum Finish : int { FIN_FIX8 = 0, FIN_FIX16 = 1, FIN_CAST = 2, FIN_SCALED = 3 };
enum Sym : int { SYM_NONE = 0, SYM_EVEN = 1, SYM_ODD = 2 };

constexpr int kMaxTaps = 128;

// adaptiveThreshold's last step (threshold.zig:340-360) folded into the
// kernel that computes the local mean m of a u8 image s: the destination
// gets on when s + idelta > m, off otherwise (idelta clamped to +-1024), in
// place of m. Off (enabled 0) for every other call.
struct AdaptPost {
    int enabled, idelta, on, off;
};
__host__ __device__ inline uint8_t adapt_out(int idelta, int on, int off, int s, uint8_t m) {
    return (uint8_t)(s + idelta > (int)m ? on : off);
}

// One separable filter call (filter.zig:120-205). Integer modes use the
// integer taps; the f32 mode uses the float taps and reproduces the
// reference's order of operations: lanes of a row before `body` fold
// symmetric taps, the scalar tail sums them in order.
struct SepParams {
    DImg src, dst;
    int sdepth, ddepth;
    int nkx, nky;
    int border;
    int fin;
    float scale;
    int f32mode;
    int vsym, hsym, const_kernel;
    long body;
    int32_t kxi[kMaxTaps], kyi[kMaxTaps];
    float kxf[kMaxTaps], kyf[kMaxTaps];
    AdaptPost adapt;  // u8 -> u8 only
};

int launch_sep(const SepParams &p, hipStream_t st);

// Register-window fast path for u8 sources with square 3/5/7 kernels
// (k_stencil.hip); returns false for shapes it does not cover. `th` is the
// number of output rows per thread (1, 2 or 4).
bool launch_sep_small_u8(const SepParams &p, hipStream_t st, int th);
int stencil_rows_per_thread(long n, int h);
struct StencilArgs;
bool stencil_args(const SepParams &p, StencilArgs &a);

// -------------------------------------------------------------- pointwise

// color.zig:324-403 gray conversion: scn 3/4, sb 0/2 (blue index).
int launch_gray(const DImg &s, const DImg &d, int depth, int scn, int sb, hipStream_t st);

struct ThreshParams {
    DImg s, d;
    int depth, kind;
    int below;  // 16-bit thresholds below the range: every value is above
    int ti, mi; // u8: floor(thresh) clamped to the int range, maxval as u8
    uint8_t lut[256];
    double t, m;  // threshold / maxval already converted to the sample type
};
int launch_threshold(const ThreshParams &p, hipStream_t st);
// Host side of threshold (capi.cpp): the automatic thresholds of a 256-bin
// histogram (threshold.zig:184-253) and the kernel's parameters for a
// threshold and maxval of depth T (threshold.zig:25-82).
double threshold_otsu(const uint32_t *hist, double total);
double threshold_triangle(const uint32_t *hist);
void threshold_setup(ThreshParams &p, uint32_t T, int kind, double thresh, double maxval);

// video/track.zig:33-81 (primed update) and the priming pass.
int launch_background(const DImg &frame, const DImg &mask, uint16_t *model, long model_stride, int shift, int thr,
                      int prime, hipStream_t st);

// ----------------------------------------------------------------- resize

int launch_copy(const DImg &s, const DImg &d, size_t elem, hipStream_t st);
int launch_half_u8(const DImg &s, const DImg &d, hipStream_t st);
// BGR / BGRA u8 -> gray shrunk by k with box averaging (videoio/frame.zig:167-177).
int launch_gray_down(const DImg &s, const DImg &d, int k, hipStream_t st);

// imgproc/pyramid.zig pyrDown (`down`, any border but constant) and pyrUp
// (reflect-101): u8, u16, i16 and f32 (depth codes 0, 2, 3, 5).
int launch_pyramid(const DImg &s, const DImg &d, int depth, int border, bool down, hipStream_t st);

// filter.zig:617-806 boxFilter on u8, 1-4 channels, odd windows up to 127:
// exact sums, sat(f32(sum) * scale); UNSUPPORTED outside that.
int launch_box_u8(const DImg &s, const DImg &d, int kw, int kh, int border, float scale, hipStream_t st,
                  const AdaptPost &adapt = AdaptPost{});
// The same for windows with a side above 127 (up to 255) and for i16 (and
// u8 / u16 / f32) samples: two passes through `tmp`, a device buffer of
// box_wide_tmp_bytes(s) bytes (column sums). `adapt` applies to u8 only.
size_t box_wide_tmp_bytes(const DImg &s);
int launch_box_wide(const DImg &s, const DImg &d, int depth, int kw, int kh, int border, float scale, void *tmp,
                    hipStream_t st, const AdaptPost &adapt = AdaptPost{});

int launch_area_int_u8(const DImg &s, const DImg &d, int ix, int iy, hipStream_t st);
int launch_nearest(const DImg &s, const DImg &d, size_t elem, hipStream_t st);

// resize.zig:186-214 taps, prepared on the host: per destination column the
// two source sample offsets (channel 0, in samples) and 11-bit weights; per
// destination row the two source rows and weights.
struct LinTap {
    int32_t o0, o1;
    int32_t w0, w1;
};
int launch_linear_u8(const DImg &s, const DImg &d, const LinTap *xt_dev, const LinTap *yt_dev, hipStream_t st, bool box,
                     bool stream = false);

// Tiles of the LDS-staged linear resize: 256 output samples x kLinTH rows,
// whose source box is at most kLinRows rows of kLinWords dwords.
constexpr int kLinTH = 16;
constexpr int kLinRows = 18;    // every upscale fits
constexpr int kLinWords = 132;  // scales down to 1/2 fit
constexpr int kLinLdsWords = kLinRows * kLinWords + kLinRows * 128;
// True when the tables (x taps in samples, premultiplied by cn) suit the
// LDS tiles: taps that never decrease, o0 <= o1, and every tile's box within
// the limits above.
bool linear_box_fits(const LinTap *xt, int dw, int cn, const LinTap *yt, int dh);

// The register-streamed linear resize (resize_stream.hpp): strips of
// kLinStreamStrip output samples, bands of tb (4 or 8) output rows, each
// wave on a run of `run` bands of one strip.
constexpr int kLinStreamStrip = 252;
struct LinStreamArgs {
    DImg s, d;  // p of each rounded down to a dword; the bytes skipped in off_s / off_d
    const LinTap *xt, *yt;
    uint32_t off_s, off_d;
    uint32_t lim_s, lim_d;    // bytes from each rounded-down base to the image's last byte + 1 (< 2^31), lim_s rounded up to 4
    int strips, ntiles, xcd;  // the tile order: tile (strip, 4 runs of bands)
    int tb, bands, run;
};
// True where the streamed kernel applies to these tables (x taps in
// samples, premultiplied by cn): weights in [0, 2050] summing to at most
// 2050, y taps that never decrease with o1 - o0 of 0 or 1, and every four
// consecutive samples' taps within 8 bytes.
bool linear_stream_fits(const LinTap *xt, int dw, int cn, const LinTap *yt, int dh);
// The streamed kernel's arguments with bands of tb rows and runs of `run`
// bands (all but the tile order; tb 0 picks from the image size, run 0 is
// 1); false where its 32-bit offsets or 12-byte windows do not fit.
bool linear_stream_args(const DImg &s, const DImg &d, const LinTap *xt, const LinTap *yt, int tb, int run, LinStreamArgs &a);

// resize.zig:470-524 area tables: entries grouped per destination index,
// start offsets in `xstart` / `ystart` (length dsize + 1).
struct AreaEntry {
    int32_t si;
    float alpha;
};
int launch_area_frac_u8(const DImg &s, const DImg &d, const AreaEntry *xt, const int32_t *xstart, const AreaEntry *yt,
                        const int32_t *ystart, hipStream_t st, bool box);

// The generic resize kernels (k_resize.hip; resize.zig:47-60): LINEAR and
// AREA for u16 / i16 / f32 and for u8 with more than 4 channels. FLinTap:
// linearCoord's taps as f32 weights; x taps in samples (premultiplied by
// cn), the second tap equal to the first at the last column; y taps are
// rows. u8 takes LinTap tables (linearTapU8).
struct FLinTap {
    int32_t i0, i1;
    float w0, w1;
};
int launch_linear_any(const DImg &s, const DImg &d, int depth, const void *xt_dev, const void *yt_dev, hipStream_t st);
int launch_area_int_any(const DImg &s, const DImg &d, int depth, int ix, int iy, hipStream_t st);
int launch_area_gen_any(const DImg &s, const DImg &d, int depth, const AreaEntry *xe, const int32_t *xs,
                        const AreaEntry *ye, const int32_t *ys, hipStream_t st);

// Tiles of the LDS-staged area resize: 256 output samples x kAreaTH rows,
// cells of at most kAreaMaxCell entries, source boxes of at most kAreaRows
// rows of kAreaWords dwords (shrinks by up to 2.2 fit; 8 workgroups' LDS
// fit one CU).
constexpr int kAreaTH = 4;
constexpr int kAreaMaxCell = 4;
constexpr int kAreaRows = 12;
constexpr int kAreaWords = 148;
constexpr int kAreaLdsWords = kAreaRows * kAreaWords + kAreaRows * 256;
// True when the area tables suit those tiles.
bool area_box_fits(const AreaEntry *xt, const int32_t *xs, int dw, int cn, const AreaEntry *yt, const int32_t *ys, int dh);

// ------------------------------------------------------------------ warps

struct AffineParams {
    DImg s, d;
    int depth, interp, border;
    float m[6];  // inverse matrix rounded to f32
    float cv;    // border value as f32 (toF of the stored sample)
    double cvd;  // border value as stored sample
};
int launch_warp_affine(const AffineParams &p, hipStream_t st);
// Bytes from the dword at or below an image's base to the end of its last
// row (the range of the warps' buffer loads), or 0 when that does not fit
// in 32 bits with room for a 16-byte load.
inline uint32_t tap_limit(const DImg &s) {
    const uint64_t lim = ((uintptr_t)s.p & 3) + (uint64_t)(s.h - 1) * (uint64_t)s.stride + (uint64_t)s.w * s.cn;
    return lim >= 16 && lim + 16 < (1ull << 31) ? (uint32_t)lim : 0;
}

struct PerspParams {
    DImg s, d;
    int depth, interp, border;
    double m[9];  // inverse matrix (f64)
    double cvd;
};
int launch_warp_perspective(const PerspParams &p, hipStream_t st);

int launch_remap(const DImg &s, const DImg &d, const DImg &mx, const DImg &my, int bv, hipStream_t st);

// ------------------------------------------------------------- morphology

// One horizontal run of "on" cells of the structuring element: element row
// ky, columns [x0, x0 + len).
struct MorphRun {
    int16_t ky, x0, len, pad;
};
constexpr int kMaxRuns = 256;  // rect, cross and ellipse have at most one run per row
struct MorphParams {
    DImg s, d;
    int depth;
    int op_max;
    int kw, kh, ax, ay;
    int nruns;
    MorphRun runs[kMaxRuns];
};
int launch_morph_pass(const MorphParams &p, hipStream_t st);

// The element as a union of centred rectangles (k_morph_chain.hip), or false:
// odd sides up to 7, the anchor at the centre, at most one run per row,
// centred on the anchor, with half-widths symmetric about the centre row
// and non-increasing away from it. Rectangle k: half-width rw[k], half-height
// rv[k]; eh / ev: the largest of each.
struct MorphRects {
    int nr, rw[4], rv[4], eh, ev;
};
inline bool morph_rects(const MorphParams &p, MorphRects &r) {
    if (p.kw > 7 || p.kh > 7 || !(p.kw & 1) || !(p.kh & 1) || p.ax != p.kw / 2 || p.ay != p.kh / 2) return false;
    int hw[7] = {-1, -1, -1, -1, -1, -1, -1};
    for (int k = 0; k < p.nruns; k++) {
        const MorphRun &run = p.runs[k];
        if (hw[run.ky] >= 0 || !(run.len & 1) || run.x0 + run.len / 2 != p.ax) return false;
        hw[run.ky] = run.len / 2;
    }
    const int ay = p.ay;
    if (hw[ay] < 0) return false;
    for (int d = 1; d <= ay; d++)
        if (hw[ay - d] != hw[ay + d] || hw[ay + d] > hw[ay + d - 1]) return false;
    r.nr = 0;
    r.ev = 0;
    for (int d = 0; d <= ay; d++) {
        const int next = d < ay ? hw[ay + d + 1] : -1;
        if (hw[ay + d] >= 0) r.ev = d;
        if (hw[ay + d] >= 0 && next < hw[ay + d]) {
            r.rw[r.nr] = hw[ay + d];
            r.rv[r.nr] = d;
            r.nr++;
        }
    }
    r.eh = hw[ay];
    return true;
}
// The element code of a union of centred rectangles (RectOf in morph_chain.hpp).
inline uint32_t morph_el_code(const MorphRects &r) {
    uint32_t el = (uint32_t)r.nr;
    for (int k = 0; k < r.nr; k++) el |= (uint32_t)(r.rw[k] | r.rv[k] << 2) << (4 + 4 * k);
    return el;
}

struct MorphSmallArgs;
int morph_small_args(const MorphParams &p, MorphSmallArgs &a);
// d = saturate(a - b) (morph.zig:574-593)
int launch_sub_sat(const DImg &a, const DImg &b, const DImg &d, int depth, hipStream_t st);

// Fused chains of erode/dilate passes on u8 images of 1-4 channels
// (k_morph_chain.hip) for elements that are unions of centred rectangles:
// rectangle k spans [-rw[k], rw[k]] samples x [-rv[k], rv[k]] rows.
// One launch stages a tile of `th` rows x `tw` bytes with `vb` rows and `hb`
// bytes of halo (rows = th + 2 vb) in LDS buffer 0 and runs `npass` passes
// between LDS buffers 0 .. nbuf - 1; the pass with dst -1 stores to `d`.
// A pass at position q of its chain writes rows [q ev, rows - q ev).
// post 1: d = saturate(C - v), 2: d = saturate(v - C), with C LDS buffer
// `comb` (0 .. 2) or, comb 3, image `c` staged after the pass buffers.
// cmask[w]: bit e + K set when a rectangle of half-width w reaches row
// offset e (K = max half-size, the kernel's frame).
struct MorphChainPass {
    int8_t src, dst, op_max, q;
};
constexpr int kChainMaxPasses = 12;
struct MorphChainArgs {
    DImg s, d, c;
    int tw, th, hb, vb, rows, ev, nbuf;
    int tiles_x, ntiles, xcd;
    int nr, rw[4], rv[4];
    uint32_t cmask[4];
    int npass, post, comb;
    MorphChainPass pass[kChainMaxPasses];
};
int morph_chain_pitch_dw(int cn, int K);  // dwords per LDS row
size_t morph_chain_lds_limit();           // LDS bytes a workgroup may use on the current device
int launch_morph_chain(const MorphChainArgs &a, int K, size_t lds, hipStream_t st);
// One pass per launch for one-channel images (morph_direct.hpp): strips
// of 512 bytes x bands of kMorphDirectRows rows, 4 bands per workgroup
// (bands4 groups).
constexpr int kMorphDirectRows = 8;
// With two passes, strips advance by 496 bytes (the lanes at both ends
// only feed the second pass). ops bit p: pass p dilates.
struct MorphDirectArgs {
    DImg s, d, c;                    // c: the combine image of post 1 / 2
    uint32_t span_s, span_d, span_c; // bytes from each image's first to last byte + 1 (< 2^31)
    int strips, bands4, ntiles, xcd;
    int ops;
};
namespace mstream {
// A set of centred rectangles (half-width a, half-height b), none inside
// another, ordered by b.
struct RectSet {
    int n;
    int a[16], b[16];
};

constexpr RectSet rects_of(uint32_t el) {
    RectSet r{};
    r.n = (int)(el & 15u);
    for (int k = 0; k < r.n; k++) {
        r.a[k] = (int)((el >> (4 + 4 * k)) & 3u);
        r.b[k] = (int)((el >> (6 + 4 * k)) & 3u);
    }
    return r;
}

// The Minkowski sum of two unions of centred rectangles: the pairwise sums,
// without those inside another.
constexpr RectSet minkowski(const RectSet &x, const RectSet &y) {
    RectSet o{};
    for (int i = 0; i < x.n; i++)
        for (int j = 0; j < y.n; j++) {
            const int a = x.a[i] + y.a[j], b = x.b[i] + y.b[j];
            bool inside = false;
            for (int k = 0; k < o.n; k++)
                if (o.a[k] >= a && o.b[k] >= b) inside = true;
            if (inside) continue;
            int m = 0;
            for (int k = 0; k < o.n; k++)
                if (!(a >= o.a[k] && b >= o.b[k])) {
                    o.a[m] = o.a[k];
                    o.b[m] = o.b[k];
                    m++;
                }
            o.a[m] = a;
            o.b[m] = b;
            o.n = m + 1;
        }
    for (int i = 1; i < o.n; i++)  // order by b
        for (int k = i; k > 0 && o.b[k] < o.b[k - 1]; k--) {
            const int ta = o.a[k], tb = o.b[k];
            o.a[k] = o.a[k - 1];
            o.b[k] = o.b[k - 1];
            o.a[k - 1] = ta;
            o.b[k - 1] = tb;
        }
    return o;
}

constexpr RectSet power(uint32_t el, int n) {
    RectSet r = rects_of(el);
    for (int k = 1; k < n; k++) r = minkowski(r, rects_of(el));
    return r;
}

constexpr int max_a(const RectSet &r) {
    int m = 0;
    for (int k = 0; k < r.n; k++) m = r.a[k] > m ? r.a[k] : m;
    return m;
}
constexpr int max_b(const RectSet &r) {
    int m = 0;
    for (int k = 0; k < r.n; k++) m = r.b[k] > m ? r.b[k] : m;
    return m;
}

// Half-width and half-height limits of the streamed kernel (12 bytes of halo).
constexpr int kMaxA = 12, kMaxB = 12;

// Accumulator rows an input row updates: the rectangles' heights summed.
constexpr int acc_rows(const RectSet &r) {
    int rows = 0;
    for (int k = 0; k < r.n; k++) rows += 2 * r.b[k] + 1;
    return rows;
}
// The streamed kernel has launches for (element, n) up to this many
// accumulator rows: measured on a 4096x4096 mask, one streamed pass beats n
// passes up to 9 (rect 3x3 taken 3 times 15 against 23 us, the 5x5
// ellipse's gradient 19 against 24 us) and not at 21 (5x5 ellipse twice 27
// against 21 us, 7x7 cross three times 61 against 34 us).
constexpr int kStreamMaxRows = 9;

}  // namespace mstream

// n = 2 or 3 iterated erode / dilate passes of a base element as one pass
// with its n-fold sum (morph_pow.hpp), for the sums up to half-width 11 and
// half-height 9 that morph_stream.hpp does not take (post as in
// launch_morph_band). False (nothing launched) for other combinations.
bool launch_morph_pow(const MorphDirectArgs &a, uint32_t el, int n, int op_max, int post, hipStream_t st);
bool morph_pow_has(uint32_t el, int n);

// One erode / dilate / gradient pass of a base element (an element code of
// morph_chain.hpp) taken n times in one launch (morph_stream.hpp), on one
// channel: mode 0 erode, 1 dilate, 2 gradient; post as above. False (nothing
// launched) when (element, n, mode, post) has no kernel.
struct MorphStreamArgs {
    DImg s, d, c;                     // c: the combine image of post 1 / 2
    uint32_t span_s, span_d, span_c;  // bytes from each image's first to last byte + 1 (< 2^31)
    int strips, ntiles, xcd;          // tiles: (strip, group of 4 bands)
    int tb;                           // output rows per band
};
bool launch_morph_stream(const MorphStreamArgs &a, uint32_t el, int n, int mode, int post, hipStream_t st);
bool launch_morph_direct1(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st);
// One erode (op_max 0) or dilate (1) pass of morph_band.hpp (one channel,
// the column window first); post 1 with dilate, 2 with erode. Strips and
// bands are set here. False (nothing launched) for other elements and posts.
bool launch_morph_band(const MorphDirectArgs &a, uint32_t el, int op_max, int post, hipStream_t st);
bool launch_morph_direct2(const MorphDirectArgs &a, uint32_t el, int post, hipStream_t st);

// --------------------------------------------------------------- histogram
int launch_hist_u8(const DImg &s, uint32_t *hist_dev, hipStream_t st);

}  // namespace visar
