/*
 * visar_oracle.h -- CPU restatement of VISAR's imgproc hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * entry point and the CPU-baseline leg of bench.py may call it; the product
 * library (visar_amd) never links, loads or calls it.
 *
 * Every function restates the reference's scalar arithmetic in plain C99 so
 * results are bit-exact with VISAR (plyght/VISAR 0.1.0, Zig). The image
 * descriptor has the layout of visar_image (include/visar.h:34-41 in the
 * reference) so the same ctypes structure drives both.
 *
 * Pinning. The reference cannot be built here (Zig), so each function is
 * checked by tests/test_oracle_golden.py against the reference's own known
 * answers, closed forms, brute force or invariants. Parity partially
 * unpinned: the exact f32 operation order of orc_warp_affine,
 * orc_warp_perspective, orc_remap, the fractional AREA path of orc_resize,
 * the generic LINEAR taps and orc_sep_filter_2d follows the cited Zig but has
 * no reference vectors (the reference checks these only against cv2 or not
 * at all); the invariants catch dropped terms, sign and index errors but not
 * FMA-versus-separate rounding.
 */
#ifndef VISAR_ORACLE_H
#define VISAR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_image {
    uint32_t width;
    uint32_t height;
    uint32_t channels;
    uint32_t depth; /* 0 u8, 2 u16, 3 i16, 5 f32 */
    size_t stride;  /* bytes between row starts */
    void *data;
} orc_image;

/* Status codes follow visar_status. */
enum { ORC_OK = 0, ORC_INVALID_ARGUMENT = -1, ORC_SIZE_MISMATCH = -2, ORC_UNSUPPORTED = -3,
       ORC_OUT_OF_MEMORY = -4, ORC_INVALID_DATA = -5, ORC_INTERNAL = -6 };

/* core semantics */
long orc_border_map(long p, long len, int border);
int orc_sat_u8_f32(float x);          /* sat.cast(u8, f32) */
int orc_sat_i16_f32(float x);         /* sat.cast(i16, f32) */
int orc_satf32_u8(float x);           /* simd.satF32(u8, .) */
int orc_satf32_i16(float x);          /* simd.satF32(i16, .) */

/* Gaussian kernels */
double orc_cv_exp(double x);
void orc_gaussian_kernel(double *out, int n, double sigma);
void orc_gaussian_kernel_fixed(uint32_t *out, int n, double sigma, int bits);
int orc_gaussian_ksize(double sigma, int depth_is_u8);
int orc_deriv_kernel(int32_t *out, int n, int order, int ksize);

/* filters */
int orc_gaussian_blur(const orc_image *src, const orc_image *dst, int kw, int kh, double sx, double sy, int border);
int orc_sobel(const orc_image *src, const orc_image *dst, int dx, int dy, int ksize, double scale, int border);
int orc_sep_filter_2d(const orc_image *src, const orc_image *dst, const float *kx, size_t nkx, const float *ky,
                      size_t nky, int border);

/* color / threshold */
int orc_cvt_color(const orc_image *src, const orc_image *dst, int code);
int orc_box_filter(const orc_image *src, const orc_image *dst, int kw, int kh, int normalize, int border);
int orc_pyr_down(const orc_image *src, const orc_image *dst, int border);
int orc_pyr_up(const orc_image *src, const orc_image *dst, int border);
int orc_adaptive_threshold(const orc_image *src, const orc_image *dst, double maxval, int method, int type, int block,
                           double delta);
int orc_undistort(const orc_image *src, const orc_image *dst, const double K[9], const double *dist, size_t nd,
                  const double *new_k);
int orc_init_undistort_rectify_map(const double K[9], const double *dist, size_t nd, const double *R, const double *new_k,
                                   const orc_image *mapx, const orc_image *mapy);
int orc_gray_downsample(const orc_image *src, const orc_image *dst, uint32_t k);
int orc_gray_downsample_rect(const orc_image *src, const orc_image *dst, uint32_t k, uint32_t x, uint32_t y, uint32_t w,
                             uint32_t h);
int orc_threshold(const orc_image *src, const orc_image *dst, double thresh, double maxval, int type, double *used);

/* resize */
int orc_resize(const orc_image *src, const orc_image *dst, int interp);
/* Timing hook: restrict this thread's row-independent functions to
 * destination rows [lo, hi) (hi < 0: all rows), so one image can be split
 * across threads in row bands. */
void orc_set_band(long lo, long hi);
/* d = a - b saturating at 0 (morph.zig:574-593). */
int orc_sub_sat(const orc_image *a, const orc_image *b, const orc_image *d);
/* Test hooks: generic u8 LINEAR (resize.zig:216-242) and fractional AREA
 * (resize.zig:527-557) whatever the size ratio, the AREA table
 * (resize.zig:470-493; returns the entry count, -1 on bad sizes) and one
 * LINEAR tap (resize.zig:186-214). */
int orc_resize_linear_generic(const orc_image *src, const orc_image *dst);
int orc_resize_area_generic(const orc_image *src, const orc_image *dst);
long orc_area_tab(uint32_t ssize, uint32_t dsize, uint32_t *di, uint32_t *si, float *alpha, long cap);
void orc_linear_tap(long d, uint32_t ssize, uint32_t dsize, int area, int pin, long o[2], int w[2]);

/* warps */
int orc_warp_affine(const orc_image *src, const orc_image *dst, const double m[6], int interp, int border, double value);
int orc_warp_perspective(const orc_image *src, const orc_image *dst, const double m[9], int interp, int border,
                         double value);
int orc_remap(const orc_image *src, const orc_image *dst, const orc_image *mapx, const orc_image *mapy, double value);

/* morphology */
int orc_structuring_element(int shape, int w, int h, uint8_t *buf);
int orc_morph_pass(const orc_image *src, const orc_image *dst, int op_max, const uint8_t *mask, int kw, int kh,
                   int ax, int ay);
int orc_morphology(const orc_image *src, const orc_image *dst, int op, int shape, int kw, int kh, int iterations);

/* video: background model. model is w*h u16 (8.8 fixed), *primed in/out. */
int orc_background_apply(uint16_t *model, int *primed, int shift, const orc_image *frame, const orc_image *mask,
                         int threshold);

/* connected components with stats (8U 1-channel source, labels 32S). */
int orc_connected_components(const orc_image *src, const orc_image *labels, int connectivity, int32_t *stats,
                             double *centroids, size_t max, size_t *count);

#ifdef __cplusplus
}
#endif
#endif
