/*
 * visar.h -- C ABI of the MI355X (gfx950) imgproc hot path.
 *
 * Drop-in for the imgproc subset of VISAR's stable C ABI: same structure
 * layout, status codes, depth numbering and prototypes as the reference
 * (plyght/VISAR include/visar.h), implemented by libvisar_amd.so with HIP
 * kernels. Every prototype cites the reference declaration it replaces.
 *
 * Memory: `data` may point to host memory (staged through device buffers,
 * results are in place when the call returns) or to device memory from
 * hipMalloc (used in place; see visar_gpu.h for the stream and the
 * synchronisation mode). Functions return VISAR_OK (0) or a negative
 * visar_status; visar_last_error() names the failure (thread-local).
 */
#ifndef VISAR_H
#define VISAR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/visar.h:22-30 */
enum visar_status {
    VISAR_OK = 0,
    VISAR_INVALID_ARGUMENT = -1,
    VISAR_SIZE_MISMATCH = -2,
    VISAR_UNSUPPORTED = -3,
    VISAR_OUT_OF_MEMORY = -4,
    VISAR_INVALID_DATA = -5,
    VISAR_INTERNAL = -6
};

/* reference include/visar.h:32 (OpenCV depth numbering) */
enum visar_depth {
    VISAR_8U = 0,
    VISAR_8S = 1,
    VISAR_16U = 2,
    VISAR_16S = 3,
    VISAR_32S = 4,
    VISAR_32F = 5,
    VISAR_64F = 6,
    VISAR_16F = 7
};

/* reference include/visar.h:34-41: interleaved HWC pixels, stride in bytes */
typedef struct visar_image {
    uint32_t width;
    uint32_t height;
    uint32_t channels;
    uint32_t depth;
    size_t stride;
    void *data;
} visar_image;

/* reference include/visar.h:43-47 */
const char *visar_version(void);
const char *visar_build_info(void);
const char *visar_last_error(void);
void visar_set_num_threads(int n);
int visar_get_num_threads(void);

/* reference include/visar.h:50. GPU codes: 6 BGR2GRAY, 7 RGB2GRAY, 10 BGRA2GRAY,
 * 11 RGBA2GRAY (8U, 16U, 32F); other codes return VISAR_UNSUPPORTED. */
int visar_cvt_color(const visar_image *src, const visar_image *dst, int code);

/* reference include/visar.h:53 (capi.zig:176-182 -> videoio/frame.zig:170 grayDownsample):
 * BT.601 gray of a BGR or BGRA 8U image shrunk by k (1..16) with box averaging
 * in one pass; dst is (width / k) x (height / k), 1 channel. k <= 0:
 * VISAR_INVALID_ARGUMENT; k > 16, bad channels or sizes: VISAR_SIZE_MISMATCH. */
int visar_gray_downsample(const visar_image *src, const visar_image *dst, int k);

/* reference include/visar.h:54. 8U with 1-4 channels: INTER_LINEAR (1) and
 * INTER_AREA (3) in OpenCV's fixed point; INTER_NEAREST (0) for every depth. */
int visar_resize(const visar_image *src, const visar_image *dst, int interpolation);

/* reference include/visar.h:55. 8U/16U bit-exact fixed point; 16S/32F in f32. */
int visar_gaussian_blur(const visar_image *src, const visar_image *dst, int kw, int kh, double sigma_x,
                        double sigma_y, int border);

/* reference include/visar.h:57. dst 16S (from 8U) or 32F; ksize -1, 1, 3, 5, 7. */
int visar_sobel(const visar_image *src, const visar_image *dst, int dx, int dy, int ksize, double scale, int border);

/* reference include/visar.h:68. Fixed thresholds (types 0-4); THRESH_OTSU and
 * THRESH_TRIANGLE on 8U one-channel images. */
int visar_threshold(const visar_image *src, const visar_image *dst, double thresh, double maxval, int type,
                    double *used);

/* reference include/visar.h:72 */
int visar_morphology(const visar_image *src, const visar_image *dst, int op, int shape, int kw, int kh,
                     int iterations);

/* reference include/visar.h:75 */
int visar_warp_affine(const visar_image *src, const visar_image *dst, const double m[6], int interpolation, int border,
                      double value);

/* reference include/visar.h:191-195: running-average background model,
 * learning rate 1 / 2^shift; the model lives in device memory. */
typedef struct visar_background visar_background;
int visar_background_create(uint32_t width, uint32_t height, int shift, visar_background **out);
int visar_background_apply(visar_background *b, const visar_image *frame, const visar_image *mask, int threshold);
void visar_background_free(visar_background *b);

/* reference include/visar.h:78 (connectedComponentsWithStats, host code) */
int visar_connected_components(const visar_image *src, const visar_image *labels, int connectivity, int32_t *stats,
                               double *centroids, size_t max, size_t *count);

/* reference include/visar.h:340-341: 8U 1-4 channels, 32F maps, INTER_LINEAR,
 * BORDER_CONSTANT with `value`. */
int visar_remap(const visar_image *src, const visar_image *dst, const visar_image *mapx, const visar_image *mapy,
                double value);

/* reference include/visar.h:56 (capi.zig:208-217 -> imgproc/filter.zig:617 boxFilter):
 * 8U, 16U and 32F, odd kw, kh; normalize != 0 divides by kw * kh. 16S returns
 * VISAR_UNSUPPORTED (the reference's i16 sums are undefined for negative
 * samples), and so do kernels wider or taller than 127 (the reference: 255). */
int visar_box_filter(const visar_image *src, const visar_image *dst, int kw, int kh, int normalize, int border);

/* reference include/visar.h:58-60 (capi.zig:234-250 -> imgproc/pyramid.zig:65, 403):
 * 8U, 16U, 16S, 32F; dst ((w + 1) / 2, (h + 1) / 2) for pyr_down (any border
 * but constant) and (2w, 2h) for pyr_up (border 4, BORDER_DEFAULT, only). */
int visar_pyr_down(const visar_image *src, const visar_image *dst, int border);
int visar_pyr_up(const visar_image *src, const visar_image *dst, int border);

/* reference include/visar.h:69-71 (capi.zig:326-337 -> imgproc/threshold.zig:306):
 * 8U, one channel; method 0 MEAN_C, 1 GAUSSIAN_C; type 0 BINARY, 1 BINARY_INV;
 * block odd, 3 to 127 on the GPU (larger: VISAR_UNSUPPORTED). */
int visar_adaptive_threshold(const visar_image *src, const visar_image *dst, double maxval, int method, int type,
                             int block, double delta);

/* reference include/visar.h:333-335 (capi_calib3d.zig:186-191 -> calib3d/undistort.zig:205):
 * 8U, 1-4 channels, dst the size of src; k and new_k (NULL: k) are row-major 3x3;
 * dist has nd = 0, 4, 5 or 8 coefficients (k1 k2 p1 p2 [k3 [k4 k5 k6]]), NULL = none.
 * Rows of a stripe whose shifted new_k is singular are left as they were. */
int visar_undistort(const visar_image *src, const visar_image *dst, const double k[9], const double *dist, size_t nd,
                    const double *new_k);

/* reference include/visar.h:336-338 (capi_calib3d.zig:195-201 -> undistort.zig:82):
 * CV_32FC1 maps (VISAR_32F, one channel, the destination size); r and new_k may be NULL. */
int visar_init_undistort_rectify_map(const double k[9], const double *dist, size_t nd, const double *r,
                                     const double *new_k, const visar_image *mapx, const visar_image *mapy);

/* ---- Extensions: Zig-API-only functions of the reference exposed here ---- */

/* reference src/imgproc/geometry.zig:586 warpPerspective(exec, src, dst, m: [9]f64,
 * interp, border, value); forward 3x3 matrix, f64 coordinates. */
int visar_warp_perspective(const visar_image *src, const visar_image *dst, const double m[9], int interpolation,
                           int border, double value);

/* reference src/imgproc/filter.zig:909 sepFilter2D(exec, gpa, src, dst, kx: []const f32,
 * ky: []const f32, border); src and dst 8U/16U/16S/32F, f32 accumulation. */
int visar_sep_filter_2d(const visar_image *src, const visar_image *dst, const float *kx, size_t nkx, const float *ky,
                        size_t nky, int border);

/* reference src/videoio/frame.zig:182 grayDownsampleRect(exec, src, dst, k, x, y, w, h):
 * only dst's (x, y, w, h) window, bit-identical to visar_gray_downsample (same
 * argument checks); the rest of dst is left as it is. A window outside dst: VISAR_INVALID_ARGUMENT
 * (the reference's OutOfBounds); w or h zero does nothing. */
int visar_gray_downsample_rect(const visar_image *src, const visar_image *dst, int k, uint32_t x, uint32_t y,
                               uint32_t w, uint32_t h);

#ifdef __cplusplus
}
#endif
#endif
