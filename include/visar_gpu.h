/*
 * visar_gpu.h -- device controls of libvisar_amd.so (no counterpart in the
 * reference, whose hot path is CPU-only).
 *
 * The imgproc entry points of visar.h accept host or device pointers. These
 * functions choose the HIP stream the library launches on, whether calls on
 * device buffers wait for their kernels, and offer plain device buffers to
 * callers that do not link a HIP runtime of their own.
 */
#ifndef VISAR_GPU_H
#define VISAR_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "visar.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Stream used by the calling thread's subsequent calls (a hipStream_t; NULL is
 * HIP's default stream). visar_gpu_reset_stream() returns the thread to the
 * library's own non-blocking stream, which is used until a stream is set. */
int visar_gpu_set_stream(void *hip_stream);
int visar_gpu_reset_stream(void);
void *visar_gpu_get_stream(void);

/* 1 (default): a call whose images are all device buffers returns after its
 * kernels finish. 0: such calls only enqueue work on the stream. Calls that
 * stage host memory always finish before returning. Per calling thread. */
int visar_gpu_set_sync(int sync);
int visar_gpu_synchronize(void);

/* Device memory for callers without a HIP runtime of their own. */
int visar_gpu_malloc(size_t bytes, void **out);
int visar_gpu_free(void *p);
int visar_gpu_memcpy(void *dst, const void *src, size_t bytes, int kind /* 1 H2D, 2 D2H, 3 D2D */);

/* Launch-shape knobs read by the kernels (the autotuner's interface), 0 =
 * automatic: "stencil_rows" output rows per thread of the small separable
 * stencils (1, 2, 4); "batch_stencil_rows" the same inside batches (1, 2, 4);
 * "batch_unit_kb" KiB of input + output per batch workgroup (1-1024);
 * "tile_order" the order of a plain call's tiles over the XCDs (0 each
 * kernel's default, 1 row-major, 2 one band of neighbouring tiles per XCD);
 * "sep_rows" output rows per tile of the separable filter engine (1-16);
 * "warp_layout" the one-channel u8 bilinear constant-border warps' kernels
 * (0 block layout with 8-byte tap loads issued per pixel, 1 the row / LDS
 * source-box kernels, 2 the block layout with the loads after a wave-wide
 * check of the footprints); "morph_chain" u8 morphology (0 automatic:
 * one-channel images stream each pass through registers with the subtraction
 * in the last pass, others run fused chains of passes in one LDS tile, where
 * the element is a union of centred rectangles, and batched erode / dilate
 * with a full rectangle runs in streamed waves; 1 the general kernels, one
 * per erode / dilate pass and one per subtraction; 2 the fused LDS chains);
 * "morph_rows" output rows per tile of the fused chains (8-128);
 * "stencil_direct" the packed u8 stencils (GaussianBlur 3x3, 5x5, 7x7 with
 * sigma 0, Sobel 3x3 to i16): 0 rows streamed through registers (one-channel
 * rows of dword multiples starting on dwords with a border other than WRAP
 * in the one-channel kernel, other one- and three-channel GaussianBlur in the
 * general one; batched 5x5 / 7x7 GaussianBlur likewise), 1 the LDS tile
 * kernels; "stencil_band" output rows per wave of the streamed kernels (1, 2,
 * 4, 8, 16 for the one-channel kernel, 2 or 4 for the general one; 0 picks
 * from the image size);
 * "stencil_wg" their waves per workgroup (1, 2, 4); "gray_layout" u8
 * cvtColor to gray on contiguous 16-byte aligned images (0 each wave reads
 * one contiguous 1024-pixel block, 1 each thread its own 16 pixels);
 * "warp_blocks" 64 x 4 pixel blocks per wave of the default u8 warps (1, or
 * 2 with the second block's loads issued before the first is blended);
 * "morph_stream" one-channel u8 morphology with iterations or gradient (0 n
 * iterations as one pass with the n-fold element where that is faster, 1
 * pass by pass), "morph_band" its output rows per band (0 = 16),
 * "morph_pow_rows" the output rows per wave of the n-fold passes whose sums
 * have too many rectangles for the accumulator kernel (4, 8; 0 = 8);
 * "morph_pass" the other one-channel u8 single passes (0 column window
 * first, 1 the two-pass-capable kernel), "morph_pass_rows" their output rows
 * per wave (4, 8, 16; 0 = 4); "lin_stream" u8 LINEAR resize (0 source rows
 * streamed through registers where the tap tables allow, 1 the LDS tiles),
 * "lin_rows" the streamed kernel's output rows per band (4, 8) and
 * "lin_run" its bands per wave (1-16), 0 picking from the image size.
 * Unknown keys or values out of range return VISAR_INVALID_ARGUMENT. */
int visar_gpu_set_tuning(const char *key, int value);
int visar_gpu_get_tuning(const char *key, int *value);

/* visar_threshold without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and errors are those of visar_threshold, except that src
 * and dst must be device memory (else VISAR_INVALID_ARGUMENT "NotDeviceMemory")
 * and dst must be src itself or not overlap it ("Overlap"). The automatic
 * modes count their histogram in a 1 KiB device array that the library holds
 * per GPU until the process ends, so they take at most 2^32 - 1 pixels
 * ("TooManyPixels"), and two such calls that overlap in time on one GPU race.
 * Nothing is staged: the automatic modes wait once for the calling thread's
 * stream (the histogram comes back to the host); the fixed-threshold form
 * only enqueues its kernel unless the thread's sync mode is on. */
int visar_threshold_noalloc(const visar_image *src, const visar_image *dst, double thresh, double maxval, int type,
                            double *used);

/* Name and compute units of the current device, e.g. "gfx950 256 CUs". */
int visar_gpu_device_info(char *buf, size_t n);

#ifdef __cplusplus
}
#endif
#endif
