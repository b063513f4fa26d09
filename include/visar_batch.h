/*
 * visar_batch.h -- many images of uneven sizes and mixed operations in one
 * GPU launch (no counterpart in the reference, whose hot path runs one CPU
 * call per image).
 *
 * Each item is one of the hot-path calls on device buffers; the result of an
 * item is bit for bit the result of the matching visar.h call. A plan
 * validates the items (with the checks of the matching visar.h call), splits
 * every image into tiles, groups the items by tile variant, cuts each image's
 * tiles into work units of about equal bytes and uploads them; a run
 * launches one kernel per variant in which every workgroup runs one unit.
 * Runs after the first replay a captured HIP graph of those launches.
 */
#ifndef VISAR_BATCH_H
#define VISAR_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "visar.h"

#ifdef __cplusplus
extern "C" {
#endif

enum visar_batch_op {
    VISAR_BATCH_GAUSSIAN_BLUR = 0, /* u8, 1/3 channels, ksize 3/5/7, any sigma and border */
    VISAR_BATCH_RESIZE = 1,        /* u8, 1/3 channels, INTER_LINEAR or INTER_AREA */
    VISAR_BATCH_WARP_AFFINE = 2,   /* u8, 1/3 channels, nearest or linear, any border */
    VISAR_BATCH_MORPHOLOGY = 3     /* u8, 1/3 channels, erode/dilate, elements up to 7 x 7 */
};

typedef struct visar_batch_item {
    int op;                /* enum visar_batch_op */
    visar_image src, dst;  /* device buffers */
    int ksize;             /* GaussianBlur kw = kh; MORPHOLOGY kw = kh */
    double sigma;          /* GaussianBlur sigma_x = sigma_y */
    int border;            /* GaussianBlur, WARP_AFFINE */
    int interpolation;     /* RESIZE: 1 LINEAR, 3 AREA; WARP_AFFINE: 0 nearest, else linear */
    double m[6];           /* WARP_AFFINE forward matrix */
    double value;          /* WARP_AFFINE constant border value */
    int morph_op;          /* MORPHOLOGY: 0 erode, 1 dilate */
    int shape;             /* MORPHOLOGY: 0 rect, 1 cross, 2 ellipse */
    int iterations;        /* MORPHOLOGY: must be 1 */
} visar_batch_item;

typedef struct visar_batch visar_batch;

int visar_batch_create(visar_batch **out);
/* Validates and plans n items; replaces any previous plan. Fails with the
 * status of the first invalid item and sets *bad_item to its index. */
int visar_batch_plan(visar_batch *b, const visar_batch_item *items, size_t n, size_t *bad_item);
/* Launches the planned batch on the calling thread's stream (visar_gpu.h);
 * waits for it when the thread's sync mode is on. use_graph != 0 replays a
 * HIP graph of the launch, captured on first use. */
int visar_batch_run(visar_batch *b, int use_graph);
/* Tiles and workgroups of the current plan. */
int visar_batch_info(const visar_batch *b, size_t *tiles, size_t *workgroups);
void visar_batch_free(visar_batch *b);

#ifdef __cplusplus
}
#endif
#endif
