/*
 * visar_serve.h -- served tracking front end: many camera streams, one GPU
 * (no counterpart in the reference, whose tracker runs one Detector per
 * stream on the CPU, demos/track_lib.zig:452-557).
 *
 * Per frame the server runs the reference detector's front end on the GPU,
 * with the results of the plain calls bit for bit:
 *   visar_gaussian_blur 5x5 (sigma 0, BORDER_REFLECT_101)
 *   -> visar_background_apply (shift, threshold; the stream's model stays
 *      on the GPU for the stream's whole life)
 *   -> visar_threshold (mask_threshold, maxval 255, THRESH_BINARY)
 *   -> visar_morphology (MORPH_OPEN, 3x3 MORPH_RECT, 1 iteration)
 * and then visar_connected_components (8-connectivity) on the host.
 *
 * Streams are admitted while their device memory fits the configured HBM
 * budget. Frames are copied at submit time; a dispatcher thread batches the
 * frames that are ready (one per stream per batch, a stream's frames in
 * order) into one fused launch while earlier batches are still on the GPU
 * (continuous batching), and host workers label each returned mask. Results
 * are collected with visar_serve_poll.
 */
#ifndef VISAR_SERVE_H
#define VISAR_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "visar.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct visar_serve_config {
    uint32_t width, height;   /* frame size, 1-channel u8 */
    int shift;                /* background learning shift, 0..15 (reference default 4) */
    int threshold;            /* background difference threshold, clamped to 0..255 (default 28) */
    int mask_threshold;       /* THRESH_BINARY threshold on the mask, 0..254 (127: identity on {0,255}) */
    uint32_t min_area;        /* components at least this large count as objects */
    uint64_t hbm_budget;      /* device bytes all admitted streams may hold; 0 = no limit */
    uint32_t max_streams;     /* 1..4096 */
    uint32_t max_batch;       /* frames per launch, 1..256 */
    uint32_t queue_depth;     /* frames a stream may have submitted and unfinished, 1..64 */
    uint32_t host_workers;    /* connected-components threads, 1..64 */
    uint32_t max_inflight;    /* batches on the GPU at once, 1..8 */
    uint32_t batch_window_us; /* with fewer than max_batch streams ready, wait up to this long
                                 after the oldest ready frame arrived for more (0: launch at once;
                                 at most 1 s) */
} visar_serve_config;

typedef struct visar_serve_result {
    int32_t stream_id;
    int32_t status;           /* VISAR_OK, or the failure of this frame */
    uint64_t tag;             /* the caller's tag from visar_serve_submit */
    uint64_t frame_index;     /* 0, 1, 2, ... per stream */
    uint32_t components;      /* labels including the background */
    uint32_t objects;         /* components with area >= min_area */
    uint64_t checksum;        /* hash of the mask and the component statistics */
    double t_submit;          /* seconds on the server's steady clock */
    double t_done;            /* when the components of this frame were done */
    uint32_t batch_size;      /* frames in the launch that carried this frame */
    uint32_t reserved;
} visar_serve_result;

typedef struct visar_serve_stats {
    uint64_t frames_done, batches, admitted, rejected;
    uint64_t device_bytes;    /* held by admitted streams */
    uint64_t bytes_per_stream;
    uint32_t max_batch_seen, active_streams;
} visar_serve_stats;

typedef struct visar_serve visar_serve;

/* Fills the defaults (1280 x 720, shift 4, threshold 28, mask threshold
 * 127, min area 250, no budget, 64 streams, batch 64, depth 4, 4 workers,
 * 2 batches in flight, no batching window). */
void visar_serve_config_default(visar_serve_config *cfg);
int visar_serve_create(const visar_serve_config *cfg, visar_serve **out);
/* VISAR_OUT_OF_MEMORY ("HbmBudget") when the stream does not fit the budget
 * or max_streams; otherwise *stream_id is the new stream. */
int visar_serve_admit(visar_serve *s, int32_t *stream_id);
/* Waits for the stream's frames to finish, then frees its memory. */
int visar_serve_release(visar_serve *s, int32_t stream_id);
/* Copies one frame (width x height u8, row stride in bytes). Blocks while
 * the stream has queue_depth frames unfinished. When mask_out is not NULL,
 * the frame's final mask is written there (row stride mask_stride) before
 * its result is posted. */
int visar_serve_submit(visar_serve *s, int32_t stream_id, const uint8_t *frame, size_t stride, uint64_t tag,
                       uint8_t *mask_out, size_t mask_stride);
/* Moves up to max finished results to out; waits up to wait_ms for the
 * first one (0: no wait). */
int visar_serve_poll(visar_serve *s, visar_serve_result *out, size_t max, size_t *n, int wait_ms);
/* Waits until every submitted frame has its result posted. */
int visar_serve_drain(visar_serve *s);
int visar_serve_get_stats(visar_serve *s, visar_serve_stats *st);
/* Measurement: times `reps` launches of the fused per-frame kernel over the
 * resident buffers of every admitted stream (one frame each, no copies, the
 * models read from one buffer and written to the other so every launch does
 * the same work) with HIP events on the server's stream. Needs an idle
 * server (every submitted frame finished). */
int visar_serve_time_kernel(visar_serve *s, uint32_t reps, double *ms_per_launch, uint32_t *frames_per_launch);
double visar_serve_now(const visar_serve *s);
void visar_serve_free(visar_serve *s);

#ifdef __cplusplus
}
#endif
#endif
