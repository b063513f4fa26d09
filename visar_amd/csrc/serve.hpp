// serve.hpp -- the fused per-frame kernel of the served front end.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace visar {

constexpr int kServeMaxBatch = 256;

// Device buffers of one stream: input frame slots, ping-pong background
// models (the model of frame f is read from bg[f & 1] and written to
// bg[(f + 1) & 1], so a tile's halo reads never see a neighbour's update)
// and mask slots. Pitches are shared by all streams of a server.
struct ServeStream {
    uint8_t *frame[2];
    uint16_t *bg[2];
    uint8_t *mask[2];
};

struct ServeLaunch {
    const ServeStream *streams;  // device table, indexed by stream slot
    int w, h;
    long p8, p16;                // row pitches in bytes: u8 planes, u16 models
    int shift, thr, mthr;
    int kx[5];                   // GaussianBlur 5x5 sigma 0 taps (8-bit fixed, sum 256)
    int n;
    // job: stream slot | frame parity << 16 | primed << 17
    uint32_t job[kServeMaxBatch];
};

int launch_serve(const ServeLaunch &a, hipStream_t st);

}  // namespace visar
