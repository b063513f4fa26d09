// runtime.hpp -- per-call staging of visar_image descriptors onto the device.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/visar.h"
#include "common.hpp"

namespace visar {

int fail(int status, const char *name);
int ensure_runtime();
hipStream_t current_stream();
bool sync_mode();
bool is_device_ptr(const void *p);
int check_image(const visar_image *c, uint32_t depth);
void *pool_get(size_t bytes);
void pool_put(void *p);

// One C-ABI call: resolves its images to device views (device buffers in
// place, host memory staged through pooled device buffers), owns the
// temporaries, and on finish() copies staged outputs back and waits when
// the call's semantics require it. Temporaries return to the pool when the
// Call is destroyed: at once when finish() waited for the stream, otherwise
// once the work queued so far on the call's stream has finished.
class Call {
  public:
    Call() = default;
    Call(const Call &) = delete;
    Call &operator=(const Call &) = delete;
    ~Call();

    int begin();
    int input(const visar_image *c, DImg &out);
    int output(const visar_image *c, DImg &out);
    // An output that the kernels may write only in part: a staged or
    // temporary buffer starts as a copy of the destination's contents.
    int output_keep(const visar_image *c, DImg &out);
    int temp_like(const DImg &like, size_t elem, DImg &out);
    void *scratch(size_t bytes);
    // A device copy of `bytes` host bytes, uploaded on the call's stream
    // through pinned memory (the host data may change once this returns).
    void *upload(const void *src, size_t bytes);
    // Pinned host memory owned by the call (returned with its buffers); the
    // HIP runtime's own copies to and from pageable memory are slow.
    void *pinned(size_t bytes);
    int finish(int launch_status);
    hipStream_t stream() const { return st_; }

  private:
    struct Pending {
        const visar_image *desc;
        void *buf;
        size_t pitch;
        bool device_target;
        void *pin;  // host targets: the pinned staging rows
    };
    DImg view(const visar_image *c, void *p, size_t stride);

    hipStream_t st_ = nullptr;  // nullptr is also HIP's default stream
    bool begun_ = false;
    bool staged_ = false;
    bool waited_ = false;
    std::vector<void *> owned_;
    std::vector<void *> pinned_;
    std::vector<Pending> pending_;
    std::vector<std::pair<const uint8_t *, size_t>> dev_ranges_;
};

}  // namespace visar
