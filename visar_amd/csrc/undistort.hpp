// undistort.hpp -- launch arguments of the undistortion kernels
// (calib3d/undistort.zig).
#pragma once

#include "common.hpp"

namespace visar {

// The lens of MapParams (undistort.zig:26-33): K's focal lengths and
// principal point and the 8 distortion coefficients (k1 k2 p1 p2 k3 k4 k5
// k6, zero-padded); `rational` when k4, k5 or k6 is nonzero.
struct UndistortLens {
    double k[8];
    double fx, fy, u0, v0;
    int rational;
};

// undistort: `ar` is new_k (or K); rows go in stripes of `stripe` rows, each
// with the inverse of `ar` whose principal point is moved up by the stripe's
// first row.
struct UndistortArgs {
    DImg s, d;
    double ar[9];
    UndistortLens lens;
    int stripe;
};

// initUndistortRectifyMap: `ir` = inv(new_k * R), f32 one-channel maps.
struct UndistortMapArgs {
    DImg mx, my;
    double ir[9];
    UndistortLens lens;
};

// calib3d/linalg.zig:417-426 inv3: cofactors times 1 / det (false when det
// is 0); host and device run the same IEEE operations.
__host__ __device__ inline bool inv3(const double a[9], double r[9]) {
    const double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[6] * a[5]) +
                       a[2] * (a[3] * a[7] - a[6] * a[4]);
    if (det == 0) return false;
    const double d = 1 / det;
    r[0] = (a[4] * a[8] - a[5] * a[7]) * d;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * d;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * d;
    r[3] = (a[5] * a[6] - a[3] * a[8]) * d;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * d;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * d;
    r[6] = (a[3] * a[7] - a[4] * a[6]) * d;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * d;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * d;
    return true;
}

int launch_undistort(const UndistortArgs &p, hipStream_t st);
int launch_undistort_map(const UndistortMapArgs &p, hipStream_t st);

}  // namespace visar
