// tables.hpp -- host-side coefficient tables of the imgproc hot path.
//
// These are computed once per call on the host in IEEE double/float exactly
// as the reference computes them, then uploaded (or passed as kernel
// arguments); kernels never recompute them.
#pragma once

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace visar {

// filter.zig:350-399 OpenCV's softfloat exp in IEEE double operations.
double cv_exp(double x);
// filter.zig:407-443 getGaussianKernelBitExact.
void gaussian_kernel(double *out, int n, double sigma);
// filter.zig:458-476 fixed point with error diffusion, taps sum to 2^bits.
void gaussian_kernel_fixed(uint32_t *out, int n, double sigma, int bits);
// filter.zig:447-451 automatic size (-1 when sigma gives no valid size).
int gaussian_ksize(double sigma, bool u8);
// filter.zig:817-854 getDerivKernels; returns false on an invalid request.
bool deriv_kernel(int32_t *out, int n, int order, int ksize);
// filter.zig:213-229
int symmetry(const float *k, int n);

// resize.zig:186-214 linearTapU8 (offsets are source indices, not scaled by cn).
LinTap linear_tap_u8(long d, uint32_t ssize, uint32_t dsize, bool area, bool pin);

// resize.zig:154-180 linearCoord: source index and fractional weight,
// clamped to the edges with weight 0 on both axes.
void linear_coord(long d, double scale, uint32_t len, bool area, long *index, double *frac);

// resize.zig:470-493 computeResizeAreaTab, grouped per destination index.
void area_tab(uint32_t ssize, uint32_t dsize, double scale, std::vector<AreaEntry> &entries,
              std::vector<int32_t> &start);

// morph.zig:32-59 getStructuringElement (w*h mask, 1 = on).
void structuring_element(int shape, int w, int h, std::vector<uint8_t> &mask);

// geometry.zig:591-613
bool invert_affine(const double m[6], double out[6]);
bool invert3(const double m[9], double out[9]);

}  // namespace visar
