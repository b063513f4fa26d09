/*
 * visar_oracle.c -- CPU restatement of VISAR's imgproc hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only tests,
 * the smoke entry point and bench.py's CPU-baseline leg use it.
 *
 * Scalar C99, built with -ffp-contract=off and without -ffast-math, so every
 * float expression is evaluated as written (one IEEE rounding per operation,
 * fma() only where the reference uses @mulAdd). Each function names the
 * reference lines it follows (paths relative to plyght/VISAR 0.1.0, src/).
 *
 * Parity pins (tests/test_oracle_golden.py): the known-answer vectors of the
 * reference's own unit tests, the closed-form Background model of
 * tests/py/test_motion_vs_opencv.py:191-197, brute-force integer
 * convolution, and invariants (constant images, identity warps, etc.).
 */
#include "visar_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ helpers */

/* Row band of the calling thread, for timing the restatement with one image
 * split across threads in row bands (the reference's Exec.parallelFor,
 * pool.zig:31-37): functions whose destination rows are independent visit
 * only rows [lo, hi) of them. The checker never sets it (the whole image). */
static __thread long g_band_lo = 0, g_band_hi = -1;
void orc_set_band(long lo, long hi) {
    g_band_lo = lo;
    g_band_hi = hi;
}
static inline long band_lo(void) { return g_band_lo; }
static inline long band_hi(long n) { return g_band_hi < 0 || g_band_hi > n ? n : g_band_hi; }

static size_t dsize(uint32_t depth) {
    switch (depth) {
    case 0: return 1;
    case 2: return 2;
    case 3: return 2;
    case 4: return 4;
    case 5: return 4;
    default: return 0;
    }
}

/* capi.zig:123-132 + image.zig:71-77, 172-176 */
static int check(const orc_image *im, uint32_t depth) {
    if (!im || !im->data) return ORC_INVALID_ARGUMENT;
    if (im->depth != depth) return ORC_UNSUPPORTED;
    if (im->channels == 0 || im->channels > 512) return ORC_INVALID_ARGUMENT;
    if (im->stride % dsize(depth) != 0) return ORC_INVALID_ARGUMENT;
    if (im->width == 0 || im->height == 0) return ORC_INVALID_ARGUMENT;
    if (im->stride < (size_t)im->width * im->channels * dsize(depth)) return ORC_INVALID_ARGUMENT;
    return ORC_OK;
}

static inline void *rowp(const orc_image *im, long y) { return (char *)im->data + (size_t)y * im->stride; }

/* Loads sample i of row y as a double (exact for every supported depth). */
static inline double ldd(const orc_image *im, long y, long i) {
    const void *r = rowp(im, y);
    switch (im->depth) {
    case 0: return ((const uint8_t *)r)[i];
    case 2: return ((const uint16_t *)r)[i];
    case 3: return ((const int16_t *)r)[i];
    case 4: return ((const int32_t *)r)[i];
    default: return ((const float *)r)[i];
    }
}
static inline int64_t ldi(const orc_image *im, long y, long i) { return (int64_t)ldd(im, y, i); }
static inline float ldf(const orc_image *im, long y, long i) {
    if (im->depth == 5) return ((const float *)rowp(im, y))[i];
    return (float)ldi(im, y, i);
}

static inline void st_i(const orc_image *im, long y, long i, int64_t v) {
    void *r = rowp(im, y);
    switch (im->depth) {
    case 0: ((uint8_t *)r)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)r)[i] = (uint16_t)v; break;
    case 3: ((int16_t *)r)[i] = (int16_t)v; break;
    case 4: ((int32_t *)r)[i] = (int32_t)v; break;
    default: break;
    }
}
static inline void st_f(const orc_image *im, long y, long i, float v) { ((float *)rowp(im, y))[i] = v; }

static void depth_range(uint32_t depth, double *lo, double *hi) {
    switch (depth) {
    case 0: *lo = 0; *hi = 255; break;
    case 2: *lo = 0; *hi = 65535; break;
    case 3: *lo = -32768; *hi = 32767; break;
    default: *lo = -2147483648.0; *hi = 2147483647.0; break;
    }
}

/* border.zig:22-39 */
long orc_border_map(long p, long n, int border) {
    if (p >= 0 && p < n) return p;
    switch (border) {
    case 0: return -1;
    case 1: return p < 0 ? 0 : n - 1;
    case 2:
    case 4: {
        if (n == 1) return 0;
        long d = border == 4 ? 1 : 0;
        long q = p;
        while (q < 0 || q >= n) {
            if (q < 0) q = -q - 1 + d;
            else q = n - 1 - (q - n) - d;
        }
        return q;
    }
    case 3: {
        long m = p % n;
        if (m < 0) m += n;
        return m;
    }
    default: return -1;
    }
}

/* saturate.zig:8-18: magic-number round half to even. */
static inline double rint64(double x) { return (x + 6755399441055744.0) - 6755399441055744.0; }
static inline float rint32(float x) { return (x + 12582912.0f) - 12582912.0f; }

/* IEEE minNum/maxNum as Zig's @min/@max: a NaN operand yields the other. */
static inline float fmaxn(float a, float b) { return isnan(a) ? b : isnan(b) ? a : (a > b ? a : b); }
static inline float fminn(float a, float b) { return isnan(a) ? b : isnan(b) ? a : (a < b ? a : b); }

/* saturate.zig:21-49 sat.cast(D, float): NaN -> 0, rint in f64, clamp. */
static int64_t sat_cast_f(double x, uint32_t depth) {
    double lo, hi;
    if (isnan(x)) return 0;
    depth_range(depth, &lo, &hi);
    double r = rint64(x);
    if (r <= lo) return (int64_t)lo;
    if (r >= hi) return (int64_t)hi;
    return (int64_t)r;
}
int orc_sat_u8_f32(float x) { return (int)sat_cast_f((double)x, 0); }
int orc_sat_i16_f32(float x) { return (int)sat_cast_f((double)x, 3); }

/* simd.zig:90-126 satF32: clamp (NaN -> low bound, then 0 for signed), rint. */
static int64_t satf32(float v, uint32_t depth) {
    double lo, hi;
    depth_range(depth, &lo, &hi);
    float c = fminn(fmaxn(v, (float)lo), (float)hi);
    if (lo != 0 && isnan(v)) c = 0;
    return (int64_t)rint64((double)c);
}
int orc_satf32_u8(float x) { return (int)satf32(x, 0); }
int orc_satf32_i16(float x) { return (int)satf32(x, 3); }

/* filter.zig:44-58 satV(D, f32 vector): rint in f32, clamp with maxNum/minNum
 * (NaN -> low bound), convert. */
static int64_t satv_f(float v, uint32_t depth) {
    double lo, hi;
    depth_range(depth, &lo, &hi);
    float r = rint32(v);
    float c = fminn(fmaxn(r, (float)lo), (float)hi);
    return (int64_t)c;
}

/* ------------------------------------------------------ Gaussian kernels */

/* filter.zig:350-399: OpenCV's softfloat exp as IEEE double operations. */
static double from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
double orc_cv_exp(double x) {
    static const uint64_t raw[64] = {
        0x3ff0000000000000ull, 0x3ff02c9a3e778061ull, 0x3ff059b0d3158574ull, 0x3ff0874518759bc8ull,
        0x3ff0b5586cf9890full, 0x3ff0e3ec32d3d1a2ull, 0x3ff11301d0125b51ull, 0x3ff1429aaea92de0ull,
        0x3ff172b83c7d517bull, 0x3ff1a35beb6fcb75ull, 0x3ff1d4873168b9aaull, 0x3ff2063b88628cd6ull,
        0x3ff2387a6e756238ull, 0x3ff26b4565e27cddull, 0x3ff29e9df51fdee1ull, 0x3ff2d285a6e4030bull,
        0x3ff306fe0a31b715ull, 0x3ff33c08b26416ffull, 0x3ff371a7373aa9cbull, 0x3ff3a7db34e59ff7ull,
        0x3ff3dea64c123422ull, 0x3ff4160a21f72e2aull, 0x3ff44e086061892dull, 0x3ff486a2b5c13cd0ull,
        0x3ff4bfdad5362a27ull, 0x3ff4f9b2769d2ca7ull, 0x3ff5342b569d4f82ull, 0x3ff56f4736b527daull,
        0x3ff5ab07dd485429ull, 0x3ff5e76f15ad2148ull, 0x3ff6247eb03a5585ull, 0x3ff6623882552225ull,
        0x3ff6a09e667f3bcdull, 0x3ff6dfb23c651a2full, 0x3ff71f75e8ec5f74ull, 0x3ff75feb564267c9ull,
        0x3ff7a11473eb0187ull, 0x3ff7e2f336cf4e62ull, 0x3ff82589994cce13ull, 0x3ff868d99b4492edull,
        0x3ff8ace5422aa0dbull, 0x3ff8f1ae99157736ull, 0x3ff93737b0cdc5e5ull, 0x3ff97d829fde4e50ull,
        0x3ff9c49182a3f090ull, 0x3ffa0c667b5de565ull, 0x3ffa5503b23e255dull, 0x3ffa9e6b5579fdbfull,
        0x3ffae89f995ad3adull, 0x3ffb33a2b84f15fbull, 0x3ffb7f76f2fb5e47ull, 0x3ffbcc1e904bc1d2ull,
        0x3ffc199bdd85529cull, 0x3ffc67f12e57d14bull, 0x3ffcb720dcef9069ull, 0x3ffd072d4a07897cull,
        0x3ffd5818dcfba487ull, 0x3ffda9e603db3285ull, 0x3ffdfc97337b9b5full, 0x3ffe502ee78b3ff6ull,
        0x3ffea4afa2a490daull, 0x3ffefa1bee615a27ull, 0x3fff50765b6e4540ull, 0x3fffa7c1819e90d8ull,
    };
    if (isnan(x)) return x;
    if (isinf(x)) return x > 0 ? x : 0;
    const double a0 = from_bits(0x3f83ce0f3e46f431ull);
    const double p5 = 1.0 / a0;
    const double p4 = from_bits(0x3fe62e42fefa39f1ull) / a0;
    const double p3 = from_bits(0x3fcebfbdff82a45aull) / a0;
    const double p2 = from_bits(0x3fac6b08d81fec75ull) / a0;
    const double p1 = from_bits(0x3f83b2a72b4f3cd3ull) / a0;
    const double p0 = from_bits(0x3f55e7aa1566c2a4ull) / a0;
    const double prescale = from_bits(0x3ff71547652b82feull) * 64.0;
    const double max_val = 3000.0 * 64.0;
    uint64_t bits;
    memcpy(&bits, &x, 8);
    double x0 = (((bits >> 52) & 0x7ff) > 1023 + 10) ? (x < 0 ? -max_val : max_val) : x * prescale;
    double r = rint64(x0);
    int32_t v = (int32_t)r;
    long t = (long)(v >> 6) + 1023;
    if (t < 0) t = 0;
    if (t > 2047) t = 2047;
    double buf = from_bits((uint64_t)t << 52);
    x0 = (x0 - r) * (1.0 / 64.0);
    double poly = ((((p0 * x0 + p1) * x0 + p2) * x0 + p3) * x0 + p4) * x0 + p5;
    return buf * a0 * from_bits(raw[v & 63]) * poly;
}

/* filter.zig:407-443 getGaussianKernelBitExact */
void orc_gaussian_kernel(double *out, int n, double sigma) {
    static const double s1[] = {1.0};
    static const double s3[] = {0.25, 0.5, 0.25};
    static const double s5[] = {0.0625, 0.25, 0.375, 0.25, 0.0625};
    static const double s7[] = {0.03125, 0.109375, 0.21875, 0.28125, 0.21875, 0.109375, 0.03125};
    static const double s9[] = {4.0 / 256.0, 13.0 / 256.0, 30.0 / 256.0, 51.0 / 256.0, 60.0 / 256.0,
                                51.0 / 256.0, 30.0 / 256.0, 13.0 / 256.0, 4.0 / 256.0};
    if (n % 2 == 1 && n <= 9 && sigma <= 0) {
        const double *t = n == 1 ? s1 : n == 3 ? s3 : n == 5 ? s5 : n == 7 ? s7 : s9;
        memcpy(out, t, sizeof(double) * (size_t)n);
        return;
    }
    double nf = (double)n;
    double sx = sigma > 0 ? sigma : fma(nf, 0.15, 0.35);
    double scale2x = -0.125 / (sx * sx);
    int half = (n - 1) / 2;
    double sum = 0;
    long x = 1 - (long)n;
    for (int i = 0; i < half; i++) {
        double t = orc_cv_exp((double)(x * x) * scale2x);
        out[i] = t;
        sum += t;
        x += 2;
    }
    sum = sum * 2 + 1;
    if (n % 2 == 0) sum += 1;
    double mul1 = 1.0 / sum;
    for (int i = 0; i < half; i++) {
        out[i] *= mul1;
        out[n - 1 - i] = out[i];
    }
    out[half] = mul1;
    if (n % 2 == 0) out[half + 1] = mul1;
}

/* filter.zig:458-476 getGaussianKernelFixedPoint_ED */
void orc_gaussian_kernel_fixed(uint32_t *out, int n, double sigma, int bits) {
    double k[128];
    orc_gaussian_kernel(k, n, sigma);
    double m = (double)((uint64_t)1 << bits);
    double err = 0;
    uint64_t sum = 0;
    for (int i = 0; i < n / 2; i++) {
        double a = k[i] * m + err;
        double v = rint64(a);
        err = a - v;
        uint64_t q = (uint64_t)(v > 0 ? v : 0);
        out[i] = (uint32_t)q;
        out[n - 1 - i] = (uint32_t)q;
        sum += q;
    }
    out[n / 2] = (uint32_t)(((uint64_t)1 << bits) - 2 * sum);
}

/* filter.zig:447-451 */
int orc_gaussian_ksize(double sigma, int depth_is_u8) {
    double k = depth_is_u8 ? 3 : 4;
    double r = round(sigma * k * 2 + 1);
    if (!(r >= 0) || r > 1e6) return -1;
    return ((int)r) | 1;
}

/* filter.zig:817-854 getDerivKernels */
int orc_deriv_kernel(int32_t *out, int n_out, int order, int ksize) {
    if (ksize == -1) {
        if (n_out != 3 || order > 1) return ORC_INVALID_ARGUMENT;
        if (order == 0) { out[0] = 3; out[1] = 10; out[2] = 3; }
        else { out[0] = -1; out[1] = 0; out[2] = 1; }
        return ORC_OK;
    }
    int n = ksize == 1 ? 3 : ksize;
    if (n_out != n || n > 31) return ORC_INVALID_ARGUMENT;
    if (ksize == 1 && order == 0) { out[0] = 0; out[1] = 1; out[2] = 0; return ORC_OK; }
    if (order >= n) return ORC_INVALID_ARGUMENT;
    int32_t a[33] = {0};
    a[0] = 1;
    int len = 1;
    for (int s = 0; s < n - order - 1; s++) {
        int32_t b[33] = {0};
        for (int j = 0; j < len; j++) { b[j] += a[j]; b[j + 1] += a[j]; }
        memcpy(a, b, sizeof a);
        len++;
    }
    for (int s = 0; s < order; s++) {
        int32_t b[33] = {0};
        for (int j = 0; j < len; j++) { b[j] -= a[j]; b[j + 1] += a[j]; }
        memcpy(a, b, sizeof a);
        len++;
    }
    memcpy(out, a, sizeof(int32_t) * (size_t)n);
    return ORC_OK;
}

/* ------------------------------------------------------ separable engine */

/* Finish kinds of filter.zig:61-116. */
enum { FIN_FIX8, FIN_FIX16, FIN_CAST, FIN_SCALED };

/*
 * Integer separable filter (filter.zig:120-205 with exact integer
 * accumulators). The reference's accumulators (u16/u32/u64/i16/i32) never
 * overflow or saturate within its own kernel bounds, so exact int64 sums
 * reproduce it; the pass order does not matter (filter.zig:163-167).
 */
static int sep_int(const orc_image *s, const orc_image *d, const int64_t *kx, int nkx, const int64_t *ky, int nky,
                   int border, int fin, double scale) {
    long W = s->width, H = s->height, cn = s->channels, n = W * cn;
    int rx = nkx / 2, ry = nky / 2;
    int64_t *V = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!V) return ORC_OUT_OF_MEMORY;
    for (long y = band_lo(); y < band_hi(H); y++) {
        for (long c = 0; c < n; c++) {
            int64_t a = 0;
            for (int j = 0; j < nky; j++) {
                long sy = orc_border_map(y + j - ry, H, border);
                a += ky[j] * (sy < 0 ? 0 : ldi(s, sy, c));
            }
            V[c] = a;
        }
        for (long e = 0; e < n; e++) {
            long x = e / cn, ch = e % cn;
            int64_t a = 0;
            for (int i = 0; i < nkx; i++) {
                long sx = orc_border_map(x + i - rx, W, border);
                a += kx[i] * (sx < 0 ? 0 : V[sx * cn + ch]);
            }
            int64_t o;
            switch (fin) {
            case FIN_FIX8: o = (a + (1 << 15)) >> 16; if (o > 255) o = 255; break;
            case FIN_FIX16: o = (int64_t)(((uint64_t)a + (1ull << 31)) >> 32); if (o > 65535) o = 65535; break;
            case FIN_CAST: {
                double lo, hi;
                if (d->depth == 5) { st_f(d, y, e, (float)a); continue; }
                depth_range(d->depth, &lo, &hi);
                o = a < (int64_t)lo ? (int64_t)lo : a > (int64_t)hi ? (int64_t)hi : a;
                break;
            }
            default: { /* FIN_SCALED: f32(v) * f32(scale) then satV */
                float f = (float)a * (float)scale;
                if (d->depth == 5) { st_f(d, y, e, f); continue; }
                o = satv_f(f, d->depth);
                break;
            }
            }
            st_i(d, y, e, o);
        }
    }
    free(V);
    return ORC_OK;
}

/* filter.zig:213-229 */
enum { SYM_NONE, SYM_EVEN, SYM_ODD };
static int symmetry_f(const float *k, int n) {
    int even = 1, odd = 1;
    for (int i = 0; i < n / 2; i++) {
        float a = k[i], b = k[n - 1 - i];
        if (a != b) even = 0;
        if (a != -b) odd = 0;
    }
    if (n % 2 == 1 && k[n / 2] != 0) odd = 0;
    return even ? SYM_EVEN : odd ? SYM_ODD : SYM_NONE;
}

/* filter.zig:232-252 tapsV for one lane: folded by symmetry. */
static float taps_fold(const float *v, const float *k, int n, int sym) {
    int r = n / 2;
    if (sym == SYM_NONE) {
        float a = 0.0f;
        for (int j = 0; j < n; j++) a = a + v[j] * k[j];
        return a;
    }
    float a = sym == SYM_EVEN ? v[r] * k[r] : 0.0f;
    for (int i = 0; i < r; i++) {
        float t = v[i], b = v[n - 1 - i];
        float pair = sym == SYM_EVEN ? t + b : b - t;
        a = a + pair * (sym == SYM_EVEN ? k[i] : k[n - 1 - i]);
    }
    return a;
}

/* filter.zig:255-259 tapsS (scalar tail); `from_zero` is the compile-time
 * kernel tail of vfirstConst/hsecondConst (filter.zig:297-301, 310-314). */
static float taps_seq(const float *v, const float *k, int n, int from_zero) {
    float a;
    int j0;
    if (from_zero) { a = 0.0f; j0 = 0; }
    else { a = k[0] * v[0]; j0 = 1; }
    for (int j = j0; j < n; j++) a = a + k[j] * v[j];
    return a;
}

/*
 * f32 separable filter (filter.zig:168-205, 262-272 with H = V = f32). The
 * vector body (lanes of 16 f32) folds symmetric taps; the scalar tail of
 * each row (elements past the last full 16-lane vector) sums taps in order.
 */
static int sep_f32(const orc_image *s, const orc_image *d, const float *kx, int nkx, const float *ky, int nky,
                   int border, int fin, double scale, int const_kernel) {
    long W = s->width, H = s->height, cn = s->channels, n = W * cn;
    int rx = nkx / 2, ry = nky / 2;
    long body = (n / 16) * 16;
    int vsym = symmetry_f(ky, nky), hsym = symmetry_f(kx, nkx);
    if (const_kernel) vsym = hsym = SYM_EVEN;
    float *V = (float *)malloc(sizeof(float) * (size_t)(n + 2 * rx * cn));
    if (!V) return ORC_OUT_OF_MEMORY;
    float *mid = V + rx * cn;
    float tv[256];
    /* const_v: value of a constant-border pixel after the vertical pass */
    float const_v = 0.0f;
    for (int j = 0; j < nky; j++) const_v = const_v + ky[j] * 0.0f;
    long rows[256];
    for (long y = band_lo(); y < band_hi(H); y++) {
        for (int j = 0; j < nky; j++) rows[j] = orc_border_map(y + j - ry, H, border);
        for (long c = 0; c < n; c++) {
            for (int j = 0; j < nky; j++) tv[j] = rows[j] < 0 ? 0.0f : ldf(s, rows[j], c);
            mid[c] = c < body ? taps_fold(tv, ky, nky, vsym) : taps_seq(tv, ky, nky, const_kernel);
        }
        for (int i = 0; i < rx; i++) {
            long lx = orc_border_map((long)i - rx, W, border);
            long rxm = orc_border_map(W + i, W, border);
            for (long ch = 0; ch < cn; ch++) {
                V[i * cn + ch] = lx < 0 ? const_v : mid[lx * cn + ch];
                V[(rx + W + i) * cn + ch] = rxm < 0 ? const_v : mid[rxm * cn + ch];
            }
        }
        for (long e = 0; e < n; e++) {
            for (int i = 0; i < nkx; i++) tv[i] = V[e + i * cn];
            float a = e < body ? taps_fold(tv, kx, nkx, hsym) : taps_seq(tv, kx, nkx, const_kernel);
            if (fin == FIN_SCALED) a = a * (float)scale;
            if (d->depth == 5) st_f(d, y, e, a);
            else st_i(d, y, e, satv_f(a, d->depth));
        }
    }
    free(V);
    return ORC_OK;
}

static int same_shape(const orc_image *a, const orc_image *b) {
    return a->width == b->width && a->height == b->height && a->channels == b->channels;
}

/* filter.zig:498-607 (Gaussian), dispatch per capi.zig:197-206 */
int orc_gaussian_blur(const orc_image *src, const orc_image *dst, int kw, int kh, double sx, double sy, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (kw < 0 || kh < 0) return ORC_INVALID_ARGUMENT;
    if (!src) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    double sy_in = sy > 0 ? sy : sx;
    if (kw == 0) kw = orc_gaussian_ksize(sx, T == 0);
    if (kh == 0) kh = orc_gaussian_ksize(sy_in, T == 0);
    if (kw <= 0 || kh <= 0 || kw % 2 == 0 || kh % 2 == 0 || kw > 127 || kh > 127) return ORC_INVALID_ARGUMENT;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    if (border != 0) {
        if (src->width == 1) kw = 1;
        if (src->height == 1) kh = 1;
    }
    if (T == 0 || T == 2) {
        int bits = T == 0 ? 8 : 16;
        uint32_t fx[128], fy[128];
        int64_t kx[128], ky[128];
        orc_gaussian_kernel_fixed(fx, kw, sx, bits);
        orc_gaussian_kernel_fixed(fy, kh, sy_in, bits);
        for (int i = 0; i < kw; i++) kx[i] = fx[i];
        for (int i = 0; i < kh; i++) ky[i] = fy[i];
        return sep_int(src, dst, kx, kw, ky, kh, border, T == 0 ? FIN_FIX8 : FIN_FIX16, 1);
    }
    double kxd[128], kyd[128];
    float kx[128], ky[128];
    orc_gaussian_kernel(kxd, kw, sx);
    orc_gaussian_kernel(kyd, kh, sy_in);
    for (int i = 0; i < kw; i++) kx[i] = (float)kxd[i];
    for (int i = 0; i < kh; i++) ky[i] = (float)kyd[i];
    int const_kernel = 0;
    if (T == 5 && kw == kh && memcmp(kx, ky, sizeof(float) * (size_t)kw) == 0) {
        static const float t3[3] = {0.25f, 0.5f, 0.25f};
        static const float t5[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
        if ((kw == 3 && memcmp(kx, t3, sizeof t3) == 0) || (kw == 5 && memcmp(kx, t5, sizeof t5) == 0))
            const_kernel = 1;
    }
    return sep_f32(src, dst, kx, kw, ky, kh, border, FIN_CAST, 1, const_kernel);
}

/* filter.zig:867-906 (Sobel), dispatch per capi.zig:218-232 */
int orc_sobel(const orc_image *src, const orc_image *dst, int dx, int dy, int ksize, double scale, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (dx < 0 || dy < 0) return ORC_INVALID_ARGUMENT;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t S = src->depth;
    if (S != 0 && S != 2 && S != 3 && S != 5) return ORC_UNSUPPORTED;
    int rc = check(src, S);
    if (rc) return rc;
    if (dst->depth == 3) { if (S != 0) return ORC_UNSUPPORTED; }
    else if (dst->depth != 5) return ORC_UNSUPPORTED;
    if ((rc = check(dst, dst->depth))) return rc;
    if (ksize != -1 && ksize != 1 && ksize != 3 && ksize != 5 && ksize != 7) return ORC_INVALID_ARGUMENT;
    int n = ksize <= 1 ? 3 : ksize;
    int nx = (ksize == 1 && dx > 0) ? 3 : n;
    int ny = (ksize == 1 && dy > 0) ? 3 : n;
    int32_t kx[32], ky[32];
    if (orc_deriv_kernel(kx, nx, dx, (ksize == 1 && dx == 0) ? 1 : (ksize == 1 ? 3 : ksize))) return ORC_INVALID_ARGUMENT;
    if (orc_deriv_kernel(ky, ny, dy, (ksize == 1 && dy == 0) ? 1 : (ksize == 1 ? 3 : ksize))) return ORC_INVALID_ARGUMENT;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    if (dst->depth == 5 || S == 5) {
        float fx[32], fy[32];
        for (int i = 0; i < nx; i++) fx[i] = (float)kx[i];
        for (int i = 0; i < ny; i++) fy[i] = (float)ky[i];
        return sep_f32(src, dst, fx, nx, fy, ny, border, FIN_SCALED, scale, 0);
    }
    int64_t bx = 0, by = 0, lx[32], ly[32];
    for (int i = 0; i < nx; i++) { bx += labs((long)kx[i]); lx[i] = kx[i]; }
    for (int i = 0; i < ny; i++) { by += labs((long)ky[i]); ly[i] = ky[i]; }
    (void)bx; (void)by;
    /* i16 and i32 accumulators are both exact here; scale != 1 takes the
     * f32-scaled finish (filter.zig:897-905). */
    return sep_int(src, dst, lx, nx, ly, ny, border, scale == 1 ? FIN_CAST : FIN_SCALED, scale);
}

/* filter.zig:909-914 (sepFilter2D: f32 kernels, f32 accumulation, FinishCast) */
int orc_sep_filter_2d(const orc_image *src, const orc_image *dst, const float *kx, size_t nkx, const float *ky,
                      size_t nky, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t S = src->depth, D = dst->depth;
    if ((S != 0 && S != 2 && S != 3 && S != 5) || (D != 0 && D != 2 && D != 3 && D != 5)) return ORC_UNSUPPORTED;
    int rc = check(src, S);
    if (rc) return rc;
    if ((rc = check(dst, D))) return rc;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    if (!kx || !ky || nkx == 0 || nky == 0 || nkx % 2 == 0 || nky % 2 == 0 || nkx > 127 || nky > 127)
        return ORC_INVALID_ARGUMENT;
    return sep_f32(src, dst, kx, (int)nkx, ky, (int)nky, border, FIN_CAST, 1, 0);
}

/* ------------------------------------------------------ color, threshold */

/* color.zig:88-105, 120-131, 324-403 (gray codes 6, 7, 10, 11) */
int orc_cvt_color(const orc_image *src, const orc_image *dst, int code) {
    int scn, sb;
    switch (code) {
    case 6: scn = 3; sb = 0; break;
    case 7: scn = 3; sb = 2; break;
    case 10: scn = 4; sb = 0; break;
    case 11: scn = 4; sb = 2; break;
    default: return ORC_UNSUPPORTED;
    }
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    if (src->width != dst->width || src->height != dst->height) return ORC_SIZE_MISMATCH;
    if ((int)src->channels != scn || dst->channels != 1) return ORC_UNSUPPORTED;
    if (T == 3) return ORC_UNSUPPORTED;
    int bi = sb, ri = 2 - sb;
    for (long y = band_lo(); y < band_hi(src->height); y++)
        for (long x = 0; x < src->width; x++) {
            if (T == 5) {
                float b = ldf(src, y, x * scn + bi), g = ldf(src, y, x * scn + 1), r = ldf(src, y, x * scn + ri);
                st_f(dst, y, x, b * 0.114f + g * 0.587f + r * 0.299f);
            } else {
                uint32_t b = (uint32_t)ldi(src, y, x * scn + bi), g = (uint32_t)ldi(src, y, x * scn + 1),
                         r = (uint32_t)ldi(src, y, x * scn + ri);
                st_i(dst, y, x, (b * 3735u + g * 19235u + r * 9798u + (1u << 14)) >> 15);
            }
        }
    return ORC_OK;
}

/* threshold.zig:184-207 getThreshVal_Otsu_8u */
static double otsu(const uint32_t *hist, double total) {
    double scale = 1.0 / total, mu = 0, mu1 = 0, q1 = 0, max_sigma = 0, max_val = 0;
    const double eps = 2.220446049250313e-16;
    for (int i = 0; i < 256; i++) mu += (double)i * (double)hist[i];
    mu *= scale;
    for (int i = 0; i < 256; i++) {
        double p_i = (double)hist[i] * scale;
        mu1 *= q1;
        q1 += p_i;
        double q2 = 1.0 - q1;
        if ((q1 < q2 ? q1 : q2) < eps || (q1 > q2 ? q1 : q2) > 1.0 - eps) continue;
        mu1 = (mu1 + (double)i * p_i) / q1;
        double mu2 = (mu - q1 * mu1) / q2;
        double sigma = q1 * q2 * (mu1 - mu2) * (mu1 - mu2);
        if (sigma > max_sigma) { max_sigma = sigma; max_val = (double)i; }
    }
    return max_val;
}

/* threshold.zig:210-253 getThreshVal_Triangle_8u */
static double triangle(const uint32_t *hist) {
    int left = 0, right = 0, max_ind = 0, flip = 0;
    uint32_t mx = 0, h[256];
    for (int i = 0; i < 256; i++) if (hist[i] > 0) { left = i; break; }
    if (left > 0) left -= 1;
    for (int i = 255; i > 0; i--) if (hist[i] > 0) { right = i; break; }
    if (right < 255) right += 1;
    for (int i = 0; i < 256; i++) if (hist[i] > mx) { mx = hist[i]; max_ind = i; }
    for (int i = 0; i < 256; i++) h[i] = hist[i];
    if (max_ind - left < right - max_ind) {
        flip = 1;
        for (int i = 0; i < 256; i++) h[i] = hist[255 - i];
        left = 255 - right;
        max_ind = 255 - max_ind;
    }
    int thresh = left;
    double a = (double)mx, b = (double)(left - max_ind), dist = 0;
    for (int j = left + 1; j <= max_ind; j++) {
        double t = a * (double)j + b * (double)h[j];
        if (t > dist) { dist = t; thresh = j; }
    }
    thresh -= 1;
    if (flip) thresh = 255 - thresh;
    return (double)thresh;
}

/* threshold.zig:25-119; THRESH_OTSU (8) / THRESH_TRIANGLE (16) on 8U one-channel images */
/* videoio/frame.zig:167-186 grayDownsample, Kernel.pairs :292-317 (k = 2:
 * integer, (sum + 2^15) >> 16) and Kernel.blocks :319-336 (other k: the
 * weighted block sum as f32 times the f32 reciprocal of k*k*2^14, plus 0.5,
 * truncated, at most 255). BGR or BGRA (alpha unused); weights 1868, 9617,
 * 4899 (:186-189); leftover columns and rows are ignored. */
int orc_gray_downsample(const orc_image *src, const orc_image *dst, uint32_t k) {
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    int rc;
    if ((rc = check(src, 0)) || (rc = check(dst, 0))) return rc;
    if (k == 0 || k > 16 || dst->channels != 1 || (src->channels != 3 && src->channels != 4)) return ORC_SIZE_MISMATCH;
    if (dst->width != src->width / k || dst->height != src->height / k) return ORC_SIZE_MISMATCH;
    const uint32_t cn = src->channels;
    const float inv = 1.0f / (float)(k * k * (1u << 14));
    for (uint32_t oy = (uint32_t)band_lo(); oy < (uint32_t)band_hi(dst->height); oy++)
        for (uint32_t ox = 0; ox < dst->width; ox++) {
            uint32_t sb = 0, sg = 0, sr = 0;
            for (uint32_t dy = 0; dy < k; dy++) {
                const uint8_t *row = (const uint8_t *)src->data + (size_t)(oy * k + dy) * src->stride;
                for (uint32_t dx = 0; dx < k; dx++) {
                    const uint8_t *p = row + (size_t)(ox * k + dx) * cn;
                    sb += p[0];
                    sg += p[1];
                    sr += p[2];
                }
            }
            const uint32_t l = sb * 1868u + sg * 9617u + sr * 4899u;
            uint8_t o;
            if (k == 2) {
                o = (uint8_t)((l + (1u << 15)) >> 16);
            } else {
                const float v = (float)l * inv + 0.5f;
                o = (uint8_t)(v < 255.0f ? v : 255.0f);
            }
            ((uint8_t *)dst->data)[(size_t)oy * dst->stride + ox] = o;
        }
    return ORC_OK;
}

/* frame.zig:191-199 grayDownsampleRect: the (x, y, w, h) window of dst from
 * the source blocks under it. */
int orc_gray_downsample_rect(const orc_image *src, const orc_image *dst, uint32_t k, uint32_t x, uint32_t y, uint32_t w,
                             uint32_t h) {
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    int rc;
    if ((rc = check(src, 0)) || (rc = check(dst, 0))) return rc;
    if (k == 0 || k > 16 || dst->channels != 1 || (src->channels != 3 && src->channels != 4)) return ORC_SIZE_MISMATCH;
    if (dst->width != src->width / k || dst->height != src->height / k) return ORC_SIZE_MISMATCH;
    if (w == 0 || h == 0) return ORC_OK;
    if ((uint64_t)x + w > dst->width || (uint64_t)y + h > dst->height) return ORC_INVALID_ARGUMENT;
    orc_image d = *dst, s = *src;
    d.data = (uint8_t *)dst->data + (size_t)y * dst->stride + x;
    d.width = w;
    d.height = h;
    s.data = (uint8_t *)src->data + (size_t)y * k * src->stride + (size_t)x * k * src->channels;
    s.width = w * k;
    s.height = h * k;
    return orc_gray_downsample(&s, &d, k);
}

int orc_threshold(const orc_image *src, const orc_image *dst, double thresh, double maxval, int type, double *used) {
    int kind = type & 7;
    if (kind > 4) return ORC_UNSUPPORTED;
    int autom = (type & 8) ? 1 : (type & 16) ? 2 : 0;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    if (autom) {
        if (T != 0 || src->channels != 1) return ORC_UNSUPPORTED;
        uint32_t hist[256] = {0};
        for (long y = 0; y < src->height; y++)
            for (long x = 0; x < src->width; x++) hist[ldi(src, y, x)]++;
        thresh = autom == 1 ? otsu(hist, (double)src->width * src->height) : triangle(hist);
    }
    if (used) *used = thresh;
    long n = (long)src->width * src->channels;
    if (T == 0) {
        double fl = floor(thresh);
        long ti = fl < -2147483648.0 ? -2147483647L - 1 : fl > 2147483647.0 ? 2147483647L : (long)fl;
        int m = (int)sat_cast_f(maxval, 0);
        int t = ti < 0 ? 0 : ti > 255 ? 255 : (int)ti;
        uint8_t lut[256];
        for (int i = 0; i < 256; i++) {
            int gt = (ti < 0 || ti >= 255) ? (i > ti) : (i > t);
            int v = i;
            switch (kind) {
            case 0: v = gt ? m : 0; break;
            case 1: v = gt ? 0 : m; break;
            case 2: v = gt ? ((ti < 0 || ti >= 255) ? (int)(ti < 0 ? 0 : 255) : t) : i; break;
            case 3: v = gt ? i : 0; break;
            default: v = gt ? 0 : i; break;
            }
            lut[i] = (uint8_t)v;
        }
        for (long y = 0; y < src->height; y++)
            for (long i = 0; i < n; i++) st_i(dst, y, i, lut[ldi(src, y, i)]);
        return ORC_OK;
    }
    if (T == 5) {
        float t = (float)thresh;
        float m = (float)maxval;
        for (long y = 0; y < src->height; y++)
            for (long i = 0; i < n; i++) {
                float v = ldf(src, y, i), o;
                int gt = v > t;
                switch (kind) {
                case 0: o = gt ? m : 0.0f; break;
                case 1: o = gt ? 0.0f : m; break;
                case 2: o = gt ? t : v; break;
                case 3: o = gt ? v : 0.0f; break;
                default: o = gt ? 0.0f : v; break;
                }
                st_f(dst, y, i, o);
            }
        return ORC_OK;
    }
    double lo, hi;
    depth_range(T, &lo, &hi);
    int64_t m = sat_cast_f(maxval, T);
    if (floor(thresh) < lo) {
        for (long y = 0; y < src->height; y++)
            for (long i = 0; i < n; i++) {
                int64_t v = ldi(src, y, i);
                st_i(dst, y, i, kind == 3 ? v : kind == 0 ? m : 0);
            }
        return ORC_OK;
    }
    int64_t t = sat_cast_f(floor(thresh), T);
    for (long y = 0; y < src->height; y++)
        for (long i = 0; i < n; i++) {
            int64_t v = ldi(src, y, i), o;
            int gt = v > t;
            switch (kind) {
            case 0: o = gt ? m : 0; break;
            case 1: o = gt ? 0 : m; break;
            case 2: o = gt ? t : v; break;
            case 3: o = gt ? v : 0; break;
            default: o = gt ? 0 : v; break;
            }
            st_i(dst, y, i, o);
        }
    return ORC_OK;
}

/* ------------------------------------------------------------- resize */

static void store_round(const orc_image *d, long y, long i, float v);


/* resize.zig:186-214 linearTapU8 */
static void linear_tap(long d, uint32_t ssize, uint32_t dsize_, int area, int pin, long o[2], int w[2]) {
    double inv = (double)dsize_ / (double)ssize;
    double scale = 1.0 / inv;
    double df = (double)d;
    long sx;
    float fx;
    if (area) {
        sx = (long)floor(df * scale);
        fx = (float)((df + 1) - (double)(sx + 1) * inv);
        fx = fx <= 0 ? 0.0f : fx - floorf(fx);
    } else {
        fx = (float)((df + 0.5) * scale - 0.5);
        float fl = floorf(fx);
        sx = (long)fl;
        fx -= fl;
    }
    long last = (long)ssize - 1;
    if (pin && sx < 0) { sx = 0; fx = 0; }
    if (pin && sx >= last) { sx = last; fx = 0; }
    w[0] = (int)sat_cast_f((double)((1.0f - fx) * 2048.0f), 3);
    w[1] = (int)sat_cast_f((double)(fx * 2048.0f), 3);
    o[0] = sx < 0 ? 0 : sx > last ? last : sx;
    o[1] = sx + 1 < 0 ? 0 : sx + 1 > last ? last : sx + 1;
}

/* resize.zig:216-242 + resize_u8.zig:641-711: fixed-point bilinear */
static void linear_u8(const orc_image *s, const orc_image *d, int area) {
    long cn = s->channels;
    for (long y = band_lo(); y < band_hi(d->height); y++) {
        long yo[2];
        int b[2];
        linear_tap(y, s->height, d->height, area, 0, yo, b);
        for (long x = 0; x < d->width; x++) {
            long xo[2];
            int w[2];
            linear_tap(x, s->width, d->width, area, 1, xo, w);
            for (long ch = 0; ch < cn; ch++) {
                int32_t h[2];
                for (int k = 0; k < 2; k++) {
                    int32_t acc = (int32_t)ldi(s, yo[k], xo[0] * cn + ch) * w[0] + (int32_t)ldi(s, yo[k], xo[1] * cn + ch) * w[1];
                    h[k] = (int16_t)(acc >> 4);
                }
                int32_t t = ((h[0] * b[0]) >> 16) + ((h[1] * b[1]) >> 16);
                int32_t o = (t + 2) >> 2;
                st_i(d, y, x * cn + ch, o < 0 ? 0 : o > 255 ? 255 : o);
            }
        }
    }
}

/* resize_u8.zig:381-447 exact half */
static void half_u8(const orc_image *s, const orc_image *d) {
    long cn = s->channels;
    for (long y = band_lo(); y < band_hi(d->height); y++)
        for (long x = 0; x < d->width; x++)
            for (long ch = 0; ch < cn; ch++) {
                long i = 2 * x * cn + ch;
                uint32_t v = (uint32_t)(ldi(s, 2 * y, i) + ldi(s, 2 * y, i + cn) + ldi(s, 2 * y + 1, i) + ldi(s, 2 * y + 1, i + cn));
                st_i(d, y, x * cn + ch, (v + 2) >> 2);
            }
}

/* resize_u8.zig:1213-1277 integer-factor area */
static void area_int_u8(const orc_image *s, const orc_image *d, long ix, long iy) {
    long cn = s->channels;
    float scale = 1.0f / (float)(ix * iy);
    for (long y = band_lo(); y < band_hi(d->height); y++)
        for (long x = 0; x < d->width; x++)
            for (long ch = 0; ch < cn; ch++) {
                uint32_t sum = 0;
                for (long dy = 0; dy < iy; dy++)
                    for (long dx = 0; dx < ix; dx++) sum += (uint32_t)ldi(s, y * iy + dy, (x * ix + dx) * cn + ch);
                st_i(d, y, x * cn + ch, satf32((float)sum * scale, 0));
            }
}

typedef struct { uint32_t di, si; float alpha; } area_tab;

/* resize.zig:470-493 computeResizeAreaTab */
static long area_tab_build(area_tab *out, uint32_t ssize, uint32_t dsz, double scale) {
    long k = 0;
    for (long dx = 0; dx < dsz; dx++) {
        double fsx1 = (double)dx * scale;
        double fsx2 = fsx1 + scale;
        double cellw = scale < (double)ssize - fsx1 ? scale : (double)ssize - fsx1;
        long sx1 = (long)ceil(fsx1);
        long sx2 = (long)floor(fsx2);
        long sx2c = sx2 < (long)ssize - 1 ? sx2 : (long)ssize - 1;
        if (sx1 > (long)ssize - 1) sx1 = (long)ssize - 1;
        if ((double)sx1 - fsx1 > 1e-3) {
            out[k].di = (uint32_t)dx; out[k].si = (uint32_t)(sx1 - 1); out[k].alpha = (float)(((double)sx1 - fsx1) / cellw); k++;
        }
        for (long sx = sx1; sx < sx2; sx++) {
            out[k].di = (uint32_t)dx; out[k].si = (uint32_t)sx; out[k].alpha = (float)(1.0 / cellw); k++;
        }
        if (fsx2 - (double)sx2 > 1e-3) {
            double a = fsx2 - (double)sx2;
            if (a > 1.0) a = 1.0;
            if (a > cellw) a = cellw;
            out[k].di = (uint32_t)dx; out[k].si = (uint32_t)sx2c; out[k].alpha = (float)(a / cellw); k++;
        }
    }
    return k;
}

/* resize.zig:527-565 areaGeneric (f32 table-order sums, no FMA, then
 * sat.cast to the sample type), any depth; the u8 kernels of
 * resize_u8.zig:1058-1208 reproduce it bit for bit. */
static int area_generic(const orc_image *s, const orc_image *d) {
    long cn = s->channels;
    double sxs = (double)s->width / (double)d->width, sys = (double)s->height / (double)d->height;
    area_tab *xt = (area_tab *)malloc(sizeof(area_tab) * (2 * (size_t)s->width + 2 * (size_t)d->width + 4));
    area_tab *yt = (area_tab *)malloc(sizeof(area_tab) * (2 * (size_t)s->height + 2 * (size_t)d->height + 4));
    long nd = (long)d->width * cn;
    float *hbuf = (float *)malloc(sizeof(float) * (size_t)nd), *sum = (float *)malloc(sizeof(float) * (size_t)nd);
    if (!xt || !yt || !hbuf || !sum) { free(xt); free(yt); free(hbuf); free(sum); return ORC_OUT_OF_MEMORY; }
    long nx = area_tab_build(xt, s->width, d->width, sxs);
    long ny = area_tab_build(yt, s->height, d->height, sys);
    long k = 0;
    while (k < ny && (long)yt[k].di < band_lo()) k++;
    for (long dy = band_lo(); dy < band_hi(d->height); dy++) {
        for (long i = 0; i < nd; i++) sum[i] = 0.0f;
        for (; k < ny && yt[k].di == (uint32_t)dy; k++) {
            for (long i = 0; i < nd; i++) hbuf[i] = 0.0f;
            for (long e = 0; e < nx; e++)
                for (long ch = 0; ch < cn; ch++)
                    hbuf[xt[e].di * cn + ch] = hbuf[xt[e].di * cn + ch] + ldf(s, yt[k].si, xt[e].si * cn + ch) * xt[e].alpha;
            for (long i = 0; i < nd; i++) sum[i] = sum[i] + hbuf[i] * yt[k].alpha;
        }
        for (long i = 0; i < nd; i++) store_round(d, dy, i, sum[i]);
    }
    free(xt); free(yt); free(hbuf); free(sum);
    return ORC_OK;
}

/* resize.zig:62-106 nearest (any depth) */
static void nearest_any(const orc_image *s, const orc_image *d) {
    long cn = s->channels;
    double ifx = (double)s->width / (double)d->width, ify = (double)s->height / (double)d->height;
    size_t es = dsize(s->depth);
    for (long y = band_lo(); y < band_hi(d->height); y++) {
        long sy = (long)floor((double)y * ify);
        if (sy > (long)s->height - 1) sy = (long)s->height - 1;
        for (long x = 0; x < d->width; x++) {
            long sx = (long)floor((double)x * ifx);
            if (sx > (long)s->width - 1) sx = (long)s->width - 1;
            memcpy((char *)rowp(d, y) + (size_t)x * cn * es, (char *)rowp(s, sy) + (size_t)sx * cn * es, (size_t)cn * es);
        }
    }
}

/* resize.zig:154-180 linearCoord: source index and fractional weight, half-
 * pixel centres (or the area-mode coordinate when enlarging), clamped to
 * the edges with weight 0. */
static void linear_coord(long d, double scale, uint32_t len, int area, long *idx, double *frac) {
    long sx;
    double f;
    if (area) {
        const double df = (double)d;
        sx = (long)floor(df * scale);
        f = (df + 1) - (double)(sx + 1) / scale;
        f = f <= 0 ? 0 : f - floor(f);
    } else {
        const double fx = ((double)d + 0.5) * scale - 0.5;
        sx = (long)floor(fx);
        f = fx - (double)sx;
    }
    if (sx < 0) { sx = 0; f = 0; }
    if (sx >= (long)len - 1) { sx = (long)len - 1; f = 0; }
    *idx = sx;
    *frac = f;
}

/* resize.zig:284-370 linear for u16 / i16 / f32 (Weights of :401-418): f32
 * weights (f32(1 - f), f32(f)), each sample converted to f32, horizontal
 * sums s0 a0 + s1 a1 and the vertical r0 b0 + r1 b1 unfused, then
 * sat.cast to the sample type. x and y both clamp to the edges. */
static int linear_any(const orc_image *s, const orc_image *d, int area) {
    const long cn = s->channels, W = s->width, H = s->height;
    const double sxs = (double)s->width / (double)d->width, sys = (double)s->height / (double)d->height;
    long *xi = (long *)malloc(sizeof(long) * (size_t)d->width);
    float *xa = (float *)malloc(sizeof(float) * 2 * (size_t)d->width);
    if (!xi || !xa) { free(xi); free(xa); return ORC_OUT_OF_MEMORY; }
    for (long x = 0; x < d->width; x++) {
        long i;
        double f;
        linear_coord(x, sxs, s->width, area, &i, &f);
        xi[x] = i;
        xa[2 * x] = (float)(1.0 - f);
        xa[2 * x + 1] = (float)f;
    }
    const long last = (W - 1) * cn;
    for (long y = band_lo(); y < band_hi(d->height); y++) {
        long iy;
        double fy;
        linear_coord(y, sys, s->height, area, &iy, &fy);
        const long y0 = iy, y1 = iy + 1 < H ? iy + 1 : H - 1;
        const float b0 = (float)(1.0 - fy), b1 = (float)fy;
        for (long x = 0; x < d->width; x++)
            for (long ch = 0; ch < cn; ch++) {
                const long i0 = xi[x] * cn + ch, i1 = i0 >= last ? i0 : i0 + cn;
                const float a0 = xa[2 * x], a1 = xa[2 * x + 1];
                const float h0 = ldf(s, y0, i0) * a0 + ldf(s, y0, i1) * a1;
                const float h1 = ldf(s, y1, i0) * a0 + ldf(s, y1, i1) * a1;
                store_round(d, y, x * cn + ch, h0 * b0 + h1 * b1);
            }
    }
    free(xi);
    free(xa);
    return ORC_OK;
}

/* resize.zig:420-457 areaInt for any depth: u8 2x2 is (a + b + c + d + 2) >> 2;
 * other unsigned samples the integer block mean rounded half up; signed
 * and float samples the f64 mean, sat.cast to the sample type. */
static void area_int_any(const orc_image *s, const orc_image *d, long ix, long iy) {
    const long cn = s->channels;
    const uint32_t T = s->depth;
    const long area = ix * iy;
    for (long y = band_lo(); y < band_hi(d->height); y++)
        for (long x = 0; x < d->width; x++)
            for (long ch = 0; ch < cn; ch++) {
                const long o = x * cn + ch;
                if (T == 0 && ix == 2 && iy == 2) {
                    const long i = 2 * x * cn + ch;
                    const int64_t v = ldi(s, 2 * y, i) + ldi(s, 2 * y, i + cn) + ldi(s, 2 * y + 1, i) + ldi(s, 2 * y + 1, i + cn);
                    st_i(d, y, o, (v + 2) >> 2);
                } else if (T == 3 || T == 5) {
                    double sum = 0;
                    for (long dy = 0; dy < iy; dy++)
                        for (long dx = 0; dx < ix; dx++) sum += (double)ldf(s, y * iy + dy, (x * ix + dx) * cn + ch);
                    const double m = sum / (double)area;
                    if (T == 5) st_f(d, y, o, (float)m);
                    else st_i(d, y, o, sat_cast_f(m, T));
                } else {
                    uint64_t sum = 0;
                    for (long dy = 0; dy < iy; dy++)
                        for (long dx = 0; dx < ix; dx++) sum += (uint64_t)ldi(s, y * iy + dy, (x * ix + dx) * cn + ch);
                    st_i(d, y, o, (int64_t)((sum + (uint64_t)area / 2) / (uint64_t)area));
                }
            }
}

/* resize.zig:28-62 dispatch: u8 with 1-4 channels takes the fixed-point fast
 * paths; other depths and channel counts the generic kernels of :47-60. */
int orc_resize(const orc_image *src, const orc_image *dst, int interp) {
    if (interp < 0 || interp > 4) return ORC_UNSUPPORTED;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    if (src->channels != dst->channels) return ORC_SIZE_MISMATCH;
    if (src->width == dst->width && src->height == dst->height) {
        size_t rl = (size_t)src->width * src->channels * dsize(T);
        for (long y = 0; y < src->height; y++) memcpy(rowp(dst, y), rowp(src, y), rl);
        return ORC_OK;
    }
    if (interp == 0) { nearest_any(src, dst); return ORC_OK; }
    if (interp != 1 && interp != 3) return ORC_UNSUPPORTED;
    if (T != 0 || src->channels > 4) {
        /* resize.zig:47-60: the generic kernels (u8 keeps its 11-bit taps). */
        const int up = dst->width > src->width || dst->height > src->height;
        if (interp == 1 || up) {
            if (T == 0) { linear_u8(src, dst, interp == 3); return ORC_OK; }
            return linear_any(src, dst, interp == 3);
        }
        long ix = src->width / dst->width, iy = src->height / dst->height;
        if (ix * dst->width == src->width && iy * dst->height == src->height) { area_int_any(src, dst, ix, iy); return ORC_OK; }
        return area_generic(src, dst);
    }
    int exact_half = src->width == 2 * dst->width && src->height == 2 * dst->height;
    if (interp == 1) {
        if (exact_half) half_u8(src, dst);
        else linear_u8(src, dst, 0);
        return ORC_OK;
    }
    if (dst->width > src->width || dst->height > src->height) { linear_u8(src, dst, 1); return ORC_OK; }
    if (exact_half) { half_u8(src, dst); return ORC_OK; }
    long ix = src->width / dst->width, iy = src->height / dst->height;
    if (ix * dst->width == src->width && iy * dst->height == src->height) { area_int_u8(src, dst, ix, iy); return ORC_OK; }
    return area_generic(src, dst);
}

/* Test hooks: the generic u8 kernels whatever the dispatch would choose, so
 * tests can hold them to the closed forms the reference's own tests tie them
 * to (resize.zig:654-680 fast paths == generic LINEAR, resize.zig:703-722
 * area kernels == areaGeneric), and the tap tables themselves. */
int orc_resize_linear_generic(const orc_image *src, const orc_image *dst) {
    if (!src || !dst || src->depth != 0 || dst->depth != 0 || src->channels != dst->channels) return ORC_INVALID_ARGUMENT;
    int rc = check(src, 0);
    if (rc || (rc = check(dst, 0))) return rc;
    linear_u8(src, dst, 0);
    return ORC_OK;
}

int orc_resize_area_generic(const orc_image *src, const orc_image *dst) {
    if (!src || !dst || src->depth != 0 || dst->depth != 0 || src->channels != dst->channels) return ORC_INVALID_ARGUMENT;
    if (dst->width > src->width || dst->height > src->height) return ORC_INVALID_ARGUMENT;
    int rc = check(src, 0);
    if (rc || (rc = check(dst, 0))) return rc;
    return area_generic(src, dst);
}

long orc_area_tab(uint32_t ssize, uint32_t dsz, uint32_t *di, uint32_t *si, float *alpha, long cap) {
    if (dsz == 0 || ssize == 0 || dsz > ssize) return -1;
    area_tab *t = (area_tab *)malloc(sizeof(area_tab) * (2 * (size_t)ssize + 2 * (size_t)dsz + 4));
    if (!t) return -1;
    long n = area_tab_build(t, ssize, dsz, (double)ssize / (double)dsz);
    for (long k = 0; k < n && k < cap; k++) { di[k] = t[k].di; si[k] = t[k].si; alpha[k] = t[k].alpha; }
    free(t);
    return n;
}

void orc_linear_tap(long d, uint32_t ssize, uint32_t dsz, int area, int pin, long o[2], int w[2]) {
    linear_tap(d, ssize, dsz, area, pin, o, w);
}

/* -------------------------------------------------------------- warps */

static float toF_any(const orc_image *s, long y, long i) { return ldf(s, y, i); }

static void store_round(const orc_image *d, long y, long i, float v) {
    if (d->depth == 5) st_f(d, y, i, v);
    else st_i(d, y, i, sat_cast_f((double)v, d->depth));
}

/* geometry.zig:591-602 */
static int invert_affine(const double m[6], double o[6]) {
    double det = m[0] * m[4] - m[1] * m[3];
    if (det == 0) return 0;
    double d = 1.0 / det;
    double a11 = m[4] * d, a22 = m[0] * d, a12 = -m[1] * d, a21 = -m[3] * d;
    double b1 = -a11 * m[2] - a12 * m[5];
    double b2 = -a21 * m[2] - a22 * m[5];
    o[0] = a11; o[1] = a12; o[2] = b1; o[3] = a21; o[4] = a22; o[5] = b2;
    return 1;
}

/* geometry.zig:604-613 */
static int invert3(const double m[9], double o[9]) {
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (det == 0) return 0;
    double d = 1.0 / det;
    o[0] = (m[4] * m[8] - m[5] * m[7]) * d; o[1] = (m[2] * m[7] - m[1] * m[8]) * d; o[2] = (m[1] * m[5] - m[2] * m[4]) * d;
    o[3] = (m[5] * m[6] - m[3] * m[8]) * d; o[4] = (m[0] * m[8] - m[2] * m[6]) * d; o[5] = (m[2] * m[3] - m[0] * m[5]) * d;
    o[6] = (m[3] * m[7] - m[4] * m[6]) * d; o[7] = (m[1] * m[6] - m[0] * m[7]) * d; o[8] = (m[0] * m[4] - m[1] * m[3]) * d;
    return 1;
}

/* geometry.zig:492-494 */
static long to_idx(float f) { return (long)fminn(fmaxn(f, -1.0e9f), 1.0e9f); }

static int warp_common_check(const orc_image *src, const orc_image *dst, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    return ORC_OK;
}

/* geometry.zig:410-533 warpAffine (warpRows / linearPixel / nearestPixel;
 * the 16-pixel fast blocks compute the same values). */
int orc_warp_affine(const orc_image *src, const orc_image *dst, const double m[6], int interp, int border, double value) {
    int rc = warp_common_check(src, dst, border);
    if (rc) return rc;
    double inv[6];
    if (!invert_affine(m, inv)) return ORC_UNSUPPORTED;
    if (src->channels != dst->channels) return ORC_SIZE_MISMATCH;
    float mf[6];
    for (int i = 0; i < 6; i++) mf[i] = (float)inv[i];
    long W = src->width, H = src->height, cn = src->channels;
    int T = (int)src->depth;
    double cvd = T == 5 ? (double)(float)value : (double)sat_cast_f(value, (uint32_t)T);
    float cv = (float)cvd;
    for (long y = band_lo(); y < band_hi(dst->height); y++) {
        float yf = (float)y;
        float bx = mf[1] * yf + mf[2];
        float by = mf[4] * yf + mf[5];
        for (long x = 0; x < dst->width; x++) {
            float xf = (float)x;
            float sx = fmaf(mf[0], xf, bx);
            float sy = fmaf(mf[3], xf, by);
            if (interp == 0) {
                long ix = to_idx(rint32(fminn(fmaxn(sx, -4.0e6f), 4.0e6f)));
                long iy = to_idx(rint32(fminn(fmaxn(sy, -4.0e6f), 4.0e6f)));
                long mx = orc_border_map(ix, W, border), my = orc_border_map(iy, H, border);
                for (long ch = 0; ch < cn; ch++) {
                    if (mx >= 0 && my >= 0) {
                        if (T == 5) st_f(dst, y, x * cn + ch, ldf(src, my, mx * cn + ch));
                        else st_i(dst, y, x * cn + ch, ldi(src, my, mx * cn + ch));
                    } else store_round(dst, y, x * cn + ch, cv);
                }
                continue;
            }
            float fx = floorf(sx), fy = floorf(sy);
            float ax = sx - fx, ay = sy - fy;
            long ix = to_idx(fx), iy = to_idx(fy);
            if (border == 0 && (ix < -1 || iy < -1 || ix >= W || iy >= H)) {
                for (long ch = 0; ch < cn; ch++) store_round(dst, y, x * cn + ch, cv);
                continue;
            }
            long xs[2] = {orc_border_map(ix, W, border), orc_border_map(ix + 1, W, border)};
            long ys[2] = {orc_border_map(iy, H, border), orc_border_map(iy + 1, H, border)};
            for (long ch = 0; ch < cn; ch++) {
                float p[4];
                for (int k = 0; k < 4; k++) {
                    long px = xs[k & 1], py = ys[k >> 1];
                    p[k] = (px >= 0 && py >= 0) ? toF_any(src, py, px * cn + ch) : cv;
                }
                float v0 = fmaf(ax, p[1] - p[0], p[0]);
                float v1 = fmaf(ax, p[3] - p[2], p[2]);
                store_round(dst, y, x * cn + ch, fmaf(ay, v1 - v0, v0));
            }
        }
    }
    return ORC_OK;
}

/* std.math.clamp(v, lo, hi) = @max(lo, @min(v, hi)) with NaN-ignoring min/max */
static double clampd(double v, double lo, double hi) {
    double a = isnan(v) ? hi : (v < hi ? v : hi);
    return isnan(a) ? lo : (a > lo ? a : lo);
}

/* geometry.zig:586-589, 615-672 warpPerspective (f64 coordinates) */
int orc_warp_perspective(const orc_image *src, const orc_image *dst, const double m[9], int interp, int border,
                         double value) {
    int rc = warp_common_check(src, dst, border);
    if (rc) return rc;
    double inv[9];
    if (!invert3(m, inv)) return ORC_UNSUPPORTED;
    if (src->channels != dst->channels) return ORC_SIZE_MISMATCH;
    long W = src->width, H = src->height, cn = src->channels;
    int T = (int)src->depth;
    double cvd = T == 5 ? (double)(float)value : (double)sat_cast_f(value, (uint32_t)T);
    for (long y = band_lo(); y < band_hi(dst->height); y++) {
        double fy = (double)y;
        for (long x = 0; x < dst->width; x++) {
            double fx = (double)x;
            double sx = inv[0] * fx + inv[1] * fy + inv[2];
            double sy = inv[3] * fx + inv[4] * fy + inv[5];
            double w = inv[6] * fx + inv[7] * fy + inv[8];
            double iw = w != 0 ? 1.0 / w : 0;
            sx *= iw;
            sy *= iw;
            if (interp == 0) {
                long ix = (long)round(clampd(sx, -1e9, 1e9));
                long iy = (long)round(clampd(sy, -1e9, 1e9));
                long mx = orc_border_map(ix, W, border), my = orc_border_map(iy, H, border);
                for (long ch = 0; ch < cn; ch++) {
                    if (mx >= 0 && my >= 0) {
                        if (T == 5) st_f(dst, y, x * cn + ch, ldf(src, my, mx * cn + ch));
                        else st_i(dst, y, x * cn + ch, ldi(src, my, mx * cn + ch));
                    } else {
                        if (T == 5) st_f(dst, y, x * cn + ch, (float)cvd);
                        else st_i(dst, y, x * cn + ch, (int64_t)cvd);
                    }
                }
                continue;
            }
            double x0f = floor(clampd(sx, -1e9, 1e9));
            double y0f = floor(clampd(sy, -1e9, 1e9));
            float ax = (float)(sx - x0f), ay = (float)(sy - y0f);
            long ix = (long)x0f, iy = (long)y0f;
            if (border == 0 && (ix < -1 || iy < -1 || ix >= W || iy >= H)) {
                for (long ch = 0; ch < cn; ch++) {
                    if (T == 5) st_f(dst, y, x * cn + ch, (float)cvd);
                    else st_i(dst, y, x * cn + ch, (int64_t)cvd);
                }
                continue;
            }
            long xs[2] = {orc_border_map(ix, W, border), orc_border_map(ix + 1, W, border)};
            long ys[2] = {orc_border_map(iy, H, border), orc_border_map(iy + 1, H, border)};
            float wt[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
            for (long ch = 0; ch < cn; ch++) {
                float acc = 0;
                for (int k = 0; k < 4; k++) {
                    long px = xs[k & 1], py = ys[k >> 1];
                    float v = (px >= 0 && py >= 0) ? toF_any(src, py, px * cn + ch) : (float)cvd;
                    acc += v * wt[k];
                }
                store_round(dst, y, x * cn + ch, acc);
            }
        }
    }
    return ORC_OK;
}

/* undistort.zig:280-394 remap (u8, f32 maps, constant border) and
 * capi_calib3d.zig:205-213 (border value u8(clamp(round(value)))) */
int orc_remap(const orc_image *src, const orc_image *dst, const orc_image *mapx, const orc_image *mapy, double value) {
    int rc;
    if ((rc = check(src, 0))) return rc;
    if ((rc = check(dst, 0))) return rc;
    if ((rc = check(mapx, 5))) return rc;
    if ((rc = check(mapy, 5))) return rc;
    if (dst->width != mapx->width || dst->height != mapx->height || mapx->width != mapy->width ||
        mapx->height != mapy->height || mapx->channels != mapy->channels || mapx->channels != 1)
        return ORC_SIZE_MISMATCH;
    if (src->channels != dst->channels || src->channels < 1 || src->channels > 4) return ORC_UNSUPPORTED;
    double rv = round(clampd(value, 0, 255));
    int bv = (int)rv;
    long W = src->width, H = src->height, cn = src->channels;
    for (long y = band_lo(); y < band_hi(dst->height); y++)
        for (long x = 0; x < dst->width; x++) {
            float sx = ldf(mapx, y, x), sy = ldf(mapy, y, x);
            float fx = floorf(sx), fy = floorf(sy);
            float ax = sx - fx, ay = sy - fy;
            if (!(fx >= -1 && fy >= -1 && fx < (float)W && fy < (float)H)) {
                for (long ch = 0; ch < cn; ch++) st_i(dst, y, x * cn + ch, bv);
                continue;
            }
            long ix = (long)fx, iy = (long)fy;
            for (long ch = 0; ch < cn; ch++) {
                float p[4];
                for (int q = 0; q < 4; q++) {
                    long px = ix + (q & 1), py = iy + (q >> 1);
                    p[q] = (px >= 0 && px < W && py >= 0 && py < H) ? (float)ldi(src, py, px * cn + ch) : (float)bv;
                }
                float v0 = fmaf(ax, p[1] - p[0], p[0]);
                float v1 = fmaf(ax, p[3] - p[2], p[2]);
                float v = fmaf(ay, v1 - v0, v0);
                float r = (v + 12582912.0f) - 12582912.0f;
                st_i(dst, y, x * cn + ch, (int64_t)fminn(fmaxn(r, 0.0f), 255.0f));
            }
        }
    return ORC_OK;
}

/* --------------------------------------------------------- morphology */

/* morph.zig:32-59 getStructuringElement */
int orc_structuring_element(int shape, int w, int h, uint8_t *buf) {
    int ax = w / 2, ay = h / 2, r = ay, c = ax;
    double inv_r2 = r > 0 ? 1.0 / (double)(r * r) : 0;
    memset(buf, 0, (size_t)w * (size_t)h);
    for (int i = 0; i < h; i++) {
        int j1 = 0, j2 = 0;
        if (shape == 0 || (shape == 1 && i == ay)) j2 = w;
        else if (shape == 1) { j1 = ax; j2 = j1 + 1; }
        else {
            long dy = (long)i - r;
            if (labs(dy) <= r) {
                double dxf = (double)c * sqrt((double)((long)r * r - dy * dy) * inv_r2);
                long dx = (long)round(dxf);
                j1 = (int)((long)c - dx > 0 ? (long)c - dx : 0);
                j2 = (int)((long)c + dx + 1 < (long)w ? (long)c + dx + 1 : (long)w);
            }
        }
        for (int j = j1; j < j2; j++) buf[i * w + j] = 1;
    }
    return ORC_OK;
}

/* morph.zig:489-519 generalPass: the per-tap definition (out-of-image taps
 * are ignored; identity = max for erode, min for dilate). */
int orc_morph_pass(const orc_image *s, const orc_image *d, int op_max, const uint8_t *mask, int kw, int kh, int ax,
                   int ay) {
    long W = s->width, H = s->height, cn = s->channels;
    int T = (int)s->depth;
    double lo, hi;
    depth_range((uint32_t)T, &lo, &hi);
    double init = T == 5 ? (op_max ? -INFINITY : INFINITY) : (op_max ? lo : hi);
    for (long y = band_lo(); y < band_hi(H); y++)
        for (long x = 0; x < W; x++)
            for (long ch = 0; ch < cn; ch++) {
                double acc = init;
                for (int ky = 0; ky < kh; ky++) {
                    long sy = y + ky - ay;
                    if (sy < 0 || sy >= H) continue;
                    for (int kx = 0; kx < kw; kx++) {
                        if (!mask[ky * kw + kx]) continue;
                        long sx = x + kx - ax;
                        if (sx < 0 || sx >= W) continue;
                        double v = ldd(s, sy, sx * cn + ch);
                        acc = op_max ? (v > acc ? v : acc) : (v < acc ? v : acc);
                    }
                }
                if (T == 5) st_f(d, y, x * cn + ch, (float)acc);
                else st_i(d, y, x * cn + ch, (int64_t)acc);
            }
    return ORC_OK;
}

/* ------------------------------------------------- box filter, pyramids */

/* capi.zig:208-217 visar_box_filter -> filter.zig:617-633 boxFilter: unit
 * taps, exact integer sums for u8 and u16 (boxU8's running and prefix sums,
 * :638-806, give the same integers), f32 through the separable engine's
 * order; then f32(sum) * f32(scale) saturated (FinishScaled), scale
 * 1 / (kw * kh) or 1. i16 is rejected: the reference widens i16 samples
 * into u32 accumulators with @intCast, undefined for negative samples. */
int orc_box_filter(const orc_image *src, const orc_image *dst, int kw, int kh, int normalize, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (kw <= 0 || kh <= 0) return ORC_INVALID_ARGUMENT;
    if (!src) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc;
    if ((rc = check(src, T)) || (rc = check(dst, T))) return rc;
    if (kw > 255 || kh > 255 || kw % 2 == 0 || kh % 2 == 0) return ORC_INVALID_ARGUMENT;
    /* i16: exact signed sums (the reference accumulates in u32, which agrees
     * wherever the samples are non-negative) */
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    const double scale = normalize ? 1.0 / (double)(kw * kh) : 1.0;
    if (T == 5) {
        float ones[255];
        for (int i = 0; i < 255; i++) ones[i] = 1.0f;
        return sep_f32(src, dst, ones, kw, ones, kh, border, FIN_SCALED, scale, 0);
    }
    int64_t ones[255];
    for (int i = 0; i < 255; i++) ones[i] = 1;
    return sep_int(src, dst, ones, kw, ones, kh, border, FIN_SCALED, scale);
}

/* The integer finish (v + 2^(s-1)) >> s (arithmetic for i16) or, for f32,
 * v * 2^-s (pyramid.zig:50-61). */
static void pyr_store(const orc_image *d, long y, long e, double v, int s) {
    if (d->depth == 5) {
        st_f(d, y, e, (float)v * (1.0f / (float)(1 << s)));
        return;
    }
    const int64_t iv = (int64_t)v;
    st_i(d, y, e, (iv + (1 << (s - 1))) >> s);
}

/* The reference's 5-tap sum (a + f) + (b + e) * 4 + c * 6, in f32 for f32
 * images and exactly otherwise (pyramid.zig:573-590 vert5, :600-607 h5). */
static double tap5(int f32, double a, double b, double c, double e, double f) {
    if (f32) return (double)(((float)a + (float)f) + ((float)b + (float)e) * 4.0f + (float)c * 6.0f);
    return (a + f) + (b + e) * 4 + c * 6;
}

/* capi.zig:234-241 visar_pyr_down -> pyramid.zig:65-122 pyrDown: dst is
 * ((w + 1) / 2, (h + 1) / 2); per output the vertical 5-tap sums of the five
 * border-mapped source columns 2x - 2 .. 2x + 2 over rows 2y - 2 .. 2y + 2,
 * then the horizontal 5-tap sum of those, then (v + 128) >> 8 (v / 256 for
 * f32). Any border but constant. */
int orc_pyr_down(const orc_image *src, const orc_image *dst, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (!src) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc;
    if ((rc = check(src, T)) || (rc = check(dst, T))) return rc;
    if (border == 0) return ORC_UNSUPPORTED;
    const long W = src->width, H = src->height, cn = src->channels;
    if (dst->channels != src->channels || dst->width != (src->width + 1) / 2 || dst->height != (src->height + 1) / 2)
        return ORC_SIZE_MISMATCH;
    const int f32 = T == 5;
    for (long y = band_lo(); y < band_hi((long)dst->height); y++) {
        long ry[5];
        for (int k = 0; k < 5; k++) ry[k] = orc_border_map(2 * y + k - 2, H, border);
        for (long x = 0; x < (long)dst->width; x++) {
            long cx[5];
            for (int i = 0; i < 5; i++) cx[i] = orc_border_map(2 * x + i - 2, W, border);
            for (long ch = 0; ch < cn; ch++) {
                double v[5];
                for (int i = 0; i < 5; i++) {
                    const long e = cx[i] * cn + ch;
                    v[i] = tap5(f32, ldd(src, ry[0], e), ldd(src, ry[1], e), ldd(src, ry[2], e), ldd(src, ry[3], e),
                                ldd(src, ry[4], e));
                }
                pyr_store(dst, y, x * cn + ch, tap5(f32, v[0], v[1], v[2], v[3], v[4]), 8);
            }
        }
    }
    return ORC_OK;
}

/* capi.zig:243-250 visar_pyr_up -> pyramid.zig:403-504 pyrUp (reflect101
 * only): dst is (2w, 2h). Source row y gives output rows 2y (a + 6b + c) and
 * 2y + 1 (4 (b + c)) from rows a = y - 1, b = y, c = y + 1 of the image
 * reflected about its upsampled edges (row -1 is row 1, row h is row h - 1;
 * columns likewise); along the row, output pixel 2p takes
 * (q[p-1] + q[p+1]) + 6 q[p] and 2p + 1 takes 4 (q[p] + q[p+1]); then
 * (v + 32) >> 6 (v / 64 for f32). */
int orc_pyr_up(const orc_image *src, const orc_image *dst, int border) {
    if (border < 0 || border > 4) return ORC_UNSUPPORTED;
    if (!src) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc;
    if ((rc = check(src, T)) || (rc = check(dst, T))) return rc;
    if (border != 4) return ORC_UNSUPPORTED;
    const long w = src->width, h = src->height, cn = src->channels;
    if (dst->channels != src->channels || (uint64_t)dst->width != 2 * (uint64_t)w || (uint64_t)dst->height != 2 * (uint64_t)h)
        return ORC_SIZE_MISMATCH;
    const int f32 = T == 5;
    for (long y = band_lo(); y < band_hi(h); y++) {
        const long ya = y > 0 ? y - 1 : (h > 1 ? 1 : 0), yc = y + 1 < h ? y + 1 : y;
        for (int odd = 0; odd < 2; odd++) {
            for (long p = 0; p < w; p++) {
                const long px[3] = {p > 0 ? p - 1 : (w > 1 ? 1 : 0), p, p + 1 < w ? p + 1 : w - 1};
                for (long ch = 0; ch < cn; ch++) {
                    double q[3];
                    for (int i = 0; i < 3; i++) {
                        const long e = px[i] * cn + ch;
                        const double a = ldd(src, ya, e), b = ldd(src, y, e), c = ldd(src, yc, e);
                        if (f32) q[i] = odd ? (double)(((float)b + (float)c) * 4.0f) : (double)(((float)a + (float)c) + (float)b * 6.0f);
                        else q[i] = odd ? (b + c) * 4 : (a + c) + b * 6;
                    }
                    double ev, od;
                    if (f32) {
                        ev = (double)(((float)q[0] + (float)q[2]) + (float)q[1] * 6.0f);
                        od = (double)(((float)q[1] + (float)q[2]) * 4.0f);
                    } else {
                        ev = (q[0] + q[2]) + q[1] * 6;
                        od = (q[1] + q[2]) * 4;
                    }
                    pyr_store(dst, 2 * y + odd, (2 * p) * cn + ch, ev, 6);
                    pyr_store(dst, 2 * y + odd, (2 * p + 1) * cn + ch, od, 6);
                }
            }
        }
    }
    return ORC_OK;
}

/* capi.zig:326-337 visar_adaptive_threshold -> threshold.zig:306-360:
 * the mean of each pixel's block x block neighbourhood (boxFilter, or the
 * Gaussian blur of an f32 copy saturated back to u8; replicated border),
 * then dst = s + idelta > m ? maxval : 0 (BINARY) or the opposite
 * (BINARY_INV), idelta = ceil(delta) or floor(delta). */
int orc_adaptive_threshold(const orc_image *src, const orc_image *dst, double maxval, int method, int type, int block,
                           double delta) {
    if (method != 0 && method != 1) return ORC_INVALID_ARGUMENT;
    if (type != 0 && type != 1) return ORC_INVALID_ARGUMENT;
    if (block < 3) return ORC_INVALID_ARGUMENT;
    int rc;
    if ((rc = check(src, 0)) || (rc = check(dst, 0))) return rc;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    if (src->channels != 1) return ORC_UNSUPPORTED;
    if (block % 2 != 1) return ORC_INVALID_ARGUMENT;
    const long W = src->width, H = src->height;
    uint8_t *mb = (uint8_t *)malloc((size_t)W * H);
    if (!mb) return ORC_OUT_OF_MEMORY;
    orc_image mean = {(uint32_t)W, (uint32_t)H, 1, 0, (size_t)W, mb};
    if (method == 0) {
        rc = orc_box_filter(src, &mean, block, block, 1, 1);
    } else {
        float *fa = (float *)malloc(sizeof(float) * (size_t)W * H * 2);
        if (!fa) { free(mb); return ORC_OUT_OF_MEMORY; }
        orc_image f = {(uint32_t)W, (uint32_t)H, 1, 5, (size_t)W * 4, fa};
        orc_image fb = {(uint32_t)W, (uint32_t)H, 1, 5, (size_t)W * 4, fa + (size_t)W * H};
        for (long y = 0; y < H; y++)
            for (long x = 0; x < W; x++) fa[y * W + x] = (float)((const uint8_t *)rowp(src, y))[x];
        rc = orc_gaussian_blur(&f, &fb, block, block, 0, 0, 1);
        if (!rc)
            for (long y = 0; y < H; y++)
                for (long x = 0; x < W; x++) mb[y * W + x] = (uint8_t)sat_cast_f((double)((float *)fb.data)[y * W + x], 0);
        free(fa);
    }
    if (rc) { free(mb); return rc; }
    const int64_t imax = sat_cast_f(maxval, 0);
    long idelta = (long)(type == 0 ? ceil(delta) : floor(delta));
    if (idelta < -1024) idelta = -1024;
    if (idelta > 1024) idelta = 1024;
    for (long y = 0; y < H; y++) {
        const uint8_t *sr = (const uint8_t *)rowp(src, y);
        uint8_t *dr = (uint8_t *)rowp(dst, y);
        for (long x = 0; x < W; x++) {
            const int gt = (long)sr[x] + idelta > (long)mb[y * W + x];
            dr[x] = (uint8_t)((type == 0) == gt ? imax : 0);
        }
    }
    free(mb);
    return ORC_OK;
}

/* ---------------------------------------------------------- undistort */

/* calib3d/linalg.zig:417-426 inv3: cofactors times 1 / det; 0 when det == 0. */
static int inv3(const double a[9], double r[9]) {
    const double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[1] * (a[3] * a[8] - a[6] * a[5]) +
                       a[2] * (a[3] * a[7] - a[6] * a[4]);
    if (det == 0) return 0;
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
    return 1;
}

/* linalg.zig:432-438 mul3 */
static void mul3(const double a[9], const double b[9], double r[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

/* calib3d/camera.zig:17-22 expandDist: 0, 4, 5 or 8 coefficients (k1 k2 p1
 * p2 [k3 [k4 k5 k6]]), zero-padded to 8. */
static int expand_dist(const double *d, size_t nd, double k[8]) {
    if (nd != 0 && nd != 4 && nd != 5 && nd != 8) return ORC_INVALID_ARGUMENT;
    for (int i = 0; i < 8; i++) k[i] = (size_t)i < nd ? d[i] : 0.0;
    return ORC_OK;
}

/* undistort.zig:44-78 MapParams.row for one pixel: destination (j, i) ->
 * distorted source (u, v), every operation rounded separately. */
static void map_px(const double ir[9], const double k[8], double fx, double fy, double u0, double v0, double i, double j,
                   double *u, double *v) {
    const double _x = (i * ir[1] + ir[2]) + j * ir[0];
    const double _y = (i * ir[4] + ir[5]) + j * ir[3];
    double x = _x, y = _y;
    if (!(ir[6] == 0 && ir[7] == 0 && ir[8] == 1)) {
        const double _w = (i * ir[7] + ir[8]) + j * ir[6];
        const double w = 1 / _w;
        x = _x * w;
        y = _y * w;
    }
    const double x2 = x * x, y2 = y * y, r2 = x2 + y2, _2xy = 2 * x * y;
    double kr = 1 + ((k[4] * r2 + k[1]) * r2 + k[0]) * r2;
    if (k[5] != 0 || k[6] != 0 || k[7] != 0) kr /= 1 + ((k[7] * r2 + k[6]) * r2 + k[5]) * r2;
    const double xd = x * kr + k[2] * _2xy + k[3] * (r2 + 2 * x2);
    const double yd = y * kr + k[2] * (r2 + 2 * y2) + k[3] * _2xy;
    *u = fx * xd + u0;
    *v = fy * yd + v0;
}

/* undistort.zig:106-115 fix: 1/32-pixel fixed point, clamped, rounded half
 * to even with the 1.5 * 2^52 trick (max before min: a NaN gives -lim). */
static int32_t fix32(double v) {
    const double s = v * 32.0, lim = 2147483647.0;
    const double c = fmin(fmax(s, -lim), lim);
    const double r = c + 6755399441055744.0 - 6755399441055744.0;
    return (int32_t)r;
}

/* undistort.zig:120-157 sample: 15-bit bilinear weights from the 1/32
 * fractions, constant border 0; all four taps inside, all outside, or per
 * tap. (linear16, :159-201, is the same integer factored by rows.) */
static void undist_px(const orc_image *src, int cn, int32_t iu, int32_t iv, uint8_t *out) {
    const int32_t sx = iu >> 5, sy = iv >> 5, a = iu & 31, b = iv & 31;
    const int32_t w[4] = {(32 - a) * (32 - b), a * (32 - b), (32 - a) * b, a * b};
    const int32_t W = (int32_t)src->width, H = (int32_t)src->height;
    if (sx >= 0 && sx < W - 1 && sy >= 0 && sy < H - 1) {
        const uint8_t *p0 = (const uint8_t *)rowp(src, sy) + (size_t)sx * cn;
        const uint8_t *p1 = p0 + src->stride;
        for (int ch = 0; ch < cn; ch++) {
            const int32_t t = p0[ch] * w[0] + p0[ch + cn] * w[1] + p1[ch] * w[2] + p1[ch + cn] * w[3];
            out[ch] = (uint8_t)((t + 512) >> 10);
        }
        return;
    }
    if (sx >= W || sx + 1 < 0 || sy >= H || sy + 1 < 0) {
        for (int ch = 0; ch < cn; ch++) out[ch] = 0;
        return;
    }
    for (int ch = 0; ch < cn; ch++) {
        int32_t t = 0;
        for (int q = 0; q < 4; q++) {
            const int32_t px = sx + (q & 1), py = sy + (q >> 1);
            const int32_t v = (px >= 0 && px < W && py >= 0 && py < H) ? ((const uint8_t *)rowp(src, py))[(size_t)px * cn + ch] : 0;
            t += v * w[q];
        }
        out[ch] = (uint8_t)((t + 512) >> 10);
    }
}

/* undistort.zig:205-272 undistort (u8, 1-4 channels, dst the shape of src,
 * new_k defaulting to K) and capi_calib3d.zig:186-191: stripes of
 * max(1, 4096 / width) rows, each with its own inverse of new_k whose
 * principal point is moved up by the stripe's first row; row i of a stripe
 * is mapped with i counted from that row. A stripe whose inverse is singular
 * is left unwritten (the reference returns from its chunk of stripes). */
int orc_undistort(const orc_image *src, const orc_image *dst, const double K[9], const double *dist, size_t nd,
                  const double *new_k) {
    int rc;
    if ((rc = check(src, 0)) || (rc = check(dst, 0))) return rc;
    if (src->width != dst->width || src->height != dst->height || src->channels != dst->channels) return ORC_SIZE_MISMATCH;
    if (src->channels < 1 || src->channels > 4) return ORC_UNSUPPORTED;
    double k[8];
    if ((rc = expand_dist(dist, nd, k))) return rc;
    const double *ar0 = new_k ? new_k : K;
    const int cn = (int)src->channels;
    const uint32_t W = src->width, H = src->height;
    uint32_t stripe = 4096 / W;
    if (stripe < 1) stripe = 1;
    if (stripe > H) stripe = H;
    for (uint32_t y0 = 0; y0 < H; y0 += stripe) {
        double ar[9], ir[9];
        memcpy(ar, ar0, sizeof ar);
        ar[5] = ar0[5] - (double)y0;
        if (!inv3(ar, ir)) continue;
        const uint32_t y1 = y0 + stripe < H ? y0 + stripe : H;
        const uint32_t ylo = y0 > (uint32_t)band_lo() ? y0 : (uint32_t)band_lo();
        const uint32_t yhi = y1 < (uint32_t)band_hi(H) ? y1 : (uint32_t)band_hi(H);
        for (uint32_t y = ylo; y < yhi; y++) {
            uint8_t *d = (uint8_t *)rowp(dst, y);
            for (uint32_t x = 0; x < W; x++) {
                double u, v;
                map_px(ir, k, K[0], K[4], K[2], K[5], (double)(y - y0), (double)x, &u, &v);
                undist_px(src, cn, fix32(u), fix32(v), d + (size_t)x * cn);
            }
        }
    }
    return ORC_OK;
}

/* undistort.zig:36-41 MapParams.init (inverse of new_k * R, singular:
 * InvalidArgument) and :82-103 initUndistortRectifyMap into f32 maps, with
 * capi_calib3d.zig:195-201 (maps checked as 32F first). */
int orc_init_undistort_rectify_map(const double K[9], const double *dist, size_t nd, const double *R, const double *new_k,
                                   const orc_image *mapx, const orc_image *mapy) {
    int rc;
    if ((rc = check(mapx, 5)) || (rc = check(mapy, 5))) return rc;
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double m[9], ir[9], k[8];
    mul3(new_k ? new_k : K, R ? R : I, m);
    if (!inv3(m, ir)) return ORC_INVALID_ARGUMENT;
    if ((rc = expand_dist(dist, nd, k))) return rc;
    if (mapx->width != mapy->width || mapx->height != mapy->height || mapx->channels != mapy->channels ||
        mapx->channels != 1)
        return ORC_SIZE_MISMATCH;
    for (uint32_t y = 0; y < mapx->height; y++) {
        float *ox = (float *)rowp(mapx, y), *oy = (float *)rowp(mapy, y);
        for (uint32_t x = 0; x < mapx->width; x++) {
            double u, v;
            map_px(ir, k, K[0], K[4], K[2], K[5], (double)y, (double)x, &u, &v);
            ox[x] = (float)u;
            oy[x] = (float)v;
        }
    }
    return ORC_OK;
}

typedef struct { orc_image im; void *buf; } tmp_image;

static int tmp_like(const orc_image *s, tmp_image *t) {
    size_t rl = (size_t)s->width * s->channels * dsize(s->depth);
    t->buf = malloc(rl * s->height);
    if (!t->buf) return ORC_OUT_OF_MEMORY;
    t->im = *s;
    t->im.stride = rl;
    t->im.data = t->buf;
    return ORC_OK;
}

static void copy_img(const orc_image *s, const orc_image *d) {
    size_t rl = (size_t)s->width * s->channels * dsize(s->depth);
    for (long y = 0; y < s->height; y++) memmove(rowp(d, y), rowp(s, y), rl);
}

/* morph.zig:63-96 erode/dilate with iterations (via the per-tap pass). */
static int morph_iter(const orc_image *s, const orc_image *d, int op_max, const uint8_t *mask, int kw, int kh,
                      int iterations) {
    if (iterations == 0) { copy_img(s, d); return ORC_OK; }
    tmp_image a, b;
    if (tmp_like(s, &a)) return ORC_OUT_OF_MEMORY;
    if (tmp_like(s, &b)) { free(a.buf); return ORC_OUT_OF_MEMORY; }
    copy_img(s, &a.im);
    orc_image *cur = &a.im, *nxt = &b.im;
    for (int it = 0; it < iterations; it++) {
        orc_morph_pass(cur, nxt, op_max, mask, kw, kh, kw / 2, kh / 2);
        orc_image *t = cur; cur = nxt; nxt = t;
    }
    copy_img(cur, d);
    free(a.buf); free(b.buf);
    return ORC_OK;
}

/* morph.zig:574-593 d = saturate(a - b) */
static void sub_sat(const orc_image *a, const orc_image *b, const orc_image *d) {
    long n = (long)a->width * a->channels;
    for (long y = band_lo(); y < band_hi(a->height); y++)
        for (long i = 0; i < n; i++) {
            if (a->depth == 5) { st_f(d, y, i, ldf(a, y, i) - ldf(b, y, i)); continue; }
            int64_t x = ldi(a, y, i), z = ldi(b, y, i);
            st_i(d, y, i, x > z ? x - z : 0);
        }
}

/* morph.zig:574-593 subInPlace as an entry point: d = a - b, saturating at 0. */
int orc_sub_sat(const orc_image *a, const orc_image *b, const orc_image *d) {
    if (!a || !b || !d || !same_shape(a, b) || !same_shape(a, d) || a->depth != b->depth || a->depth != d->depth)
        return ORC_INVALID_ARGUMENT;
    sub_sat(a, b, d);
    return ORC_OK;
}

/* morph.zig:532-571 morphologyEx, C ABI checks of capi.zig:339-350 */
int orc_morphology(const orc_image *src, const orc_image *dst, int op, int shape, int kw, int kh, int iterations) {
    if (op < 0 || op > 6 || shape < 0 || shape > 2) return ORC_UNSUPPORTED;
    if (kw <= 0 || kh <= 0 || kw > 255 || kh > 255 || iterations < 0) return ORC_INVALID_ARGUMENT;
    if (!src || !dst) return ORC_INVALID_ARGUMENT;
    uint32_t T = src->depth;
    if (T != 0 && T != 2 && T != 3 && T != 5) return ORC_UNSUPPORTED;
    int rc = check(src, T);
    if (rc) return rc;
    if ((rc = check(dst, T))) return rc;
    if (!same_shape(src, dst)) return ORC_SIZE_MISMATCH;
    uint8_t *mask = (uint8_t *)malloc((size_t)kw * (size_t)kh);
    if (!mask) return ORC_OUT_OF_MEMORY;
    orc_structuring_element(shape, kw, kh, mask);
    tmp_image t1, t2, s0;
    if (tmp_like(src, &t1) || tmp_like(src, &t2) || tmp_like(src, &s0)) { free(mask); return ORC_OUT_OF_MEMORY; }
    copy_img(src, &s0.im);
    const orc_image *S = &s0.im;
    switch (op) {
    case 0: morph_iter(S, dst, 0, mask, kw, kh, iterations); break;
    case 1: morph_iter(S, dst, 1, mask, kw, kh, iterations); break;
    case 2: morph_iter(S, &t1.im, 0, mask, kw, kh, iterations); morph_iter(&t1.im, dst, 1, mask, kw, kh, iterations); break;
    case 3: morph_iter(S, &t1.im, 1, mask, kw, kh, iterations); morph_iter(&t1.im, dst, 0, mask, kw, kh, iterations); break;
    case 4:
        morph_iter(S, &t1.im, 0, mask, kw, kh, iterations);
        morph_iter(S, &t2.im, 1, mask, kw, kh, iterations);
        sub_sat(&t2.im, &t1.im, dst);
        break;
    case 5:
        morph_iter(S, &t1.im, 0, mask, kw, kh, iterations);
        morph_iter(&t1.im, &t2.im, 1, mask, kw, kh, iterations);
        sub_sat(S, &t2.im, dst);
        break;
    default:
        morph_iter(S, &t1.im, 1, mask, kw, kh, iterations);
        morph_iter(&t1.im, &t2.im, 0, mask, kw, kh, iterations);
        sub_sat(&t2.im, S, dst);
        break;
    }
    free(mask); free(t1.buf); free(t2.buf); free(s0.buf);
    return ORC_OK;
}

/* ----------------------------------------------------------- background */

/* video/track.zig:33-81 Background.apply; capi.zig:446-469 */
int orc_background_apply(uint16_t *model, int *primed, int shift, const orc_image *frame, const orc_image *mask,
                         int threshold) {
    int rc;
    if ((rc = check(frame, 0))) return rc;
    if ((rc = check(mask, 0))) return rc;
    if (frame->channels != 1 || !same_shape(frame, mask)) return ORC_SIZE_MISMATCH;
    int t = threshold < 0 ? 0 : threshold > 255 ? 255 : threshold;
    long W = frame->width, H = frame->height;
    if (!*primed) {
        for (long y = 0; y < H; y++)
            for (long x = 0; x < W; x++) {
                model[y * W + x] = (uint16_t)(ldi(frame, y, x) << 8);
                st_i(mask, y, x, 0);
            }
        *primed = 1;
        return ORC_OK;
    }
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            int32_t fv = (int32_t)ldi(frame, y, x), bv = model[y * W + x];
            int32_t diff = fv - (bv >> 8);
            int fg = (diff < 0 ? -diff : diff) > t;
            st_i(mask, y, x, fg ? 255 : 0);
            int sh = shift + (fg ? 2 : 0);
            model[y * W + x] = (uint16_t)(bv + ((((int32_t)fv << 8) - bv) >> sh));
        }
    return ORC_OK;
}

/* ------------------------------------------------- connected components */

static long uf_find(long *p, long x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}
static void uf_union(long *p, long a, long b) {
    a = uf_find(p, a); b = uf_find(p, b);
    if (a == b) return;
    if (a < b) p[b] = a; else p[a] = b;
}

/*
 * imgproc/components.zig:408-535 and capi.zig:533-550, restated per pixel:
 * 4-connected components are numbered in raster order of their first
 * pixel; 8-connected ones by (first 2x2 block row, smallest block column
 * of the component inside that block row). Statistics are integer sums
 * (twice the x sum), centroids sx2 / 2 / area and sy / area.
 */
int orc_connected_components(const orc_image *src, const orc_image *labels, int connectivity, int32_t *stats,
                             double *centroids, size_t max, size_t *count) {
    int rc;
    if ((rc = check(src, 0))) return rc;
    if ((rc = check(labels, 4))) return rc;
    if (src->channels != 1 || labels->channels != 1 || src->width != labels->width || src->height != labels->height)
        return ORC_SIZE_MISMATCH;
    long W = src->width, H = src->height, N = W * H;
    int eight = connectivity != 4;
    long *p = (long *)malloc(sizeof(long) * (size_t)N);
    long *lab = (long *)malloc(sizeof(long) * (size_t)N);
    if (!p || !lab) { free(p); free(lab); return ORC_OUT_OF_MEMORY; }
    for (long i = 0; i < N; i++) p[i] = i;
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            if (!ldi(src, y, x)) continue;
            long i = y * W + x;
            if (x > 0 && ldi(src, y, x - 1)) uf_union(p, i, i - 1);
            if (y > 0) {
                if (ldi(src, y - 1, x)) uf_union(p, i, i - W);
                if (eight && x > 0 && ldi(src, y - 1, x - 1)) uf_union(p, i, i - W - 1);
                if (eight && x + 1 < W && ldi(src, y - 1, x + 1)) uf_union(p, i, i - W + 1);
            }
        }
    /* ordering key per root */
    long *key = (long *)malloc(sizeof(long) * (size_t)N);
    long *order = (long *)malloc(sizeof(long) * (size_t)N);
    if (!key || !order) { free(p); free(lab); free(key); free(order); return ORC_OUT_OF_MEMORY; }
    for (long i = 0; i < N; i++) key[i] = -1;
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            if (!ldi(src, y, x)) continue;
            long r = uf_find(p, y * W + x);
            long k;
            if (!eight) k = y * W + x;
            else {
                long ry = r / W;
                if ((y >> 1) != (ry >> 1)) continue; /* only the block row of the first pixel */
                k = (y >> 1) * (W + 2) + (x >> 1);
            }
            if (key[r] < 0 || k < key[r]) key[r] = k;
        }
    long nroots = 0;
    for (long i = 0; i < N; i++)
        if (key[i] >= 0) order[nroots++] = i;
    /* sort roots by key (insertion into a keyed array: keys are unique) */
    long maxkey = eight ? ((H + 1) / 2 + 1) * (W + 2) : N;
    long *byk = (long *)malloc(sizeof(long) * (size_t)(maxkey + 1));
    if (!byk) { free(p); free(lab); free(key); free(order); return ORC_OUT_OF_MEMORY; }
    for (long i = 0; i <= maxkey; i++) byk[i] = -1;
    for (long i = 0; i < nroots; i++) byk[key[order[i]]] = order[i];
    long next = 1;
    for (long i = 0; i < N; i++) lab[i] = 0;
    for (long k = 0; k <= maxkey; k++)
        if (byk[k] >= 0) lab[byk[k]] = next++;
    size_t ncomp = (size_t)next;
    uint64_t *acc = (uint64_t *)calloc(ncomp * 7, sizeof(uint64_t));
    if (!acc) { free(p); free(lab); free(key); free(order); free(byk); return ORC_OUT_OF_MEMORY; }
    for (size_t i = 0; i < ncomp; i++) { acc[i * 7 + 0] = UINT64_MAX; acc[i * 7 + 1] = UINT64_MAX; }
    for (long y = 0; y < H; y++)
        for (long x = 0; x < W; x++) {
            long l = ldi(src, y, x) ? lab[uf_find(p, y * W + x)] : 0;
            ((int32_t *)rowp(labels, y))[x] = (int32_t)l;
            uint64_t *a = acc + (size_t)l * 7;
            if ((uint64_t)x < a[0]) a[0] = (uint64_t)x;
            if ((uint64_t)y < a[1]) a[1] = (uint64_t)y;
            if ((uint64_t)x > a[2]) a[2] = (uint64_t)x;
            if ((uint64_t)y > a[3]) a[3] = (uint64_t)y;
            a[4] += 1;
            a[5] += 2 * (uint64_t)x;
            a[6] += (uint64_t)y;
        }
    *count = ncomp;
    for (size_t i = 0; i < ncomp && i < max; i++) {
        uint64_t *a = acc + i * 7;
        if (stats) {
            if (a[4] == 0) {
                stats[i * 5 + 0] = -1; stats[i * 5 + 1] = 2147483647; stats[i * 5 + 2] = 0; stats[i * 5 + 3] = 0; stats[i * 5 + 4] = 0;
            } else {
                stats[i * 5 + 0] = (int32_t)a[0]; stats[i * 5 + 1] = (int32_t)a[1];
                stats[i * 5 + 2] = (int32_t)(a[2] - a[0] + 1); stats[i * 5 + 3] = (int32_t)(a[3] - a[1] + 1);
                stats[i * 5 + 4] = (int32_t)a[4];
            }
        }
        if (centroids) {
            double ar = (double)(a[4] > 0 ? a[4] : 1);
            if (a[4] == 0) { centroids[i * 2] = NAN; centroids[i * 2 + 1] = NAN; }
            else { centroids[i * 2] = (double)a[5] / 2.0 / ar; centroids[i * 2 + 1] = (double)a[6] / ar; }
        }
    }
    free(p); free(lab); free(key); free(order); free(byk); free(acc);
    return ORC_OK;
}
