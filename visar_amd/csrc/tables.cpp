// tables.cpp -- host-side coefficient tables (built with -ffp-contract=off).
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace visar {

Tuning &tuning() {
    static Tuning t;
    return t;
}

namespace {
struct Knob {
    const char *key;
    int Tuning::*field;
    int lo, hi;
};
const Knob kKnobs[] = {
    {"stencil_rows", &Tuning::stencil_rows, 0, 4},
    {"batch_stencil_rows", &Tuning::batch_stencil_rows, 0, 4},
    {"batch_unit_kb", &Tuning::batch_unit_kb, 0, 1024},
    {"tile_order", &Tuning::tile_order, 0, 2},
    {"sep_rows", &Tuning::sep_rows, 0, 16},
    {"warp_layout", &Tuning::warp_layout, 0, 2},
    {"morph_chain", &Tuning::morph_chain, 0, 2},
    {"morph_rows", &Tuning::morph_rows, 0, 128},
    {"stencil_direct", &Tuning::stencil_direct, 0, 1},
    {"stencil_band", &Tuning::stencil_band, 0, 16},
    {"stencil_wg", &Tuning::stencil_wg, 0, 4},
    {"gray_layout", &Tuning::gray_layout, 0, 1},
    {"lin_stream", &Tuning::lin_stream, 0, 1},
    {"lin_rows", &Tuning::lin_rows, 0, 8},
    {"lin_run", &Tuning::lin_run, 0, 16},
    {"warp_blocks", &Tuning::warp_blocks, 0, 2},
    {"morph_stream", &Tuning::morph_stream, 0, 1},
    {"morph_band", &Tuning::morph_band, 0, 256},
    {"morph_pass", &Tuning::morph_pass, 0, 1},
    {"morph_pass_rows", &Tuning::morph_pass_rows, 0, 16},
    {"morph_pow_rows", &Tuning::morph_pow_rows, 0, 8},
};
}  // namespace

bool set_tuning(const char *key, int value) {
    for (const Knob &k : kKnobs)
        if (std::strcmp(key, k.key) == 0) {
            if (value < k.lo || value > k.hi || (k.field == &Tuning::stencil_rows && value == 3)) return false;
            tuning().*(k.field) = value;
            return true;
        }
    return false;
}

bool get_tuning(const char *key, int *value) {
    for (const Knob &k : kKnobs)
        if (std::strcmp(key, k.key) == 0) {
            *value = tuning().*(k.field);
            return true;
        }
    return false;
}

namespace {
double bits_to_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}
}  // namespace

// 2^(i/64) for i in [0, 64), the table of OpenCV's softfloat exp (the same
// numbers as filter.zig:357-373; any correctly rounded 2^(i/64) table is it).
static const uint64_t kExp2Tab[64] = {
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

double cv_exp(double x) {
    if (std::isnan(x)) return x;
    if (std::isinf(x)) return x > 0 ? x : 0.0;
    // Polynomial coefficients of the reduced argument, scaled by 1/a0.
    static const double a0 = bits_to_double(0x3f83ce0f3e46f431ull);
    static const double c[6] = {
        bits_to_double(0x3f55e7aa1566c2a4ull) / a0, bits_to_double(0x3f83b2a72b4f3cd3ull) / a0,
        bits_to_double(0x3fac6b08d81fec75ull) / a0, bits_to_double(0x3fcebfbdff82a45aull) / a0,
        bits_to_double(0x3fe62e42fefa39f1ull) / a0, 1.0 / a0,
    };
    static const double prescale = bits_to_double(0x3ff71547652b82feull) * 64.0;  // 64 / ln 2
    const double limit = 3000.0 * 64.0;
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    double t = ((b >> 52) & 0x7ff) > 1033 ? (x < 0 ? -limit : limit) : x * prescale;
    double r = rint_magic(t);
    int32_t v = (int32_t)r;
    long e = (long)(v >> 6) + 1023;
    e = e < 0 ? 0 : (e > 2047 ? 2047 : e);
    double scale = bits_to_double((uint64_t)e << 52);
    t = (t - r) * (1.0 / 64.0);
    double poly = c[0];
    for (int i = 1; i < 6; i++) poly = poly * t + c[i];
    return scale * a0 * bits_to_double(kExp2Tab[v & 63]) * poly;
}

void gaussian_kernel(double *out, int n, double sigma) {
    if (n % 2 == 1 && n <= 9 && sigma <= 0) {
        // Binomial rows (OpenCV's small fixed kernels).
        static const double k1[] = {1.0};
        static const double k3[] = {0.25, 0.5, 0.25};
        static const double k5[] = {0.0625, 0.25, 0.375, 0.25, 0.0625};
        static const double k7[] = {0.03125, 0.109375, 0.21875, 0.28125, 0.21875, 0.109375, 0.03125};
        static const double k9[] = {4.0 / 256, 13.0 / 256, 30.0 / 256, 51.0 / 256, 60.0 / 256,
                                    51.0 / 256, 30.0 / 256, 13.0 / 256, 4.0 / 256};
        const double *t[] = {k1, k3, k5, k7, k9};
        std::memcpy(out, t[n / 2], sizeof(double) * (size_t)n);
        return;
    }
    double s = sigma > 0 ? sigma : std::fma((double)n, 0.15, 0.35);
    double scale2x = -0.125 / (s * s);
    int half = (n - 1) / 2;
    double sum = 0;
    for (int i = 0; i < half; i++) {
        long x = 1 - (long)n + 2 * i;
        out[i] = cv_exp((double)(x * x) * scale2x);
        sum += out[i];
    }
    sum = sum * 2 + 1;
    if (n % 2 == 0) sum += 1;
    double inv = 1.0 / sum;
    for (int i = 0; i < half; i++) {
        out[i] *= inv;
        out[n - 1 - i] = out[i];
    }
    out[half] = inv;
    if (n % 2 == 0) out[half + 1] = inv;
}

void gaussian_kernel_fixed(uint32_t *out, int n, double sigma, int bits) {
    double k[kMaxTaps];
    gaussian_kernel(k, n, sigma);
    const double one = (double)((uint64_t)1 << bits);
    double carry = 0;
    uint64_t total = 0;
    for (int i = 0; i < n / 2; i++) {
        double a = k[i] * one + carry;
        double q = rint_magic(a);
        carry = a - q;
        uint64_t v = q > 0 ? (uint64_t)q : 0;
        out[i] = out[n - 1 - i] = (uint32_t)v;
        total += v;
    }
    out[n / 2] = (uint32_t)(((uint64_t)1 << bits) - 2 * total);
}

int gaussian_ksize(double sigma, bool u8) {
    double r = std::round(sigma * (u8 ? 3.0 : 4.0) * 2 + 1);
    if (!(r >= 0) || r > 1e6) return -1;
    return ((int)r) | 1;
}

bool deriv_kernel(int32_t *out, int n_out, int order, int ksize) {
    if (ksize == -1) {
        if (n_out != 3 || order > 1) return false;
        const int32_t s[3] = {3, 10, 3}, d[3] = {-1, 0, 1};
        std::memcpy(out, order == 0 ? s : d, sizeof s);
        return true;
    }
    int n = ksize == 1 ? 3 : ksize;
    if (n_out != n || n > 31) return false;
    if (ksize == 1 && order == 0) {
        out[0] = 0, out[1] = 1, out[2] = 0;
        return true;
    }
    if (order >= n) return false;
    // Pascal smoothing (n - order - 1 convolutions with [1 1]), then
    // differencing (order convolutions with [-1 1]).
    std::vector<int32_t> a{1};
    for (int s = 0; s < n - 1; s++) {
        std::vector<int32_t> b(a.size() + 1, 0);
        bool diff = s >= n - order - 1;
        for (size_t j = 0; j < a.size(); j++) {
            b[j] += diff ? -a[j] : a[j];
            b[j + 1] += a[j];
        }
        a.swap(b);
    }
    std::memcpy(out, a.data(), sizeof(int32_t) * (size_t)n);
    return true;
}

int symmetry(const float *k, int n) {
    bool even = true, odd = true;
    for (int i = 0; i < n / 2; i++) {
        if (k[i] != k[n - 1 - i]) even = false;
        if (k[i] != -k[n - 1 - i]) odd = false;
    }
    if (n % 2 == 1 && k[n / 2] != 0) odd = false;
    return even ? SYM_EVEN : odd ? SYM_ODD : SYM_NONE;
}

void linear_coord(long d, double scale, uint32_t len, bool area, long *index, double *frac) {
    long sx;
    double f;
    if (area) {
        const double df = (double)d;
        sx = (long)std::floor(df * scale);
        f = (df + 1) - (double)(sx + 1) / scale;
        f = f <= 0 ? 0 : f - std::floor(f);
    } else {
        const double fx = ((double)d + 0.5) * scale - 0.5;
        sx = (long)std::floor(fx);
        f = fx - (double)sx;
    }
    if (sx < 0) sx = 0, f = 0;
    if (sx >= (long)len - 1) sx = (long)len - 1, f = 0;
    *index = sx;
    *frac = f;
}

LinTap linear_tap_u8(long d, uint32_t ssize, uint32_t dsize, bool area, bool pin) {
    double inv = (double)dsize / (double)ssize;
    double scale = 1.0 / inv;
    double df = (double)d;
    long sx;
    float fx;
    if (area) {
        sx = (long)std::floor(df * scale);
        fx = (float)((df + 1) - (double)(sx + 1) * inv);
        fx = fx <= 0 ? 0.0f : fx - std::floor(fx);
    } else {
        fx = (float)((df + 0.5) * scale - 0.5);
        float fl = std::floor(fx);
        sx = (long)fl;
        fx -= fl;
    }
    long last = (long)ssize - 1;
    if (pin && sx < 0) sx = 0, fx = 0;
    if (pin && sx >= last) sx = last, fx = 0;
    LinTap t;
    t.w0 = sat_cast<int16_t>((double)((1.0f - fx) * 2048.0f));
    t.w1 = sat_cast<int16_t>((double)(fx * 2048.0f));
    t.o0 = (int32_t)(sx < 0 ? 0 : sx > last ? last : sx);
    t.o1 = (int32_t)(sx + 1 < 0 ? 0 : sx + 1 > last ? last : sx + 1);
    return t;
}

bool linear_box_fits(const LinTap *xt, int dw, int cn, const LinTap *yt, int dh) {
    auto mono = [](const LinTap *t, int n) {
        for (int i = 0; i < n; i++) {
            if (t[i].o0 > t[i].o1) return false;
            if (i && (t[i].o0 < t[i - 1].o0 || t[i].o1 < t[i - 1].o1)) return false;
        }
        return true;
    };
    if (dw <= 0 || dh <= 0 || cn <= 0 || !mono(xt, dw) || !mono(yt, dh)) return false;
    const long n = (long)dw * cn;
    for (long e = 0; e < n; e += 256) {
        const long last = std::min(n, e + 256) - 1;
        const long span = (long)xt[last / cn].o1 + cn - 1 - xt[e / cn].o0 + 1;
        if ((span + 6) / 4 > kLinWords) return false;
    }
    for (int y = 0; y < dh; y += kLinTH) {
        const int last = std::min(dh, y + kLinTH) - 1;
        if (yt[last].o1 - yt[y].o0 + 1 > kLinRows) return false;
    }
    return true;
}

bool linear_stream_fits(const LinTap *xt, int dw, int cn, const LinTap *yt, int dh) {
    if (tuning().lin_stream == 1 || dw <= 0 || dh <= 0 || cn < 1 || cn > 4) return false;
    auto weights = [](const LinTap &t) { return t.w0 >= 0 && t.w1 >= 0 && t.w0 + t.w1 <= 2050; };
    for (int x = 0; x < dw; x++)
        if (!weights(xt[x]) || xt[x].o0 < 0 || xt[x].o1 < xt[x].o0) return false;
    for (int y = 0; y < dh; y++) {
        if (!weights(yt[y]) || yt[y].o0 < 0 || yt[y].o1 - yt[y].o0 < 0 || yt[y].o1 - yt[y].o0 > 1) return false;
        if (y && (yt[y].o0 < yt[y - 1].o0 || yt[y].o1 < yt[y - 1].o1)) return false;
    }
    const long n = (long)dw * cn;
    for (long g = 0; g < n; g += 4) {
        long lo = 0, hi = 0;
        for (int j = 0; j < 4; j++) {
            const long e = std::min(g + j, n - 1), x = e / cn, ch = e - x * cn;
            const long a = xt[x].o0 + ch, b = xt[x].o1 + ch;
            lo = j ? std::min(lo, a) : a;
            hi = j ? std::max(hi, b) : b;
        }
        if (hi - lo > 7) return false;
    }
    return true;
}

bool area_box_fits(const AreaEntry *xt, const int32_t *xs, int dw, int cn, const AreaEntry *yt, const int32_t *ys, int dh) {
    if (dw <= 0 || dh <= 0 || cn <= 0) return false;
    // One band of cells [c0, c1]: its entries' range must be that of its
    // first and last entries.
    auto band = [](const AreaEntry *t, const int32_t *st, int c0, int c1, long *lo, long *hi) {
        const int a = st[c0], b = st[c1 + 1];
        if (b <= a) return false;
        long mn = t[a].si, mx = t[a].si;
        for (int q = a; q < b; q++) mn = std::min<long>(mn, t[q].si), mx = std::max<long>(mx, t[q].si);
        *lo = mn;
        *hi = mx;
        return mn == t[a].si && mx == t[b - 1].si;
    };
    for (int x = 0; x < dw; x++)
        if (xs[x + 1] - xs[x] < 1 || xs[x + 1] - xs[x] > kAreaMaxCell) return false;
    for (int y = 0; y < dh; y++)
        if (ys[y + 1] - ys[y] < 1) return false;
    const long n = (long)dw * cn;
    for (long e = 0; e < n; e += 256) {
        long lo, hi;
        if (!band(xt, xs, (int)(e / cn), (int)((std::min(n, e + 256) - 1) / cn), &lo, &hi)) return false;
        if (((hi - lo) * cn + cn + 6) / 4 > kAreaWords) return false;
    }
    for (int y = 0; y < dh; y += kAreaTH) {
        long lo, hi;
        if (!band(yt, ys, y, std::min(dh, y + kAreaTH) - 1, &lo, &hi)) return false;
        if (hi - lo + 1 > kAreaRows) return false;
    }
    return true;
}

void area_tab(uint32_t ssize, uint32_t dsize, double scale, std::vector<AreaEntry> &e, std::vector<int32_t> &start) {
    e.clear();
    start.assign(dsize + 1, 0);
    const long last = (long)ssize - 1;
    for (uint32_t dx = 0; dx < dsize; dx++) {
        start[dx] = (int32_t)e.size();
        double f1 = (double)dx * scale, f2 = f1 + scale;
        double cellw = std::fmin(scale, (double)ssize - f1);
        long s1 = (long)std::ceil(f1), s2 = (long)std::floor(f2);
        long s2c = s2 < last ? s2 : last;
        if (s1 > last) s1 = last;
        if ((double)s1 - f1 > 1e-3) e.push_back({(int32_t)(s1 - 1), (float)(((double)s1 - f1) / cellw)});
        for (long s = s1; s < s2; s++) e.push_back({(int32_t)s, (float)(1.0 / cellw)});
        if (f2 - (double)s2 > 1e-3) {
            double a = std::fmin(std::fmin(f2 - (double)s2, 1.0), cellw);
            e.push_back({(int32_t)s2c, (float)(a / cellw)});
        }
    }
    start[dsize] = (int32_t)e.size();
}

void structuring_element(int shape, int w, int h, std::vector<uint8_t> &m) {
    m.assign((size_t)w * h, 0);
    const int ax = w / 2, ay = h / 2, r = ay, c = ax;
    const double inv_r2 = r > 0 ? 1.0 / (double)(r * r) : 0.0;
    for (int i = 0; i < h; i++) {
        int j1 = 0, j2 = 0;
        if (shape == 0 || (shape == 1 && i == ay)) {
            j2 = w;
        } else if (shape == 1) {
            j1 = ax, j2 = ax + 1;
        } else {
            long dy = (long)i - r;
            if (std::labs(dy) <= r) {
                double dxf = (double)c * std::sqrt((double)((long)r * r - dy * dy) * inv_r2);
                long dx = std::lround(dxf);
                j1 = (int)std::max(0L, (long)c - dx);
                j2 = (int)std::min((long)w, (long)c + dx + 1);
            }
        }
        for (int j = j1; j < j2; j++) m[(size_t)i * w + j] = 1;
    }
}

bool invert_affine(const double m[6], double o[6]) {
    double det = m[0] * m[4] - m[1] * m[3];
    if (det == 0) return false;
    double d = 1.0 / det;
    double a11 = m[4] * d, a22 = m[0] * d, a12 = -m[1] * d, a21 = -m[3] * d;
    o[0] = a11;
    o[1] = a12;
    o[2] = -a11 * m[2] - a12 * m[5];
    o[3] = a21;
    o[4] = a22;
    o[5] = -a21 * m[2] - a22 * m[5];
    return true;
}

bool invert3(const double m[9], double o[9]) {
    double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[3] * m[8] - m[5] * m[6], c02 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * c00 - m[1] * c01 + m[2] * c02;
    if (det == 0) return false;
    double d = 1.0 / det;
    o[0] = c00 * d;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * d;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * d;
    o[3] = (m[5] * m[6] - m[3] * m[8]) * d;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * d;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * d;
    o[6] = c02 * d;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * d;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * d;
    return true;
}

}  // namespace visar
