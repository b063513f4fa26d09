"""Pins for the oracle's float-inside paths (CPU): generic LINEAR resize, the
AREA tables and fractional AREA, warpAffine, warpPerspective and remap.

None of these has a known-answer vector in the reference. Each is held to
what the mathematics fixes, computed independently of the oracle's own
operation sequence:

- the reference's own cross-path identities (resize.zig:654-680: the exact
  half and the exact 2x fast paths equal the generic LINEAR kernel, so the
  generic 11-bit pipeline must reduce to their closed forms);
- exact rational constructions (the cells of computeResizeAreaTab with its
  1e-3 sliver rule, the half-pixel LINEAR coordinates);
- dyadic matrices and maps, where every float operation of the reference is
  exact, so the output is the exact bilinear value rounded half to even;
- for general matrices and maps, the float64 bilinear value, which the
  reference's f32 chain must round to the same integer wherever that value
  is not within a small margin of a rounding tie or a coordinate of an
  integer (the only places float rounding can decide the result).

What stays open: at exact ties of a general (non-dyadic) case, whether a
product was fused or rounded separately; those bits follow the cited Zig by
restatement only.
"""
from fractions import Fraction

import numpy as np
import pytest


def lcg_fill(shape, seed):
    """The reference tests' byte generator (resize.zig:657-662, :699-704):
    seed = seed * 1103515245 + 12345 (mod 2^32), byte = seed >> 16."""
    n = int(np.prod(shape))
    out = np.empty(n, np.uint8)
    s = seed
    for i in range(n):
        s = (s * 1103515245 + 12345) & 0xFFFFFFFF
        out[i] = (s >> 16) & 0xFF
    return out.reshape(shape)


@pytest.fixture(scope="module")
def src150():
    return {cn: lcg_fill((64, 150, cn), 12345) for cn in (1, 2, 3, 4)}


@pytest.fixture(scope="module")
def src97():
    return {cn: lcg_fill((61, 97, cn), 777) for cn in (1, 2, 3, 4)}


def squeeze(a):
    return a[:, :, 0] if a.shape[2] == 1 else a


# ---------------------------------------------------- LINEAR (generic u8)

@pytest.mark.parametrize("cn", [1, 2, 3, 4])
def test_generic_linear_at_exact_half_is_rounded_block_mean(orc, src150, cn):
    """resize.zig:654-680 holds ru8.half == generic LINEAR at 150x64 -> 75x32."""
    a = squeeze(src150[cn])
    got = orc.resize_linear_generic(a, (75, 32)).astype(np.int32)
    s = a.astype(np.int32)
    want = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(orc.resize(a, (75, 32), 1), want)


def up2_closed_form(a):
    """Bilinear 2x from the integer algebra of resize_u8.zig:796-800: with
    OpenCV's weights (512, 1536) the horizontal sums are 512 h, where h =
    s[k-1] + 3 s[k] (even outputs) or 3 s[k] + s[k+1] (odd), edges
    replicated; output row 2k mixes source rows (k-1, k) as (1, 3), row 2k+1
    rows (k, k+1) as (3, 1), rows clamped; out = min(255, ((h0 m0 >> 2) +
    (h1 m1 >> 2) + 2) >> 2)."""
    s = a.astype(np.int64)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, cn = s.shape
    left = np.concatenate([s[:, :1], s[:, :-1]], axis=1)
    right = np.concatenate([s[:, 1:], s[:, -1:]], axis=1)
    hs = np.empty((h, 2 * w, cn), np.int64)
    hs[:, 0::2] = left + 3 * s
    hs[:, 1::2] = 3 * s + right
    out = np.empty((2 * h, 2 * w, cn), np.int64)
    for y in range(2 * h):
        k = y // 2
        if y % 2 == 0:
            r0, r1, m0, m1 = max(k - 1, 0), k, 1, 3
        else:
            r0, r1, m0, m1 = k, min(k + 1, h - 1), 3, 1
        out[y] = np.minimum((((hs[r0] * m0) >> 2) + ((hs[r1] * m1) >> 2) + 2) >> 2, 255)
    return out[:, :, 0] if a.ndim == 2 else out


@pytest.mark.parametrize("cn", [1, 2, 3, 4])
def test_generic_linear_at_exact_2x_is_up2_closed_form(orc, src150, cn):
    """resize.zig:654-680 holds up2Linear == generic LINEAR at 150x64 -> 300x128."""
    a = squeeze(src150[cn])
    want = up2_closed_form(a)
    np.testing.assert_array_equal(orc.resize_linear_generic(a, (300, 128)), want)
    np.testing.assert_array_equal(orc.resize(a, (300, 128), 1), want)


# Destination sizes of the reference test (resize.zig:668), applied per axis.
LINEAR_SIZES = [(75, 32), (300, 128), (113, 47), (120, 90), (112, 48), (37, 20), (20, 9), (9, 5), (199, 70),
                (151, 64), (1, 1), (2, 64), (112, 42), (100, 43), (225, 96), (141, 61), (6, 40)]


def exact_linear_coord(d, ssize, dsize, area):
    """The source coordinate of destination sample d with exact rationals:
    half-pixel centres, or OpenCV's area-mode coordinate when enlarging."""
    if area:
        scale = Fraction(ssize, dsize)
        sx = (d * scale).numerator // (d * scale).denominator
        f = (d + 1) - Fraction(sx + 1) / scale
        f = Fraction(0) if f <= 0 else f - (f.numerator // f.denominator)
        return sx, f
    fx = (d + Fraction(1, 2)) * Fraction(ssize, dsize) - Fraction(1, 2)
    sx = fx.numerator // fx.denominator
    return sx, fx - sx


@pytest.mark.parametrize("area", [False, True])
@pytest.mark.parametrize("pin", [True, False])
def test_linear_taps_quantise_the_exact_coordinate(orc, area, pin):
    """Each 11-bit weight is the exact fractional coordinate times 2048,
    rounded to nearest (within the f32 coordinate's rounding), taps lie on
    floor(coordinate) and its neighbour, and the edges follow the pinning
    rule: along x (pin) an outside tap takes the edge sample with weight
    2048; along y (no pin) the indices are clamped and the weights kept."""
    pairs = [(150, d) for d, _ in LINEAR_SIZES] + [(64, d) for _, d in LINEAR_SIZES] + [(97, 53), (61, 37), (7, 300)]
    for ssize, dsize in pairs:
        if area and dsize <= ssize:
            continue  # area-mode taps are used only when enlarging
        last = ssize - 1
        for d in range(dsize):
            (o0, o1), (w0, w1) = orc.linear_tap(d, ssize, dsize, area, pin)
            sx, f = exact_linear_coord(d, ssize, dsize, area)
            if f < Fraction(1, 10 ** 4) or f > 1 - Fraction(1, 10 ** 4):
                continue  # floor decided by the f32 coordinate's last bit
            if pin and (sx < 0 or sx >= last):
                p = 0 if sx < 0 else last  # the tap moves to the edge sample
                assert (o0, o1, w0, w1) == (p, min(p + 1, last), 2048, 0), (ssize, dsize, d)
                continue
            assert (o0, o1) == (min(max(sx, 0), last), min(max(sx + 1, 0), last)), (ssize, dsize, d)
            # The f32 coordinate (below 2^e, e = bits of ssize) is within half
            # an f32 ulp, 2^(e - 25), of the exact one; its fraction is exact.
            tol = 0.5 + 2048 * 2.0 ** (int(ssize).bit_length() - 25)
            assert abs(w1 - 2048 * f) <= tol, (ssize, dsize, d, w1, float(f))
            assert abs(w0 - 2048 * (1 - f)) <= tol, (ssize, dsize, d, w0, float(f))


# ------------------------------------------------------------------ AREA

AREA_SIZES = [(53, 37), (64, 40), (30, 20), (13, 9), (11, 61), (97, 29), (7, 3), (1, 1), (2, 60), (96, 60)]


def exact_area_table(ssize, dsize):
    """computeResizeAreaTab built from exact rationals: cell dx covers
    [dx s, (dx + 1) s), s = ssize / dsize; its entries are the partial source
    pixel on the left if the sliver exceeds 1e-3, the whole pixels, and the
    partial pixel on the right if that sliver exceeds 1e-3, each weighted by
    its overlap over the cell width (resize.zig:470-493)."""
    s = Fraction(ssize, dsize)
    eps = Fraction(1, 1000)
    out = []
    for dx in range(dsize):
        a, b = dx * s, (dx + 1) * s
        cellw = min(s, ssize - a)
        lo = -((-a.numerator) // a.denominator)  # ceil
        hi = b.numerator // b.denominator        # floor
        lo_c = min(lo, ssize - 1)
        if lo_c - a > eps:
            out.append((dx, lo_c - 1, (lo_c - a) / cellw))
        for sx in range(lo_c, hi):
            out.append((dx, sx, 1 / cellw))
        if b - hi > eps:
            out.append((dx, min(hi, ssize - 1), min(b - hi, 1, cellw) / cellw))
    return out


def area_pairs():
    for dw, dh in AREA_SIZES:
        yield 97, dw
        yield 61, dh
    yield 1001, 500
    yield 1001, 501
    yield 2048, 1365


@pytest.mark.parametrize("ssize,dsize", sorted(set(p for p in area_pairs() if p[1] <= p[0])))
def test_area_table_is_the_exact_cell_partition(orc, ssize, dsize):
    di, si, al = orc.area_tab(ssize, dsize)
    want = exact_area_table(ssize, dsize)
    assert [(int(a), int(b)) for a, b in zip(di, si)] == [(d, s_) for d, s_, _ in want]
    for (d, s_, w), got in zip(want, al):
        # alpha is the f32 rounding of an f64 value within 1e-15 of w.
        assert abs(float(got) - float(w)) <= float(w) * 2.0 ** -23, (d, s_, float(got), float(w))
    # Every kept weight of a cell sums to one less the dropped slivers (each below 1e-3 of a pixel).
    sums = np.zeros(dsize)
    for d, _, w in want:
        sums[d] += float(w)
    s = ssize / dsize
    assert np.all(sums <= 1 + 1e-12) and np.all(sums >= 1 - 2e-3 / s - 1e-12)


def area_matrix(ssize, dsize):
    m = np.zeros((dsize, ssize))
    for d, s_, w in exact_area_table(ssize, dsize):
        m[d, s_] += float(w)
    return m


def exact_area_values(a, dw, dh):
    """The table-weighted averages of areaGeneric, in float64 (within 1e-10
    of the exact rational values)."""
    s = a.astype(np.float64)
    if s.ndim == 2:
        s = s[:, :, None]
    my, mx = area_matrix(s.shape[0], dh), area_matrix(s.shape[1], dw)
    return np.einsum("yi,ijc,xj->yxc", my, s, mx)


@pytest.mark.parametrize("cn", [1, 2, 3, 4])
def test_area_generic_rounds_the_exact_weighted_average(orc, src97, cn):
    """Fractional AREA equals the exact table-weighted average rounded half
    to even wherever that average is more than 1e-3 from a tie (the f32
    sums of areaGeneric stay within 1e-4 of it)."""
    a = squeeze(src97[cn])
    for dw, dh in AREA_SIZES:
        if dw > 97 or dh > 61:
            continue
        got = orc.resize_area_generic(a, (dw, dh)).reshape(dh, dw, cn).astype(np.float64)
        v = exact_area_values(a, dw, dh)
        clear = np.abs(v - np.floor(v) - 0.5) > 1e-3
        assert clear.mean() > 0.97  # cells of two equal weights make exact ties common
        np.testing.assert_array_equal(got[clear], np.rint(v)[clear], err_msg=f"{dw}x{dh}")
        assert np.abs(got - v).max() <= 0.5 + 1e-3
        if (dw, dh) != (96, 60) and not (97 % dw == 0 and 61 % dh == 0):
            np.testing.assert_array_equal(orc.resize(a, (dw, dh), 3).reshape(dh, dw, cn), got)


@pytest.mark.parametrize("f", [2, 4])
def test_area_generic_dyadic_factor_is_mean_rounded_half_even(orc, f):
    """With a power-of-two factor every weight is 1/f (exact in f32) and every
    sum exact, so areaGeneric is the block mean rounded half to even, where
    the exact half path rounds ties up: the two are told apart on ties."""
    a = lcg_fill((60, 96, 3), 99)
    got = orc.resize_area_generic(a, (96 // f, 60 // f))
    mean = a.astype(np.float64).reshape(60 // f, f, 96 // f, f, 3).mean(axis=(1, 3))
    np.testing.assert_array_equal(got, np.rint(mean))
    if f == 2:
        half = orc.resize(a, (48, 30), 3).astype(np.float64)
        ties = mean - np.floor(mean) == 0.5
        assert ties.any()
        np.testing.assert_array_equal(half[ties], np.floor(mean[ties]) + 1)


# ----------------------------------------------------------------- warps

def border_index(p, n, border):
    """border.zig:22-39 restated on integers: -1 for a constant border."""
    if 0 <= p < n:
        return p
    if border == 0:
        return -1
    if border == 1:
        return min(max(p, 0), n - 1)
    if n == 1:
        return 0
    if border == 3:
        return p % n
    period = 2 * n if border == 2 else 2 * n - 2
    q = p % period
    if border == 2:
        return q if q < n else period - 1 - q
    return q if q < n else period - q


def exact_bilinear(a, sx, sy, border, value):
    """Bilinear value at source coordinates (sx, sy) (exact rationals or
    floats), warp semantics of geometry.zig:508-533: with a constant border a
    footprint wholly outside (floor(x) < -1 or >= W, same for y) gives the
    value, other taps outside the image give the value per tap."""
    h, w = a.shape[:2]
    x0, y0 = int(np.floor(sx)), int(np.floor(sy))
    ax, ay = sx - x0, sy - y0
    if border == 0 and (x0 < -1 or y0 < -1 or x0 >= w or y0 >= h):
        return None
    def tap(y, x):
        my, mx = border_index(y, h, border), border_index(x, w, border)
        if my < 0 or mx < 0:
            return np.full(a.shape[2:], value, dtype=object)
        return a[my, mx].astype(object)
    top = tap(y0, x0) * (1 - ax) + tap(y0, x0 + 1) * ax
    bot = tap(y0 + 1, x0) * (1 - ax) + tap(y0 + 1, x0 + 1) * ax
    return top * (1 - ay) + bot * ay


def round_half_even(v):
    """Round a rational (or object array of them) half to even, clamp to u8."""
    def one(q):
        q = Fraction(q)
        f = q.numerator // q.denominator
        r = q - f
        n = f + (1 if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1) else 0)
        return min(max(n, 0), 255)
    return np.vectorize(one, otypes=[np.int64])(v)


def exact_warp(a, inv, size, border, value, persp=False):
    """Warp of `a` by the inverse map `inv` (2x3 affine or 3x3 projective, all
    entries exact rationals), evaluated exactly and rounded half to even."""
    dw, dh = size
    cn = a.shape[2] if a.ndim == 3 else 1
    s = a if a.ndim == 3 else a[:, :, None]
    out = np.zeros((dh, dw, cn), np.int64)
    for y in range(dh):
        for x in range(dw):
            sx = inv[0][0] * x + inv[0][1] * y + inv[0][2]
            sy = inv[1][0] * x + inv[1][1] * y + inv[1][2]
            if persp:
                wv = inv[2][0] * x + inv[2][1] * y + inv[2][2]
                sx, sy = sx / wv, sy / wv
            v = exact_bilinear(s, sx, sy, border, Fraction(value))
            out[y, x] = value if v is None else round_half_even(v)
    return out if a.ndim == 3 else out[:, :, 0]


def frac_matrix(m):
    return [[Fraction(v) for v in row] for row in m]


# Forward matrices whose inverses have dyadic entries with few bits: the
# reference's f32 inverse, coordinates, differences and fused blends are
# then all exact.
DYADIC_AFFINE = [
    [[2, 0, -3], [0, 4, 1]],          # inverse scales 1/2, 1/4 with dyadic shifts
    [[0.5, 0, 0.25], [0, 0.5, -0.75]],  # upscale: every coordinate a half-pixel tie
    [[0, 2, 1.5], [-4, 0, 70]],       # rotation by 90 degrees with scaling
    [[1, 0, 0.25], [0, 1, 0.625]],    # sub-pixel shifts of 1/4 and 5/8
    [[-2, 0, 50], [0, 0.5, 3.25]],    # a mirror
]


@pytest.mark.parametrize("k", range(len(DYADIC_AFFINE)))
@pytest.mark.parametrize("border,cn", [(0, 1), (0, 3), (1, 1), (4, 3), (2, 1)])
def test_warp_affine_dyadic_is_exact_bilinear_rounded_half_even(orc, k, border, cn):
    m = np.array(DYADIC_AFFINE[k], np.float64)
    a = lcg_fill((23, 29, cn), 5 + k)[:, :, 0] if cn == 1 else lcg_fill((23, 29, cn), 5 + k)
    inv = np.linalg.inv(np.vstack([m, [0, 0, 1]]))[:2]
    assert all(float(Fraction(v).limit_denominator(64)) == v for v in inv.ravel())
    got = orc.warp_affine(a, m, (31, 19), 1, border, 7)
    np.testing.assert_array_equal(got, exact_warp(a, frac_matrix(inv), (31, 19), border, 7))


DYADIC_PERSP = [
    [[2, 0, -3], [0, 4, 1], [0, 0, 1]],
    [[1, 0, 0.25], [0, 1, 0.625], [0, 0, 1]],
    [[1, 0, 0.5], [0, 1, -1.5], [0, 0, 2]],    # w = 1/2 everywhere: the division is exact
    [[0, 1, 0.5], [-1, 0, 30.25], [0, 0, 0.25]],
]


@pytest.mark.parametrize("k", range(len(DYADIC_PERSP)))
@pytest.mark.parametrize("border,cn", [(0, 1), (0, 3), (4, 1), (1, 3)])
def test_warp_perspective_dyadic_is_exact_bilinear_rounded_half_even(orc, k, border, cn):
    m = np.array(DYADIC_PERSP[k], np.float64)
    a = lcg_fill((23, 29, cn), 15 + k)[:, :, 0] if cn == 1 else lcg_fill((23, 29, cn), 15 + k)
    inv = np.linalg.inv(m)
    assert all(float(Fraction(v).limit_denominator(64)) == v for v in inv.ravel())
    got = orc.warp_perspective(a, m, (31, 19), 1, border, 9)
    np.testing.assert_array_equal(got, exact_warp(a, frac_matrix(inv), (31, 19), border, 9, persp=True))


def test_remap_dyadic_maps_are_exact_bilinear_rounded_half_even(orc):
    """remap (undistort.zig:280-394, f32 maps, fused blend, magic-number
    rounding) on maps with quarter- and eighth-pixel coordinates, some
    outside the image: every operation is exact."""
    a = lcg_fill((21, 27, 3), 31)
    yy, xx = np.mgrid[0:17, 0:33].astype(np.float64)
    mx = (xx * 0.75 - 2.125 + (yy % 3) * 0.25).astype(np.float32)
    my = (yy * 1.25 - 1.5 + (xx % 4) * 0.125).astype(np.float32)
    got = orc.remap(a, mx, my, value=13)
    want = np.zeros_like(got, dtype=np.int64)
    h, w = a.shape[:2]
    for y in range(17):
        for x in range(33):
            sx, sy = Fraction(float(mx[y, x])), Fraction(float(my[y, x]))
            x0, y0 = sx.numerator // sx.denominator, sy.numerator // sy.denominator
            if not (x0 >= -1 and y0 >= -1 and x0 < w and y0 < h):
                want[y, x] = 13
                continue
            v = exact_bilinear(a, sx, sy, 0, Fraction(13))
            want[y, x] = round_half_even(v)
    np.testing.assert_array_equal(got, want)


def float64_warp(a, inv, size, persp):
    """The float64 bilinear value of every destination pixel with constant
    border 0, the coordinates, and whether the footprint was wholly outside."""
    dw, dh = size
    yy, xx = np.mgrid[0:dh, 0:dw].astype(np.float64)
    sx = inv[0, 0] * xx + inv[0, 1] * yy + inv[0, 2]
    sy = inv[1, 0] * xx + inv[1, 1] * yy + inv[1, 2]
    if persp:
        wv = inv[2, 0] * xx + inv[2, 1] * yy + inv[2, 2]
        sx, sy = sx / wv, sy / wv
    h, w = a.shape[:2]
    x0, y0 = np.floor(sx).astype(np.int64), np.floor(sy).astype(np.int64)
    ax, ay = sx - x0, sy - y0
    s = a.astype(np.float64)
    def tap(yi, xi):
        ok = (yi >= 0) & (yi < h) & (xi >= 0) & (xi < w)
        return np.where(ok, s[np.clip(yi, 0, h - 1), np.clip(xi, 0, w - 1)], 0.0)
    v = (tap(y0, x0) * (1 - ax) + tap(y0, x0 + 1) * ax) * (1 - ay) + \
        (tap(y0 + 1, x0) * (1 - ax) + tap(y0 + 1, x0 + 1) * ax) * ay
    outside = (x0 < -1) | (y0 < -1) | (x0 >= w) | (y0 >= h)
    return np.where(outside, 0.0, v), sx, sy


def rounding_is_clear(v, sx, sy, margin):
    near_tie = np.abs(v - np.floor(v) - 0.5) <= margin
    near_int = (np.abs(sx - np.rint(sx)) <= 2e-3) | (np.abs(sy - np.rint(sy)) <= 2e-3)
    return ~near_tie & ~near_int


@pytest.mark.parametrize("seed", range(4))
def test_warp_affine_general_rounds_the_float64_bilinear(orc, seed):
    """The benchmark's seeded matrices (rotation +-30 deg, scale 0.8-1.2,
    shift +-10 %) on a 320x180 frame: the f32 chain gives round(v) of the
    float64 bilinear v at every pixel not within 0.02 of a tie or 2e-3 px
    of a tap boundary (the f32 coordinates carry at most ~1e-4 px of error,
    which moves v by less than 0.02); every pixel within 0.5 + 0.02."""
    from visar_amd import synth
    w, h = 320, 180
    a = lcg_fill((h, w), 100 + seed)
    m = synth.affine_matrix(seed, w, h)
    inv = np.linalg.inv(np.vstack([m, [0, 0, 1]]))
    got = orc.warp_affine(a, m, (w, h), 1, 0, 0).astype(np.float64)
    v, sx, sy = float64_warp(a, inv, (w, h), False)
    clear = rounding_is_clear(v, sx, sy, 0.02)
    assert clear.mean() > 0.9
    np.testing.assert_array_equal(got[clear], np.rint(v[clear]))
    near = (np.abs(sx - np.rint(sx)) > 2e-3) & (np.abs(sy - np.rint(sy)) > 2e-3)
    assert np.abs(got - v)[near].max() <= 0.52


@pytest.mark.parametrize("seed", range(4))
def test_warp_perspective_general_rounds_the_float64_bilinear(orc, seed):
    """The benchmark's seeded perspective (corners jittered +-5 %): f64
    coordinates as in the reference, f32 weights and non-fused sums, whose
    error stays below 0.01; round(v) wherever v is 0.02 clear of a tie."""
    from visar_amd import synth
    w, h = 320, 180
    a = lcg_fill((h, w), 200 + seed)
    m = synth.perspective_matrix(seed, w, h)
    inv = np.linalg.inv(m)
    got = orc.warp_perspective(a, m, (w, h), 1, 0, 0).astype(np.float64)
    v, sx, sy = float64_warp(a, inv, (w, h), True)
    clear = rounding_is_clear(v, sx, sy, 0.02)
    assert clear.mean() > 0.9
    np.testing.assert_array_equal(got[clear], np.rint(v[clear]))


def test_remap_general_rounds_the_float64_bilinear(orc):
    """A smooth non-dyadic map (a radial distortion) in f32: the fused f32
    blend stays within 1e-4 of the float64 bilinear value at the same f32
    coordinates, so the output is round(v) wherever v is 0.01 clear of a tie."""
    w, h = 240, 160
    a = lcg_fill((h, w), 301)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    cx, cy = w / 2 + 0.3, h / 2 - 0.7
    r2 = ((xx - cx) ** 2 + (yy - cy) ** 2) / (w * w)
    k = 1 + 0.21 * r2 - 0.05 * r2 * r2
    mx = (cx + (xx - cx) * k).astype(np.float32)
    my = (cy + (yy - cy) * k).astype(np.float32)
    got = orc.remap(a, mx, my).astype(np.float64)
    sx, sy = mx.astype(np.float64), my.astype(np.float64)
    x0, y0 = np.floor(sx).astype(np.int64), np.floor(sy).astype(np.int64)
    ax, ay = sx - x0, sy - y0
    s = a.astype(np.float64)
    def tap(yi, xi):
        ok = (yi >= 0) & (yi < h) & (xi >= 0) & (xi < w)
        return np.where(ok, s[np.clip(yi, 0, h - 1), np.clip(xi, 0, w - 1)], 0.0)
    v = (tap(y0, x0) * (1 - ax) + tap(y0, x0 + 1) * ax) * (1 - ay) + \
        (tap(y0 + 1, x0) * (1 - ax) + tap(y0 + 1, x0 + 1) * ax) * ay
    outside = ~((x0 >= -1) & (y0 >= -1) & (x0 < w) & (y0 < h))
    v = np.where(outside, 0.0, v)
    clear = np.abs(v - np.floor(v) - 0.5) > 0.01
    assert clear.mean() > 0.97
    np.testing.assert_array_equal(got[clear], np.rint(v[clear]))


# ------------------------------------------- generic resize (any depth)

def exact_linear_coord_clamped(d, ssize, dsize, area):
    """linearCoord (resize.zig:154-180) with exact rationals: both axes
    clamp to the edges with weight 0."""
    if area:
        scale = Fraction(ssize, dsize)
        sx = (d * scale).numerator // (d * scale).denominator
        f = (d + 1) - Fraction(sx + 1) / scale
        f = Fraction(0) if f <= 0 else f - (f.numerator // f.denominator)
    else:
        fx = (d + Fraction(1, 2)) * Fraction(ssize, dsize) - Fraction(1, 2)
        sx = fx.numerator // fx.denominator
        f = fx - sx
    if sx < 0:
        sx, f = 0, Fraction(0)
    if sx >= ssize - 1:
        sx, f = ssize - 1, Fraction(0)
    return sx, f


def linear64(a, dw, dh, area=False):
    """Float64 bilinear with linearCoord's taps on both axes."""
    s = a.astype(np.float64)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w = s.shape[:2]
    xs = [exact_linear_coord_clamped(x, w, dw, area) for x in range(dw)]
    ys = [exact_linear_coord_clamped(y, h, dh, area) for y in range(dh)]
    x0 = np.array([i for i, _ in xs]); x1 = np.minimum(x0 + 1, w - 1)
    fx = np.array([float(f) for _, f in xs])[None, :, None]
    y0 = np.array([i for i, _ in ys]); y1 = np.minimum(y0 + 1, h - 1)
    fy = np.array([float(f) for _, f in ys])[:, None, None]
    top = s[y0][:, x0] * (1 - fx) + s[y0][:, x1] * fx
    bot = s[y1][:, x0] * (1 - fx) + s[y1][:, x1] * fx
    v = top * (1 - fy) + bot * fy
    return v[:, :, 0] if a.ndim == 2 else v


@pytest.mark.parametrize("dtype", [np.uint16, np.int16, np.float32])
@pytest.mark.parametrize("cn", [1, 3])
def test_generic_linear_2x_exact_on_integer_samples(orc, dtype, cn):
    """Bilinear 2x with integer-valued samples: every weight is 0, 1/4, 3/4
    or 1 and every product and sum exact in f32, so the output is the exact
    bilinear value (rounded half to even for integer samples)."""
    r = np.random.default_rng(40 + cn)
    lo, hi = (0, 4000) if dtype == np.uint16 else (-2000, 2000)
    a = r.integers(lo, hi, (23, 31, cn)).astype(dtype)
    a = a[:, :, 0] if cn == 1 else a
    got = orc.resize(a, (62, 46), 1)
    v = linear64(a, 62, 46)
    want = v.astype(np.float32) if dtype == np.float32 else np.rint(v).astype(dtype)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.uint16, np.int16, np.float32])
@pytest.mark.parametrize("size,area", [((53, 37), False), ((150, 90), False), ((40, 61), False), ((150, 90), True),
                                       ((101, 61), True)])
def test_generic_linear_rounds_the_float64_bilinear(orc, dtype, size, area):
    """LINEAR (and AREA enlarging, area-style weights) on u16 / i16 / f32 at
    general scales: f32 within 1e-5 relative of the float64 bilinear of
    linearCoord's taps; integers round(v) wherever v is 0.01 clear of a tie."""
    r = np.random.default_rng(7)
    if dtype == np.float32:
        a = r.random((61, 97, 3), dtype=np.float32) * 100
    else:
        info = np.iinfo(dtype)
        a = r.integers(info.min // 2, info.max // 2, (61, 97, 3)).astype(dtype)
    got = orc.resize(a, size, 3 if area else 1).astype(np.float64)
    v = linear64(a, size[0], size[1], area)
    if dtype == np.float32:
        np.testing.assert_allclose(got, v, rtol=1e-5, atol=1e-5)
    else:
        clear = np.abs(v - np.floor(v) - 0.5) > 0.01
        assert clear.mean() > 0.95
        np.testing.assert_array_equal(got[clear], np.rint(v)[clear])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.float32])
def test_generic_area_integer_factors(orc, dtype):
    """areaInt (resize.zig:420-457) on the generic path (u8 with 5 channels,
    u16, i16, f32): unsigned samples the block mean rounded half up (u8 2x2
    as (a + b + c + d + 2) >> 2), signed and float the f64 mean rounded half
    to even / to f32; integer-valued f32 samples keep every sum exact."""
    r = np.random.default_rng(11)
    cn = 5 if dtype == np.uint8 else 3
    if dtype == np.float32:
        a = r.integers(-5000, 5000, (36, 60, cn)).astype(np.float32)
    else:
        info = np.iinfo(dtype)
        a = r.integers(info.min, info.max, (36, 60, cn), endpoint=True).astype(dtype)
    for fx, fy in [(2, 2), (3, 2), (4, 3), (5, 6)]:
        got = orc.resize(a, (60 // fx, 36 // fy), 3)
        blocks = a.astype(np.int64).reshape(36 // fy, fy, 60 // fx, fx, cn).sum(axis=(1, 3))
        n = fx * fy
        if dtype in (np.uint8, np.uint16):
            want = (blocks + n // 2) // n
        elif dtype == np.int16:
            want = np.rint(blocks / n)
        else:
            want = (blocks / n).astype(np.float32)
        np.testing.assert_array_equal(got, want.astype(dtype), err_msg=f"{fx}x{fy}")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.float32])
def test_generic_area_fractional_rounds_the_exact_average(orc, dtype):
    """areaGeneric on the generic path: the exact table-weighted average,
    rounded half to even for integer samples wherever it is clear (by the
    f32 sums' rounding) clear of a tie, within 1e-5 relative for f32."""
    r = np.random.default_rng(12)
    cn = 5 if dtype == np.uint8 else 3
    if dtype == np.float32:
        a = r.random((61, 97, cn), dtype=np.float32) * 50
    else:
        info = np.iinfo(dtype)
        a = r.integers(max(info.min, -30000), min(info.max, 30000), (61, 97, cn)).astype(dtype)
    for dw, dh in [(53, 37), (30, 20), (13, 9), (97, 29), (7, 3)]:
        got = orc.resize(a, (dw, dh), 3).astype(np.float64)
        v = exact_area_values(a, dw, dh)
        if dtype == np.float32:
            np.testing.assert_allclose(got, v, rtol=1e-5, atol=1e-4)
            continue
        # f32 sums of up to ~10 terms of magnitude M carry about 10 M 2^-24 of error.
        tol = 1e-3 + 10 * float(np.abs(v).max()) * 2.0 ** -24
        clear = np.abs(v - np.floor(v) - 0.5) > tol
        assert clear.mean() > 0.9
        np.testing.assert_array_equal(got[clear], np.rint(v)[clear], err_msg=f"{dw}x{dh}")


def test_generic_u8_many_channels_linear_is_the_11_bit_pipeline(orc):
    """u8 with more than 4 channels takes the generic LINEAR kernel, which for
    u8 keeps linearTapU8's 11-bit taps: per channel it equals the fast path
    on that channel alone, and at the exact half it is the block mean."""
    r = np.random.default_rng(13)
    a = r.integers(0, 256, (64, 150, 6), dtype=np.uint8)
    for size in [(75, 32), (113, 47), (300, 128), (37, 20)]:
        got = orc.resize(a, size, 1)
        for c in range(6):
            np.testing.assert_array_equal(got[:, :, c], orc.resize(np.ascontiguousarray(a[:, :, c]), size, 1))
    s = a.astype(np.int32)
    half = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(orc.resize(a, (75, 32), 1), half)
