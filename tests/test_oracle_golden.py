"""Pins the CPU oracle (CPU only).

Each check is independent of the oracle's own code: the reference's unit-test
known answers (tests/golden/known_answers.json), brute-force integer
convolution and per-tap morphology in NumPy, closed forms (the Background
model of the reference's tests/py/test_motion_vs_opencv.py:191-197), exact
geometric identities and float64 references within the reference's own
tolerances.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))

PAD = {0: "constant", 1: "edge", 2: "symmetric", 3: "wrap", 4: "reflect"}


def rnd(shape, seed, dtype=np.uint8):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        return rng.uniform(-100, 300, shape).astype(np.float32)
    info = np.iinfo(dtype)
    return rng.integers(info.min, info.max + 1, shape, dtype=np.int64).astype(dtype)


# ----------------------------------------------------------- core pins

def test_border_map(orc):
    for p, n, b, want in KA["border_map"]["cases"]:
        assert orc.border_map(p, n, b) == want


def test_border_map_matches_numpy_pad(orc):
    # numpy's pad modes restate the same extrapolations, including pads
    # wider than the image.
    for b, mode in PAD.items():
        if b == 0:
            continue
        for n in (1, 2, 3, 5, 8):
            if n == 1 and b in (2, 4):
                continue
            idx = np.pad(np.arange(n), 11, mode=mode)
            got = [orc.border_map(p, n, b) for p in range(-11, n + 11)]
            assert got == idx.tolist(), (b, n)


def test_saturate(orc):
    lib = orc.lib()
    for x, want in KA["sat_cast"]["u8_from_f32"]:
        assert lib.orc_sat_u8_f32(x) == want
    for x, want in KA["sat_cast"]["i16_from_f32"]:
        assert lib.orc_sat_i16_f32(x) == want
    vals = [float("nan") if v == "nan" else v for v in KA["satF32"]["in"]]
    assert [lib.orc_satf32_u8(v) for v in vals] == KA["satF32"]["u8"]
    s = KA["satF32"]["i16_scale"]
    assert [lib.orc_satf32_i16(np.float32(v) * np.float32(s)) for v in vals] == KA["satF32"]["i16"]


def test_gaussian_kernels_known_answers(orc):
    g = KA["gaussian_kernel"]
    k = orc.gaussian_kernel(3, 0.9)
    bits = [hex(struct.unpack("<Q", struct.pack("<d", v))[0]) for v in k[:2]]
    assert bits == g["k3_s0p9_bits"]
    assert orc.gaussian_kernel_fixed(13, 0, 8).tolist() == g["fixed8_k13"]
    assert orc.gaussian_kernel_fixed(13, 0, 16).tolist() == g["fixed16_k13"]
    assert orc.gaussian_kernel_fixed(9, 0, 8).tolist() == g["fixed8_k9"]
    for n in range(1, 64, 2):
        for s in (0, 0.6, 1.3, 4.1):
            assert orc.gaussian_kernel_fixed(n, s, 8).sum() == 256
            assert orc.gaussian_kernel_fixed(n, s, 16).sum() == 65536


def test_cv_exp_close_to_exp(orc):
    for x in np.linspace(-40, 5, 997):
        want = math.exp(x)
        assert abs(orc.cv_exp(x) - want) <= 1e-14 * want
    assert orc.cv_exp(0.0) == 1.0


def test_deriv_kernels(orc):
    for order, ksize, want in KA["deriv_kernel"]["cases"]:
        assert orc.deriv_kernel(order, ksize).tolist() == want
    assert orc.deriv_kernel(0, -1).tolist() == [3, 10, 3]


# ---------------------------------------------------------- filters

def test_gaussian_impulse(orc):
    g = KA["gaussian_impulse_3x3"]
    a = np.zeros(g["size"], np.uint8)
    a[g["impulse"][1], g["impulse"][0]] = g["impulse"][2]
    d = orc.gaussian_blur(a, (3, 3), 0)
    for x, y, v in g["expect"]:
        assert d[y, x] == v


def conv2d_int(a, kx, ky, border):
    """Exact 2-D correlation with border extrapolation (channels last)."""
    rx, ry = len(kx) // 2, len(ky) // 2
    a3 = a if a.ndim == 3 else a[:, :, None]
    p = np.pad(a3.astype(np.int64), ((ry, ry), (rx, rx), (0, 0)), mode=PAD[border])
    h, w = a3.shape[:2]
    out = np.zeros(a3.shape, np.int64)
    for j, kyv in enumerate(ky):
        for i, kxv in enumerate(kx):
            out += int(kyv) * int(kxv) * p[j:j + h, i:i + w]
    return out if a.ndim == 3 else out[:, :, 0]


@pytest.mark.parametrize("border", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("cn", [1, 3])
def test_gaussian_u8_is_exact_convolution(orc, border, cn):
    a = rnd((37, 53, cn) if cn > 1 else (37, 53), 1 + border)
    for k, s in [(3, 0), (5, 0), (7, 0), (9, 0), (13, 0), (11, 1.3), (21, 2.5)]:
        taps = orc.gaussian_kernel_fixed(k, s, 8)
        want = np.minimum((conv2d_int(a, taps, taps, border) + (1 << 15)) >> 16, 255)
        got = orc.gaussian_blur(a, (k, k), s, border=border)
        np.testing.assert_array_equal(got, want, err_msg=f"k{k} s{s}")


def test_gaussian_u16_is_exact_convolution(orc):
    a = rnd((29, 31), 5, np.uint16)
    for k, s in [(3, 0), (7, 0), (15, 2.0)]:
        taps = orc.gaussian_kernel_fixed(k, s, 16)
        want = np.minimum((conv2d_int(a, taps, taps, 4) + (1 << 31)) >> 32, 65535)
        np.testing.assert_array_equal(orc.gaussian_blur(a, (k, k), s), want)


def test_gaussian_asymmetric_sizes(orc):
    a = rnd((23, 41), 9)
    for kw, kh, sx, sy in [(3, 15, 0, 0), (21, 5, 0, 0), (11, 7, 1.4, 3.2), (1, 9, 0, 0)]:
        kx = orc.gaussian_kernel_fixed(kw, sx, 8)
        ky = orc.gaussian_kernel_fixed(kh, sy if sy > 0 else sx, 8)
        want = np.minimum((conv2d_int(a, kx, ky, 4) + (1 << 15)) >> 16, 255)
        np.testing.assert_array_equal(orc.gaussian_blur(a, (kw, kh), sx, sy), want)


def test_gaussian_f32_close_to_float64(orc):
    a = rnd((33, 47), 3, np.float32)
    for k, s in [(3, 0), (5, 0), (9, 2.0), (15, 0)]:
        kd = orc.gaussian_kernel(k, s)
        want = conv2d_float(a.astype(np.float64), kd, kd, 4)
        got = orc.gaussian_blur(a, (k, k), s)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-3)


def conv2d_float(a, kx, ky, border):
    rx, ry = len(kx) // 2, len(ky) // 2
    p = np.pad(a, ((ry, ry), (rx, rx)), mode=PAD[border])
    h, w = a.shape
    out = np.zeros(a.shape)
    for j in range(len(ky)):
        for i in range(len(kx)):
            out += float(ky[j]) * float(kx[i]) * p[j:j + h, i:i + w]
    return out


@pytest.mark.parametrize("dx,dy,ks", [(1, 0, 3), (0, 1, 3), (1, 1, 3), (1, 0, 5), (2, 0, 5), (1, 0, -1), (1, 0, 7)])
def test_sobel_is_exact_convolution(orc, dx, dy, ks):
    a = rnd((31, 45), 11)
    kx = orc.deriv_kernel(dx, ks)
    ky = orc.deriv_kernel(dy, ks)
    full = conv2d_int(a, kx, ky, 4)
    np.testing.assert_array_equal(orc.sobel(a, np.int16, dx, dy, ks), np.clip(full, -32768, 32767))
    np.testing.assert_array_equal(orc.sobel(a, np.float32, dx, dy, ks), full.astype(np.float32))


def test_sep_filter_f32_close_to_float64(orc):
    a = rnd((21, 35), 4).astype(np.float32)
    kx = np.array([0.1, -0.7, 0.3, 1.25, 0.05], np.float32)
    ky = np.array([0.2, 0.5, -0.15], np.float32)
    want = conv2d_float(a.astype(np.float64), kx, ky, 1)
    got = orc.sep_filter_2d(a, np.float32, kx, ky, border=1)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)
    # a transposed pair must differ (pins kx horizontal, ky vertical)
    assert not np.allclose(orc.sep_filter_2d(a, np.float32, ky, kx, border=1), want, atol=1e-3)


# ------------------------------------------------------- color, threshold

def test_gray_bt601_primaries(orc):
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [0, 0, 0]]], np.uint8)
    # BT.601: 0.114 B + 0.587 G + 0.299 R
    assert orc.cvt_color(px, 6).ravel().tolist() == [29, 150, 76, 255, 0]
    assert orc.cvt_color(px, 7).ravel().tolist() == [76, 150, 29, 255, 0]


def test_gray_lattice(orc):
    v = np.minimum(np.arange(52) * 5, 255)
    b, g, r = np.meshgrid(v, v, v, indexing="ij")
    img = np.stack([b.ravel(), g.ravel(), r.ravel()], -1).astype(np.uint8)[None]
    want = (b.ravel() * 3735 + g.ravel() * 19235 + r.ravel() * 9798 + 16384) >> 15
    assert np.array_equal(orc.cvt_color(img, 6)[0], want)
    exact = 0.114 * b.ravel() + 0.587 * g.ravel() + 0.299 * r.ravel()
    assert np.abs(orc.cvt_color(img, 6)[0] - exact).max() <= 0.5 + 1e-9


# frame.zig grayDownsample (the tracker's producer of the path's input)

GD_W = np.array([1868, 9617, 4899], np.int64)  # frame.zig:186-189, B, G, R


def _block_sums(a, k):
    h, w = a.shape[0] // k, a.shape[1] // k
    b = a[:h * k, :w * k, :3].astype(np.int64).reshape(h, k, w, k, 3)
    return b.sum(axis=(1, 3))


@pytest.mark.parametrize("cn", [4, 3])
def test_gray_downsample_matches_gray_then_area(orc, cn):
    """frame.zig:384-416: within 0.51 of the exact mean of the full-size lumas,
    and within 1 of cvtColor's gray at k = 1."""
    for k in (1, 2, 3, 4, 6, 8):
        a = rnd((9 * k, 70 * k + 1, cn), k)
        got = orc.gray_downsample(a, k).astype(np.float64)
        want = (_block_sums(a, k) @ GD_W) / 16384.0 / (k * k)
        assert np.abs(got - want).max() <= 0.51
        if k == 1:
            gray = orc.cvt_color(a, 6 if cn == 3 else 10)
            assert np.abs(gray.astype(np.int64) - got.astype(np.int64)).max() <= 1


def test_gray_downsample_rounding(orc):
    """k = 2 is (sum + 2^15) >> 16 exactly; other k round the exact mean half
    up except within 1e-3 of a half (the f32 reciprocal's error is below 1e-4),
    and never exceed 255; constant images stay constant."""
    for cn in (3, 4):
        for k in (2, 3, 5, 7, 16):
            a = rnd((5 * k + 3, 23 * k + 2, cn), 100 + k + cn)
            got = orc.gray_downsample(a, k).astype(np.int64)
            l = _block_sums(a, k) @ GD_W
            if k == 2:
                assert np.array_equal(got, (l + (1 << 15)) >> 16)
            else:
                exact = l / float(k * k * 16384) + 0.5
                frac = exact - np.floor(exact)
                ok = np.minimum(frac, 1.0 - frac) > 1e-3
                assert np.array_equal(got[ok], np.minimum(255, np.floor(exact[ok])).astype(np.int64))
            for v in (0, 1, 128, 254, 255):
                assert np.unique(orc.gray_downsample(np.full((2 * k, 3 * k, cn), v, np.uint8), k)).tolist() == [v]


def test_gray_downsample_rect_equals_window(orc):
    """frame.zig:418-446: the window of the whole-image result, the rest untouched."""
    rng = np.random.default_rng(5)
    for cn in (4, 3):
        for k in (1, 2, 3, 5, 6, 8):
            a = rnd((23 * k + 1, 61 * k + 3, cn), k * 7 + cn)
            want = orc.gray_downsample(a, k)
            for _ in range(10):
                x, y = int(rng.integers(0, want.shape[1])), int(rng.integers(0, want.shape[0]))
                w, h = 1 + int(rng.integers(0, want.shape[1] - x)), 1 + int(rng.integers(0, want.shape[0] - y))
                got = np.zeros_like(want)
                orc.gray_downsample_rect(a, k, got, x, y, w, h)
                inside = np.zeros(want.shape, bool)
                inside[y:y + h, x:x + w] = True
                assert np.array_equal(got[inside], want[inside]) and not got[~inside].any()


def test_gray_downsample_rejects(orc):
    a = rnd((8, 8, 3), 1)
    for k, shape in [(0, (8, 8)), (17, (0, 0)), (2, (4, 5)), (2, (3, 4))]:
        with pytest.raises(Exception):
            orc.gray_downsample(a, k, np.zeros(shape if shape[0] else (1, 1), np.uint8))
    with pytest.raises(Exception):
        orc.gray_downsample(rnd((8, 8), 1), 2)


def test_threshold_known(orc):
    t = KA["threshold"]
    a = np.array([t["in"]], np.uint8)
    _, d = orc.threshold(a, t["thresh"], t["maxval"], 0)
    assert d.ravel().tolist() == t["binary"]


def test_threshold_kinds_match_numpy(orc):
    a = rnd((17, 19, 3), 2)
    for kind in range(5):
        _, d = orc.threshold(a, 127.5, 200, kind)
        gt = a > 127
        want = [np.where(gt, 200, 0), np.where(gt, 0, 200), np.where(gt, 127, a), np.where(gt, a, 0),
                np.where(gt, 0, a)][kind]
        np.testing.assert_array_equal(d, want)


# ------------------------------------------------------------- resize

def test_resize_constant(orc):
    c = KA["resize_constant"]
    a = np.full((c["src"][1], c["src"][0], c["src"][2]), c["src"][3], np.uint8)
    for interp in (0, 1, 3):
        for (w, h) in [(7, 13), (5, 5), (20, 3), (3, 20)]:
            assert (orc.resize(a, (w, h), interp) == c["src"][3]).all()


def test_resize_half_and_area_int(orc):
    a = rnd((40, 62, 3), 7)
    want = (a[0::2, 0::2].astype(np.int32) + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2] + 2) >> 2
    np.testing.assert_array_equal(orc.resize(a, (31, 20), 1), want)
    np.testing.assert_array_equal(orc.resize(a, (31, 20), 3), want)
    b = rnd((60, 96, 1), 8)[:, :, 0]
    s = b.reshape(20, 3, 24, 4).astype(np.int64).sum(axis=(1, 3))
    want = np.rint(s.astype(np.float32) * np.float32(1.0 / 12.0))
    np.testing.assert_array_equal(orc.resize(b, (24, 20), 3), want)


def bilinear64(a, dw, dh):
    h, w = a.shape[:2]
    def coords(d, s):
        f = (np.arange(d) + 0.5) * (s / d) - 0.5
        i0 = np.floor(f).astype(int)
        t = f - i0
        return np.clip(i0, 0, s - 1), np.clip(i0 + 1, 0, s - 1), t
    x0, x1, tx = coords(dw, w)
    tx = np.where((np.floor((np.arange(dw) + 0.5) * (w / dw) - 0.5) < 0) |
                  (np.floor((np.arange(dw) + 0.5) * (w / dw) - 0.5) >= w - 1), 0, tx)
    y0, y1, ty = coords(dh, h)
    a = a.astype(np.float64)
    top = a[y0][:, x0] * (1 - tx)[None, :, None] + a[y0][:, x1] * tx[None, :, None]
    bot = a[y1][:, x0] * (1 - tx)[None, :, None] + a[y1][:, x1] * tx[None, :, None]
    return top * (1 - ty)[:, None, None] + bot * ty[:, None, None]


def test_resize_linear_within_one_of_bilinear(orc):
    a = rnd((61, 97, 3), 12)
    for dw, dh in [(53, 37), (150, 90), (113, 71), (97, 30), (40, 61)]:
        got = orc.resize(a, (dw, dh), 1).astype(np.float64)
        want = bilinear64(a, dw, dh)
        assert np.abs(got - want).max() <= 1.0, (dw, dh)


def area64(a, dw, dh):
    h, w = a.shape[:2]
    out = np.zeros((dh, dw, a.shape[2]))
    sx, sy = w / dw, h / dh
    for y in range(dh):
        for x in range(dw):
            x0, x1, y0, y1 = x * sx, (x + 1) * sx, y * sy, (y + 1) * sy
            acc = 0
            for yy in range(int(np.floor(y0)), int(np.ceil(y1))):
                wy = min(y1, yy + 1) - max(y0, yy)
                for xx in range(int(np.floor(x0)), int(np.ceil(x1))):
                    wx = min(x1, xx + 1) - max(x0, xx)
                    acc = acc + wx * wy * a[min(yy, h - 1), min(xx, w - 1)].astype(np.float64)
            out[y, x] = acc / (sx * sy)
    return out


def test_resize_area_fractional_within_one(orc):
    a = rnd((61, 97, 3), 13)
    for dw, dh in [(53, 37), (30, 20), (13, 9), (97, 29), (7, 3)]:
        got = orc.resize(a, (dw, dh), 3).astype(np.float64)
        assert np.abs(got - area64(a, dw, dh)).max() <= 1.0, (dw, dh)


# ------------------------------------------------------------- warps

def test_warp_affine_identities(orc):
    a = rnd((40, 40, 3), 21)
    eye = np.array([[1, 0, 0], [0, 1, 0]], np.float64)
    for interp in (0, 1):
        for border in range(5):
            np.testing.assert_array_equal(orc.warp_affine(a, eye, (40, 40), interp, border), a)
    # translation by (3, -2) with a constant border of 7
    t = np.array([[1, 0, 3], [0, 1, -2]], np.float64)
    d = orc.warp_affine(a, t, (40, 40), 1, 0, 7)
    np.testing.assert_array_equal(d[:38, 3:], a[2:, :37])
    assert (d[:, :3] == 7).all() and (d[38:] == 7).all()
    # rotation by 90 degrees: dst(x, y) = src(y, 39 - x), i.e. np.rot90
    r = np.array([[0, 1, 0], [-1, 0, 39]], np.float64)
    np.testing.assert_array_equal(orc.warp_affine(a, r, (40, 40), 1, 4), np.rot90(a))


def test_warp_affine_half_pixel_rounds_half_even(orc):
    a = rnd((9, 30), 22)
    m = np.array([[1, 0, 0.5], [0, 1, 0]], np.float64)  # dst(x) = src(x - 0.5)
    d = orc.warp_affine(a, m, (30, 9), 1, 1)
    mid = (a[:, :-1].astype(np.float64) + a[:, 1:]) / 2
    np.testing.assert_array_equal(d[:, 1:], np.rint(mid))  # np.rint is half to even
    dn = orc.warp_affine(a, m, (30, 9), 0, 1)  # nearest, rint(x - 0.5): ties to even
    xs = np.rint(np.arange(30) - 0.5).astype(int).clip(0, 29)
    np.testing.assert_array_equal(dn, a[:, xs])


def test_warp_perspective_identities(orc):
    a = rnd((33, 45), 23)
    np.testing.assert_array_equal(orc.warp_perspective(a, np.eye(3), (45, 33), 1, 4), a)
    t = np.array([[1, 0, -4], [0, 1, 5], [0, 0, 1]], np.float64)
    d = orc.warp_perspective(a, t, (45, 33), 1, 0, 9)
    np.testing.assert_array_equal(d[5:, :41], a[:28, 4:])
    # nearest uses round-half-away: +0.5 shift maps x to x (x >= 1), x = 0 to -1 (border)
    h = np.array([[1, 0, 0.5], [0, 1, 0], [0, 0, 1]], np.float64)
    dn = orc.warp_perspective(a, h, (45, 33), 0, 0, 3)
    np.testing.assert_array_equal(dn[:, 1:], a[:, 1:])
    assert (dn[:, 0] == 3).all()
    # a pure scaling by 2 (w = 1) equals the affine scaling on integer grids
    s = np.array([[2, 0, 0], [0, 2, 0], [0, 0, 1]], np.float64)
    np.testing.assert_array_equal(orc.warp_perspective(a, s, (45, 33), 1, 0),
                                  orc.warp_affine(a, s[:2], (45, 33), 1, 0))


def test_remap_identity_and_shift(orc):
    a = rnd((20, 30, 3), 24)
    yy, xx = np.mgrid[0:20, 0:30].astype(np.float32)
    np.testing.assert_array_equal(orc.remap(a, xx, yy), a)
    d = orc.remap(a, xx + np.float32(0.5), yy, value=11)
    mid = (a[:, :-1].astype(np.float64) + a[:, 1:]) / 2
    np.testing.assert_array_equal(d[:, :-1], np.rint(mid))
    out = orc.remap(a, xx + 100, yy, value=11.6)
    assert (out == 12).all()


# --------------------------------------------------------- morphology

def test_structuring_elements(orc):
    e = orc.structuring_element(2, 5, 5)
    assert e.tolist() == [[0, 0, 1, 0, 0], [1, 1, 1, 1, 1], [1, 1, 1, 1, 1], [1, 1, 1, 1, 1], [0, 0, 1, 0, 0]]
    c = orc.structuring_element(1, 7, 7)
    assert c.sum() == 13 and c[3].all() and c[:, 3].all()
    assert orc.structuring_element(0, 3, 5).all()


def morph_np(a, mask, op_max):
    kh, kw = mask.shape
    ay, ax = kh // 2, kw // 2
    a3 = a if a.ndim == 3 else a[:, :, None]
    ident = 0 if op_max else 255
    p = np.pad(a3, ((ay, kh - 1 - ay), (ax, kw - 1 - ax), (0, 0)), constant_values=ident)
    h, w = a3.shape[:2]
    out = np.full(a3.shape, ident, np.uint8)
    for j in range(kh):
        for i in range(kw):
            if mask[j, i]:
                win = p[j:j + h, i:i + w]
                out = np.maximum(out, win) if op_max else np.minimum(out, win)
    return out if a.ndim == 3 else out[:, :, 0]


@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("ks", [(3, 3), (5, 5), (5, 7), (7, 3)])
def test_morph_matches_per_tap_numpy(orc, shape, ks):
    a = rnd((29, 37, 3), 31 + shape)
    m = orc.structuring_element(shape, ks[0], ks[1])
    er = morph_np(a, m, False)
    di = morph_np(a, m, True)
    np.testing.assert_array_equal(orc.morphology(a, 0, shape, ks), er)
    np.testing.assert_array_equal(orc.morphology(a, 1, shape, ks), di)
    op = morph_np(er, m, True)
    cl = morph_np(di, m, False)
    np.testing.assert_array_equal(orc.morphology(a, 2, shape, ks), op)
    np.testing.assert_array_equal(orc.morphology(a, 3, shape, ks), cl)
    np.testing.assert_array_equal(orc.morphology(a, 4, shape, ks), di - er)
    np.testing.assert_array_equal(orc.morphology(a, 5, shape, ks), a - op)
    np.testing.assert_array_equal(orc.morphology(a, 6, shape, ks), cl - a)
    e2 = morph_np(morph_np(a, m, False), m, False)
    np.testing.assert_array_equal(orc.morphology(a, 0, shape, ks, iterations=2), e2)
    np.testing.assert_array_equal(orc.morphology(a, 1, shape, ks, iterations=0), a)


def test_morph_off_centre_general_mask(orc):
    a = rnd((15, 21), 41)
    rng = np.random.default_rng(3)
    mask = (rng.random((4, 6)) > 0.5).astype(np.uint8)
    for anchor in [(0, 0), (5, 3), (2, 1)]:
        d = orc.morph_pass(a, 0, mask, anchor)
        ax, ay = anchor
        want = np.full(a.shape, 255, np.uint8)
        for j in range(4):
            for i in range(6):
                if not mask[j, i]:
                    continue
                for y in range(15):
                    for x in range(21):
                        sy, sx = y + j - ay, x + i - ax
                        if 0 <= sy < 15 and 0 <= sx < 21:
                            want[y, x] = min(want[y, x], a[sy, sx])
        np.testing.assert_array_equal(d, want)


# ---------------------------------------------------------- background

def test_background_closed_form(orc):
    rng = np.random.default_rng(21)
    for (h, w) in [(24, 32), (37, 53), (1, 1), (7, 1)]:
        for shift, thr in ((4, 28), (2, 5), (8, 0)):
            bg = orc.Background(w, h, shift)
            ref = None
            for _ in range(6):
                frame = rng.integers(0, 256, (h, w)).astype(np.uint8)
                mask = np.empty_like(frame)
                bg.apply(frame, mask, thr)
                fi = frame.astype(np.int32)
                if ref is None:
                    ref, want = fi << 8, np.zeros_like(frame)
                else:
                    fg = np.abs(fi - (ref >> 8)) > thr
                    want = np.where(fg, 255, 0).astype(np.uint8)
                    ref = ref + (((fi << 8) - ref) >> np.where(fg, shift + 2, shift))
                np.testing.assert_array_equal(mask, want)
                np.testing.assert_array_equal(bg.model, ref.astype(np.uint16))


# --------------------------------------------------- connected components

def test_components_two(orc):
    c = KA["components_two"]
    a = np.zeros((c["size"][1], c["size"][0]), np.uint8)
    for x, y in c["on"]:
        a[y, x] = 255
    n, lab, st, ce = orc.connected_components_with_stats(a, 8)
    assert n == c["count"]
    for x, y, l in c["label_at"]:
        assert lab[y, x] == l
    assert st[1, 4] == c["area1"] and st[1, 2] == c["width1"]


def flood_labels(a, conn):
    """components.zig:563-590 referenceLabels: flood fill visited in raster
    order (4) or 2x2 block-raster order (8)."""
    h, w = a.shape
    out = np.zeros((h, w), np.int32)
    nxt = 1
    if conn == 8:
        order = [(by * 2 + dy, bx * 2 + dx) for by in range((h + 1) // 2) for bx in range((w + 1) // 2)
                 for dy in range(2) for dx in range(2)]
    else:
        order = [(y, x) for y in range(h) for x in range(w)]
    for y, x in order:
        if y >= h or x >= w or a[y, x] == 0 or out[y, x]:
            continue
        out[y, x] = nxt
        st = [(y, x)]
        while st:
            py, px = st.pop()
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if (dy == 0 and dx == 0) or (conn == 4 and dx and dy):
                        continue
                    ny, nx = py + dy, px + dx
                    if 0 <= ny < h and 0 <= nx < w and a[ny, nx] and not out[ny, nx]:
                        out[ny, nx] = nxt
                        st.append((ny, nx))
        nxt += 1
    return nxt, out


@pytest.mark.parametrize("pattern", ["random", "sparse", "checker", "stripes", "diag", "spiral", "ones", "zeros"])
def test_components_match_flood_fill(orc, pattern):
    rng = np.random.default_rng(42)
    for (w, h) in [(1, 1), (1, 30), (30, 1), (37, 29), (64, 64)]:
        yy, xx = np.mgrid[0:h, 0:w]
        d = np.minimum(np.minimum(xx, yy), np.minimum(w - 1 - xx, h - 1 - yy))
        m = {"random": rng.random((h, w)) < 0.5, "sparse": rng.random((h, w)) < 0.1, "checker": (xx + yy) % 2 == 0,
             "stripes": (xx % 3 == 0) | (yy % 50 == 0), "diag": ((xx + yy) % 7 == 0) | ((xx + 2 * h - yy) % 11 == 0),
             "spiral": (d % 2 == 0) & ~((xx == d + 1) & (yy == d) & (d + 2 < w)), "ones": np.ones((h, w), bool),
             "zeros": np.zeros((h, w), bool)}[pattern]
        a = (m * 255).astype(np.uint8)
        for conn in (4, 8):
            n, lab, st, ce = orc.connected_components_with_stats(a, conn)
            rn, rlab = flood_labels(a, conn)
            assert n == rn
            np.testing.assert_array_equal(lab, rlab)
            for l in range(n):
                ys, xs = np.nonzero(lab == l)
                if len(xs) == 0:
                    assert st[l].tolist() == [-1, 2147483647, 0, 0, 0]
                    continue
                assert st[l].tolist() == [xs.min(), ys.min(), xs.max() - xs.min() + 1, ys.max() - ys.min() + 1, len(xs)]
                assert ce[l, 0] == pytest.approx(xs.mean(), rel=1e-12)
                assert ce[l, 1] == pytest.approx(ys.mean(), rel=1e-12)


def test_otsu_maximises_between_class_variance(orc):
    t = KA["threshold"]
    a = np.array([t["in"]], np.uint8)
    th = orc.threshold_auto(a, 8)
    assert 30 <= th < 200  # threshold.zig:297-298
    img = (np.clip(np.random.default_rng(5).normal(90, 30, (64, 64)), 0, 255)).astype(np.uint8)
    img[:, 32:] = np.clip(img[:, 32:].astype(int) + 90, 0, 255)
    hist = np.bincount(img.ravel(), minlength=256).astype(np.float64)
    p = hist / hist.sum()
    best, arg = -1.0, 0
    for k in range(256):
        q1 = p[:k + 1].sum()
        q2 = 1 - q1
        if min(q1, q2) < 1e-12:
            continue
        m1 = (np.arange(k + 1) * p[:k + 1]).sum() / q1
        m2 = (np.arange(k + 1, 256) * p[k + 1:]).sum() / q2
        s = q1 * q2 * (m1 - m2) ** 2
        if s > best * (1 + 1e-12):
            best, arg = s, k
    assert orc.threshold_auto(img, 8) == arg


# ------------------------------------------------------------ undistort

UD_DISTS = {  # tests/py/test_calib3d_vs_opencv.py:31-36
    "none": np.zeros(5),
    "k1k2": np.array([-0.28, 0.09, 0, 0]),
    "full5": np.array([-0.28, 0.09, 0.001, -0.0005, -0.01]),
    "rational8": np.array([0.1, -0.2, 0.002, 0.001, 0.05, 0.01, 0.002, 0.003]),
}


def ud_camera(w, h):
    """The reference test's camera for a w x h image (test_calib3d_vs_opencv.py:232-235)."""
    K = np.array([[900.0, 0, 640.5], [0, 880.0, 359.2], [0, 0, 1]])
    K[0, 2], K[1, 2] = w / 2 - 0.5, h / 2 + 0.3
    K[0, 0] = K[1, 1] = 0.7 * w
    return K


def rodrigues(r):
    """Rotation vector -> matrix (the textbook formula)."""
    r = np.asarray(r, np.float64)
    th = np.linalg.norm(r)
    k = r / th
    kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.cos(th) * np.eye(3) + (1 - np.cos(th)) * np.outer(k, k) + np.sin(th) * kx


def textbook_map(K, dist, R, new_k, w, h):
    """OpenCV's documented model: (X, Y, W) = R^T newK^-1 (u, v, 1); x' = X/W;
    x'' = x' (1 + k1 r2 + k2 r4 + k3 r6) / (1 + k4 r2 + k5 r4 + k6 r6) + 2 p1 x'y' + p2 (r2 + 2x'^2),
    y'' likewise with p1 and p2 exchanged; map = f x'' + c."""
    k = np.zeros(8)
    k[:len(dist)] = dist
    k1, k2, p1, p2, k3, k4, k5, k6 = k
    nk = K if new_k is None else new_k
    Rm = np.eye(3) if R is None else R
    v, u = np.mgrid[0:h, 0:w].astype(np.float64)
    P = np.stack([u.ravel(), v.ravel(), np.ones(u.size)])
    X, Y, Wc = Rm.T @ np.linalg.inv(nk) @ P
    x, y = X / Wc, Y / Wc
    r2 = x * x + y * y
    rad = (1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3) / (1 + k4 * r2 + k5 * r2 ** 2 + k6 * r2 ** 3)
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return (K[0, 0] * xd + K[0, 2]).reshape(h, w), (K[1, 1] * yd + K[1, 2]).reshape(h, w)


def fixed_point_remap(src, mx, my):
    """OpenCV's fixed-point INTER_LINEAR remap with a zero border: coordinates
    rounded to 1/32 px, weights (32 - a)(32 - b), a(32 - b), (32 - a)b, ab,
    sum + 512 >> 10."""
    h, w = src.shape[:2]
    s3 = src.reshape(h, w, -1).astype(np.int64)
    iu = np.rint(mx * 32).astype(np.int64)
    iv = np.rint(my * 32).astype(np.int64)
    sx, sy, a, b = iu >> 5, iv >> 5, iu & 31, iv & 31
    acc = np.zeros(mx.shape + (s3.shape[2],), np.int64)
    for dx, dy, wt in [(0, 0, (32 - a) * (32 - b)), (1, 0, a * (32 - b)), (0, 1, (32 - a) * b), (1, 1, a * b)]:
        px, py = sx + dx, sy + dy
        ok = (px >= 0) & (px < w) & (py >= 0) & (py < h)
        val = s3[np.clip(py, 0, h - 1), np.clip(px, 0, w - 1)] * ok[..., None]
        acc += val * wt[..., None]
    return ((acc + 512) >> 10).astype(np.uint8).reshape(src.shape)


def test_undistort_zero_distortion_is_identity(orc):
    # undistort.zig:396-407, then every pixel (the identity map is exact).
    src = (np.arange(23 * 37 * 3) * 7 % 256).astype(np.uint8).reshape(23, 37, 3)
    K = np.array([[40, 0, 18], [0, 40, 11], [0, 0, 1.0]])
    dst = orc.undistort(src, K)
    assert np.array_equal(dst[:22, :36], src[:22, :36])
    assert np.array_equal(dst, src)
    for cn in (1, 2, 4):
        a = np.random.default_rng(cn).integers(0, 256, (9, 700, cn), dtype=np.uint8)[..., 0 if cn == 1 else slice(None)]
        assert np.array_equal(orc.undistort(a, ud_camera(700, 9), np.zeros(4)), a)


@pytest.mark.parametrize("tx,ty", [(3, 0), (-2, 5), (0.5, 0), (0, 0.5), (0.25, 0), (1.75, -0.25)])
def test_undistort_shifts_closed_form(orc, tx, ty):
    """With no distortion and new_k = K moved by (-tx, -ty) the map is
    (x + tx, y + ty): integer shifts copy with a zero border; a half pixel
    gives (p0 + p1 + 1) >> 1 and a quarter (3 p0 + p1 + 2) >> 2."""
    h, w = 29, 41
    a = np.random.default_rng(int(10 * tx + ty + 50)).integers(0, 256, (h, w), dtype=np.uint8)
    K = ud_camera(w, h)
    nk = K.copy()
    nk[0, 2] -= tx
    nk[1, 2] -= ty
    got = orc.undistort(a, K, None, nk).astype(np.int64)
    pad = np.zeros((h + 12, w + 12), np.int64)
    pad[6:6 + h, 6:6 + w] = a

    def at(dx, dy):  # src(x + dx, y + dy), 0 outside
        return pad[6 + dy:6 + dy + h, 6 + dx:6 + dx + w]

    def frac_blend(p0, p1, f):  # 1/32 weights along one axis
        return ((32 - f) * 32 * p0 + f * 32 * p1 + 512) >> 10

    ix, fx = int(np.floor(tx)), int(round((tx - np.floor(tx)) * 32))
    iy, fy = int(np.floor(ty)), int(round((ty - np.floor(ty)) * 32))
    if fy == 0:
        want = frac_blend(at(ix, iy), at(ix + 1, iy), fx)
    elif fx == 0:
        want = frac_blend(at(ix, iy), at(ix, iy + 1), fy)
    else:  # both axes: the product of the 1-D weights
        want = ((32 - fx) * (32 - fy) * at(ix, iy) + fx * (32 - fy) * at(ix + 1, iy) + (32 - fx) * fy * at(ix, iy + 1)
                + fx * fy * at(ix + 1, iy + 1) + 512) >> 10
    if tx == 0.5 and ty == 0:
        assert np.array_equal(want, (at(0, 0) + at(1, 0) + 1) >> 1)
    if tx == 0.25:
        assert np.array_equal(want, (3 * at(0, 0) + at(1, 0) + 2) >> 2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", list(UD_DISTS))
def test_undistort_map_matches_documented_model(orc, name):
    for (h, w) in ((720, 1280), (37, 53)):
        K = ud_camera(w, h)
        nk = K.copy()
        nk[0, 0] *= 0.8
        R = rodrigues([0.01, -0.02, 0.005])
        for Rm, nkm in [(None, None), (R, nk), (None, nk)]:
            mx, my = orc.init_undistort_rectify_map(K, UD_DISTS[name], Rm, nkm, (w, h))
            tx, ty = textbook_map(K, UD_DISTS[name], Rm, nkm, w, h)
            assert np.abs(mx - tx).max() <= 1e-4 * w / 1280 + 2e-4 and np.abs(my - ty).max() <= 2e-4, name
            # f32 maps: within half an f32 ulp plus the f64 difference.
            assert np.abs(mx.astype(np.float64) - tx).max() < 1e-3


@pytest.mark.parametrize("name", list(UD_DISTS))
def test_undistort_equals_fixed_point_remap_of_model(orc, name):
    """undistort (striped inverses) against the fixed-point remap of the
    documented model's map computed in one piece, except where a coordinate
    lies within 1e-6 / 32 px of a rounding boundary."""
    for (h, w), cn in (((720, 1280), 3), ((37, 53), 1), ((37, 53), 4), ((150, 91), 2)):
        K = ud_camera(w, h)
        a = np.random.default_rng(w + cn).integers(0, 256, (h, w, cn), dtype=np.uint8)
        a = a[..., 0] if cn == 1 else a
        for nk in (None, np.array([[0.8 * K[0, 0], 0, K[0, 2]], [0, K[1, 1], K[1, 2]], [0, 0, 1]])):
            got = orc.undistort(a, K, UD_DISTS[name], nk)
            mx, my = textbook_map(K, UD_DISTS[name], None, nk, w, h)
            want = fixed_point_remap(a, mx, my)
            near = lambda m: np.abs(m * 32 - np.floor(m * 32) - 0.5) < 1e-6
            ok = ~(near(mx) | near(my))
            assert ok.mean() > 0.999
            assert np.array_equal(got[ok], want[ok]), (name, w, h, cn)


def test_undistort_rejects(orc):
    a = np.zeros((10, 12, 3), np.uint8)
    K = ud_camera(12, 10)
    with pytest.raises(orc.OracleError) as e:
        orc.undistort(a, K, np.zeros(3))
    assert e.value.rc == -1
    with pytest.raises(orc.OracleError) as e:
        orc.undistort(a, K, None, None, np.zeros((10, 13, 3), np.uint8))
    assert e.value.rc == -2
    with pytest.raises(orc.OracleError) as e:
        orc.undistort(np.zeros((10, 12, 5), np.uint8), K)
    assert e.value.rc == -3
    with pytest.raises(orc.OracleError) as e:
        orc.init_undistort_rectify_map(K, None, None, np.zeros((3, 3)), (12, 10))
    assert e.value.rc == -1
    # A singular stripe inverse leaves its rows as they were (fx = 0: all).
    K0 = K.copy()
    K0[0, 0] = 0
    out = np.full((10, 12, 3), 7, np.uint8)
    orc.undistort(np.ones((10, 12, 3), np.uint8), K, None, K0, out)
    assert (out == 7).all()


# ------------------------------------------------ pyramids, box, adaptive

def bmap(p, n, b):
    """border.zig:22-39 for one coordinate: 1 replicate, 2 reflect, 3 wrap, 4 reflect101."""
    if 0 <= p < n:
        return p
    if b == 1:
        return 0 if p < 0 else n - 1
    if b == 3:
        return p % n
    if n == 1:
        return 0
    d = 1 if b == 4 else 0
    while p < 0 or p >= n:
        p = -p - 1 + d if p < 0 else n - 1 - (p - n) - d
    return p


def fill_random(dtype, h, w, cn, seed):
    """pyramid.zig:597-610 fillRandom: an LCG over the samples in row order."""
    n = h * w * cn
    out = np.empty(n, np.uint32)
    s = seed & 0xFFFFFFFF
    for i in range(n):
        s = (s * 1103515245 + 12345) & 0xFFFFFFFF
        out[i] = (s >> 12) & 0xFFFF
    if dtype == np.uint8:
        a = (out & 0xFF).astype(np.uint8)
    elif dtype == np.uint16:
        a = out.astype(np.uint16)
    elif dtype == np.int16:
        a = out.astype(np.uint16).view(np.int16)
    else:
        a = (out.astype(np.float32) / np.float32(65536.0)).astype(np.float32)
    a = a.reshape(h, w, cn)
    return a[..., 0] if cn == 1 else a


PK = np.array([1, 4, 6, 4, 1], np.float64)


def ref_down(a, b):
    """pyramid.zig:570-580 refDown: the direct 5x5 sum of border-mapped samples."""
    h, w = a.shape[:2]
    a3 = a.reshape(h, w, -1).astype(np.float64)
    dh, dw = (h + 1) // 2, (w + 1) // 2
    rows = np.array([[bmap(2 * y + j - 2, h, b) for j in range(5)] for y in range(dh)])
    cols = np.array([[bmap(2 * x + i - 2, w, b) for i in range(5)] for x in range(dw)])
    s = np.zeros((dh, dw, a3.shape[2]))
    for j in range(5):
        for i in range(5):
            s += PK[j] * PK[i] * a3[rows[:, j]][:, cols[:, i]]
    return s


def ref_up(a):
    """pyramid.zig:584-595 refUp: the kernel over the zero-upsampled image,
    reflect101 in upsampled coordinates, odd coordinates contributing 0."""
    h, w = a.shape[:2]
    a3 = a.reshape(h, w, -1).astype(np.float64)
    s = np.zeros((2 * h, 2 * w, a3.shape[2]))
    for j in range(5):
        uy = np.array([bmap(y + j - 2, 2 * h, 4) for y in range(2 * h)])
        for i in range(5):
            ux = np.array([bmap(x + i - 2, 2 * w, 4) for x in range(2 * w)])
            ok = (uy[:, None] % 2 == 0) & (ux[None, :] % 2 == 0)
            s += PK[j] * PK[i] * a3[uy // 2][:, ux // 2] * ok[..., None]
    return s


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int16, np.float32])
def test_pyr_down_up_direct_sum(orc, dtype):
    # pyramid.zig:612-644, every size, channel count and border it uses.
    for (w, h) in ((1, 1), (2, 3), (5, 4), (37, 19), (70, 9), (129, 7)):
        for cn in (1, 2, 3, 4, 5):
            a = fill_random(dtype, h, w, cn, w * 31 + cn)
            for b in (4, 2, 1, 3):
                got = orc.pyr_down(a, b).reshape((h + 1) // 2, (w + 1) // 2, cn).astype(np.float64)
                s = ref_down(a, b)
                if dtype == np.float32:
                    assert np.abs(got - s / 256).max() <= 1e-5, (w, h, cn, b)
                else:
                    assert np.array_equal(got, np.floor((s + 128) / 256)), (w, h, cn, b, dtype)
            got = orc.pyr_up(a).reshape(2 * h, 2 * w, cn).astype(np.float64)
            s = ref_up(a)
            if dtype == np.float32:
                assert np.abs(got - s / 64).max() <= 1e-5, (w, h, cn)
            else:
                assert np.array_equal(got, np.floor((s + 32) / 64)), (w, h, cn, dtype)


def test_pyr_down_f32_order(orc):
    # pyramid.zig:646-672: the f32 sums in the scalar formula's order.
    f = np.float32
    for (w, h) in ((131, 12), (64, 9), (70, 7), (9, 5)):
        for cn in (3, 4):
            a = fill_random(np.float32, h, w, cn, w * 7 + cn)
            for b in (4, 1):
                got = orc.pyr_down(a, b)
                dh, dw = (h + 1) // 2, (w + 1) // 2
                want = np.empty_like(got)
                for y in range(dh):
                    rs = [bmap(2 * y + k - 2, h, b) for k in range(5)]
                    for x in range(dw):
                        cs = [bmap(2 * x + i - 2, w, b) for i in range(5)]
                        r = a[rs][:, cs]  # (5 rows, 5 cols, cn)
                        p = (r[0] + r[4]) + (r[1] + r[3]) * f(4) + r[2] * f(6)
                        want[y, x] = ((p[0] + p[4]) + (p[1] + p[3]) * f(4) + p[2] * f(6)) * f(1.0 / 256.0)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (w, h, cn, b)


def test_pyr_sizes_and_rejects(orc):
    # pyramid.zig:674-686 buildPyramid sizes on a constant 101x50 image.
    lv = np.full((50, 101), 7, np.uint8)
    sizes = []
    for _ in range(4):
        lv = orc.pyr_down(lv)
        sizes.append(lv.shape[::-1])
    assert sizes[0] == (51, 25) and sizes[3] == (7, 4) and lv[2, 3] == 7 and (lv == 7).all()
    with pytest.raises(orc.OracleError) as e:
        orc.pyr_down(np.zeros((50, 101), np.uint8), 4, np.zeros((12, 26), np.uint8))
    assert e.value.rc == -2
    with pytest.raises(orc.OracleError) as e:
        orc.pyr_down(np.zeros((8, 8), np.uint8), 0)
    assert e.value.rc == -3
    with pytest.raises(orc.OracleError) as e:
        orc.pyr_up(np.zeros((8, 8), np.uint8), 1)
    assert e.value.rc == -3


def naive_box(a, kw, kh, b):
    h, w = a.shape[:2]
    a3 = a.reshape(h, w, -1).astype(np.int64)
    s = np.zeros(a3.shape, np.int64)
    for j in range(kh):
        rows = [bmap(y + j - kh // 2, h, b) if b else y + j - kh // 2 for y in range(h)]
        for i in range(kw):
            cols = [bmap(x + i - kw // 2, w, b) if b else x + i - kw // 2 for x in range(w)]
            ry = np.array(rows)
            cx = np.array(cols)
            ok = ((ry >= 0) & (ry < h))[:, None] & ((cx >= 0) & (cx < w))[None, :]
            s += a3[np.clip(ry, 0, h - 1)][:, np.clip(cx, 0, w - 1)] * ok[..., None]
    return s


def sat_round(v, lo, hi):
    return np.clip(np.rint(v), lo, hi)


def test_box_filter_naive_sum(orc):
    # filter.zig:960-985 (67 x 23, cn 1-4, its seven windows, reflect101),
    # then every other border, u16, f32 and unnormalised sums.
    for cn in (1, 2, 3, 4):
        seed = 99
        vals = np.empty(67 * 23 * cn, np.uint8)
        for i in range(vals.size):
            seed = (seed * 1103515245 + 12345) & 0xFFFFFFFF
            vals[i] = (seed >> 16) & 0xFF
        a = vals.reshape(23, 67, cn)
        a = a[..., 0] if cn == 1 else a
        for kw, kh in ((3, 3), (5, 5), (7, 3), (9, 1), (15, 15), (31, 31), (1, 9)):
            scale = np.float32(1.0 / (kw * kh))
            for b in (4, 1, 2, 3, 0):
                s = naive_box(a, kw, kh, b)
                want = sat_round(s.astype(np.float32) * scale, 0, 255).reshape(a.shape)
                assert np.array_equal(orc.box_filter(a, (kw, kh), True, b), want), (cn, kw, kh, b)
            s = naive_box(a, kw, kh, 4)
            assert np.array_equal(orc.box_filter(a, (kw, kh), False), np.clip(s, 0, 255).reshape(a.shape))
    u = fill_random(np.uint16, 19, 37, 3, 5)
    for kw, kh in ((3, 3), (255, 1), (5, 255)):
        s = naive_box(u, kw, kh, 1)
        want = sat_round(s.astype(np.float32) * np.float32(1.0 / (kw * kh)), 0, 65535).reshape(u.shape)
        assert np.array_equal(orc.box_filter(u, (kw, kh), True, 1), want), (kw, kh)
    f = fill_random(np.float32, 19, 37, 1, 6)
    for kw, kh in ((3, 3), (9, 5), (255, 3)):
        h, w = f.shape
        ref = np.zeros((h, w))
        for j in range(kh):
            for i in range(kw):
                ry = np.array([bmap(y + j - kh // 2, h, 4) for y in range(h)])
                cx = np.array([bmap(x + i - kw // 2, w, 4) for x in range(w)])
                ref += f.astype(np.float64)[ry][:, cx]
        got = orc.box_filter(f, (kw, kh))
        assert np.abs(got - ref / (kw * kh)).max() < 1e-5, (kw, kh)
    # i16, negative samples included: exact signed window sums.
    r = np.random.default_rng(7)
    i16 = r.integers(-32768, 32767, (19, 37, 2), endpoint=True).astype(np.int16)
    for kw, kh in ((3, 3), (255, 1), (5, 255), (129, 129)):
        for b in (4, 0):
            s = naive_box(i16, kw, kh, b)
            want = sat_round(s.astype(np.float32) * np.float32(1.0 / (kw * kh)), -32768, 32767).reshape(i16.shape)
            assert np.array_equal(orc.box_filter(i16, (kw, kh), True, b), want), (kw, kh, b)
        s = naive_box(i16, kw, kh, 4)
        assert np.array_equal(orc.box_filter(i16, (kw, kh), False), np.clip(s, -32768, 32767).reshape(i16.shape))


def test_box_filter_rejects(orc):
    a = np.zeros((5, 6), np.uint8)
    for ks, border, rc in [((3, 3), 7, -3), ((0, 3), 4, -1), ((4, 3), 4, -1), ((257, 3), 4, -1)]:
        with pytest.raises(orc.OracleError) as e:
            orc.box_filter(a, ks, True, border)
        assert e.value.rc == rc, (ks, border)


@pytest.mark.parametrize("kind", [0, 1])
def test_adaptive_threshold(orc, kind):
    """threshold.zig:306-360: MEAN_C against a naive replicate-border box mean,
    GAUSSIAN_C against the Gaussian blur of an f32 copy; closed forms on a
    constant image."""
    a = fill_random(np.uint8, 41, 57, 1, 21)
    for block in (3, 5, 11, 31):
        for delta in (0, 5, -3.5, 2.25, 300, -300):
            s = naive_box(a, block, block, 1)
            mean = sat_round(s.astype(np.float32) * np.float32(1.0 / (block * block)), 0, 255).reshape(a.shape)
            idelta = int(np.ceil(delta) if kind == 0 else np.floor(delta))
            gt = a.astype(np.int64) + idelta > mean
            want = np.where(gt == (kind == 0), 200, 0).astype(np.uint8)
            assert np.array_equal(orc.adaptive_threshold(a, 200, 0, kind, block, delta), want), (block, delta)
            g = orc.gaussian_blur(a.astype(np.float32), (block, block), 0, 0, 1)
            gm = sat_round(g, 0, 255)
            want = np.where((a.astype(np.int64) + idelta > gm) == (kind == 0), 255, 0).astype(np.uint8)
            assert np.array_equal(orc.adaptive_threshold(a, 255.4, 1, kind, block, delta), want), (block, delta)
    # MEAN_C blocks beyond the Gaussian's 127 (boxFilter takes up to 255).
    for block in (129, 255):
        s = naive_box(a, block, block, 1)
        mean = sat_round(s.astype(np.float32) * np.float32(1.0 / (block * block)), 0, 255).reshape(a.shape)
        want = np.where((a.astype(np.int64) + 2 > mean) == (kind == 0), 255, 0).astype(np.uint8)
        assert np.array_equal(orc.adaptive_threshold(a, 255, 0, kind, block, 2), want), block
    c = np.full((9, 13), 77, np.uint8)
    for delta, want_on in ((0.5, True), (0, False), (-0.5, False), (-1, False), (1, True)):
        on = orc.adaptive_threshold(c, 9, 0, 0, 3, delta)
        assert (on == (9 if want_on else 0)).all(), delta


def test_adaptive_threshold_rejects(orc):
    a = np.zeros((5, 6), np.uint8)
    # GaussianBlur's and boxFilter's InvalidKernel beyond 127 and 255 taps.
    for args, rc in [((255, 2, 0, 3, 0), -1), ((255, 0, 2, 3, 0), -1), ((255, 0, 0, 1, 0), -1), ((255, 0, 0, 4, 0), -1),
                     ((255, 1, 0, 129, 0), -1), ((255, 0, 0, 257, 0), -1)]:
        with pytest.raises(orc.OracleError) as e:
            orc.adaptive_threshold(a, *args)
        assert e.value.rc == rc, args
    with pytest.raises(orc.OracleError) as e:
        orc.adaptive_threshold(np.zeros((5, 6, 3), np.uint8), 255, 0, 0, 3, 0)
    assert e.value.rc == -3
