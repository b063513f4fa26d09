"""HIP path vs the CPU oracle, bit for bit (GPU).

Every case runs the product through the C ABI twice: with host NumPy arrays
(staged by the library) and with device tensors (used in place), and must
equal the oracle exactly, integer and float paths alike (the float paths
reproduce the reference's operation order, so no tolerance is needed).
Case lists follow the reference's difftests (tests/py/test_core_vs_opencv.py,
test_calib3d_vs_opencv.py, test_motion_vs_opencv.py): every border, 1/3/4
channels, strided views and degenerate sizes.
"""
import contextlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def va():
    import visar_amd
    assert torch.cuda.is_available()
    # The library's calls go on torch's stream, ordered with the tensors' own
    # fills and copies (the library's default stream is a separate one).
    visar_amd.set_stream(torch.cuda.current_stream())
    yield visar_amd
    visar_amd.set_stream(None)


def rnd(shape, seed, dtype=np.uint8):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        return rng.uniform(-50, 300, shape).astype(np.float32)
    info = np.iinfo(dtype)
    return rng.integers(info.min, info.max + 1, shape, dtype=np.int64).astype(dtype)


def natural(h, w, cn, seed, dtype=np.uint8):
    """Smooth content with noise (like the reference's natural() inputs)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = 128 + 60 * np.sin(xx / 7.0) * np.cos(yy / 11.0)
    img = base[:, :, None] + rng.normal(0, 20, (h, w, cn))
    if dtype == np.float32:
        out = (img / 255.0).astype(np.float32)
    elif dtype == np.uint16:
        out = np.clip(img * 257, 0, 65535).astype(np.uint16)
    elif dtype == np.int16:
        out = np.clip((img - 128) * 200, -32768, 32767).astype(np.int16)
    else:
        out = np.clip(img, 0, 255).astype(np.uint8)
    return out[:, :, 0] if cn == 1 else out


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def both(fn_va, fn_or, *arrays, n_images=1):
    """Runs the product on host arrays and on device copies of the first
    `n_images` arguments; returns (host, device, oracle)."""
    host = fn_va(*arrays)
    dres = fn_va(*[dev(a) if i < n_images else a for i, a in enumerate(arrays)])
    dres = dres.cpu().numpy() if hasattr(dres, "cpu") else dres
    return host, dres, fn_or(*arrays)


def assert_same(got, want, msg=""):
    assert got.shape == want.shape and got.dtype == want.dtype, msg
    if got.dtype.kind == "f":
        ok = np.array_equal(got.view(np.uint32), want.view(np.uint32))
        if not ok:
            d = np.abs(got.astype(np.float64) - want)
            raise AssertionError(f"{msg}: f32 bits differ, max |d| {d.max()} rel {np.nanmax(d / (np.abs(want) + 1e-30))}")
    else:
        if not np.array_equal(got, want):
            idx = np.argwhere(got != want)
            raise AssertionError(f"{msg}: {len(idx)} differ, first {idx[:3].tolist()}")


def check(va, orc, name, *arrays, **kw):
    f_va = getattr(va, name)
    f_or = getattr(orc, name)
    n_images = 3 if name == "remap" else 1
    host, dres, want = both(lambda *a: f_va(*a, **kw), lambda *a: f_or(*a, **kw), *arrays, n_images=n_images)
    assert_same(host, want, f"{name} host {kw}")
    assert_same(dres, want, f"{name} device {kw}")


# ------------------------------------------------------------ Gaussian

GAUSS_SIZES = [(37, 53), (101, 67), (1, 1), (7, 1), (1, 9), (2, 40), (31, 3)]


@pytest.mark.parametrize("cn", [1, 3, 4])
def test_gaussian_u8_sweep(va, orc, cn):
    for (h, w) in GAUSS_SIZES:
        a = natural(h, w, cn, h * 31 + w)
        for (kw, kh), sx, sy in [((3, 15), 0, 0), ((21, 5), 0, 0), ((1, 9), 0, 0), ((13, 1), 0, 0), ((15, 15), 0, 0),
                                 ((31, 31), 0, 0), ((11, 7), 1.4, 3.2), ((9, 9), 2.0, 0.7), ((0, 0), 1.2, 2.9),
                                 ((5, 5), 0.4, 0), ((3, 3), 0, 0), ((5, 5), 0, 0), ((7, 7), 0, 0)]:
            for border in range(5):
                check(va, orc, "gaussian_blur", a, (kw, kh), sx, sigma_y=sy, border=border)


@pytest.mark.parametrize("dt", [np.uint16, np.int16, np.float32])
def test_gaussian_other_depths(va, orc, dt):
    for cn in (1, 3):
        a = natural(47, 61, cn, 5, dt)
        for k, s in [(3, 0), (5, 0), (7, 0), (9, 2.0), (0, 1.5), (13, 0), (31, 0)]:
            for border in (0, 1, 2, 4):
                check(va, orc, "gaussian_blur", a, (k, k), s, border=border)


def test_gaussian_1080p_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(0, (1080, 1920))
    check(va, orc, "gaussian_blur", a, (5, 5), 0, border=4)


def test_gaussian_strided_views(va, orc):
    for cn in (1, 3):
        big = natural(120, 160, cn, 77)
        for y0, x0, h, w in [(3, 5, 37, 53), (1, 1, 97, 131), (10, 7, 9, 150)]:
            v = big[y0:y0 + h, x0:x0 + w]
            for k, s in [(3, 0), (7, 0), (13, 0), (19, 2.3)]:
                want = orc.gaussian_blur(np.ascontiguousarray(v), (k, k), s)
                got = va.gaussian_blur(v, (k, k), s)
                assert_same(got, want, f"view {y0},{x0} k{k}")
                dbig = dev(big)
                dv = dbig[y0:y0 + h, x0:x0 + w]
                assert_same(va.gaussian_blur(dv, (k, k), s).cpu().numpy(), want, f"device view k{k}")


def test_stencil_and_morph_misaligned_rows(va, orc):
    """Sources and destinations at every byte offset mod 4 with odd row
    strides, wide enough for several 256-byte tiles: every output byte equals
    the oracle and the bytes around the destination view stay untouched."""
    for cn in (1, 3):
        for (h, w) in [(19, 347), (6, 601), (33, 97)]:
            a = natural(h, w, cn, h + w + cn)
            for off in range(4):
                for pad in (0, 1, 5):
                    pitch = w * cn + pad
                    sbig = torch.zeros(h * pitch + 8, dtype=torch.uint8, device="cuda")
                    sv = sbig[off:off + h * pitch].view(h, pitch)[:, :w * cn]
                    sv.copy_(dev(a.reshape(h, w * cn)))
                    sv = sv.view(h, w, cn) if cn == 3 else sv
                    dbig = torch.full((h * pitch + 8,), 77, dtype=torch.uint8, device="cuda")
                    dv = dbig[(off + 1) % 4:(off + 1) % 4 + h * pitch].view(h, pitch)[:, :w * cn]
                    dv = dv.view(h, w, cn) if cn == 3 else dv
                    keep = torch.ones_like(dbig, dtype=torch.bool)
                    o = (off + 1) % 4
                    for r in range(h):
                        keep[o + r * pitch:o + r * pitch + w * cn] = False
                    for k in (3, 5, 7):
                        want = orc.gaussian_blur(a, (k, k), 0, border=4)
                        dbig.fill_(77)
                        va.gaussian_blur(sv, (k, k), 0, border=4, dst=dv)
                        assert_same(dv.cpu().numpy(), want, f"gauss k{k} cn{cn} {h}x{w} off{off} pad{pad}")
                        assert bool((dbig[keep] == 77).all()), f"gauss k{k} wrote outside its view"
                    for op, shape, k in [(1, 0, 5), (0, 2, 7), (1, 1, 3)]:
                        want = orc.morphology(a, op, shape, (k, k))
                        dbig.fill_(77)
                        va.morphology(sv, op, shape, (k, k), dst=dv)
                        assert_same(dv.cpu().numpy(), want, f"morph op{op} k{k} cn{cn} off{off} pad{pad}")
                        assert bool((dbig[keep] == 77).all()), f"morph op{op} wrote outside its view"


# --------------------------------------------------------------- Sobel

@pytest.mark.parametrize("dx,dy,ks", [(1, 0, 3), (0, 1, 3), (1, 1, 3), (1, 0, 5), (2, 0, 5), (1, 0, -1), (0, 1, 1),
                                      (1, 0, 7), (2, 1, 7)])
def test_sobel(va, orc, dx, dy, ks):
    for cn in (1, 3):
        a = natural(67, 83, cn, 9)
        for border in range(5):
            check(va, orc, "sobel", a, np.int16, dx, dy, ksize=ks, border=border)
            check(va, orc, "sobel", a, np.float32, dx, dy, ksize=ks, border=border)
        check(va, orc, "sobel", a, np.int16, dx, dy, ksize=ks, scale=0.37)
        check(va, orc, "sobel", a, np.float32, dx, dy, ksize=ks, scale=0.37)
    f = natural(35, 41, 1, 10, np.float32)
    check(va, orc, "sobel", f, np.float32, dx, dy, ksize=ks)
    u = natural(35, 41, 1, 11, np.uint16)
    check(va, orc, "sobel", u, np.float32, dx, dy, ksize=ks)


def _stencil_cases():
    """The packed stencils the streamed one-channel kernels take: GaussianBlur
    3/5/7 with sigma 0 and Sobel 3x3 dx / dy to i16."""
    return [("gaussian_blur", ((k, k), 0), {}) for k in (3, 5, 7)] + \
           [("sobel", (np.int16, dx, dy), {"ksize": 3}) for dx, dy in ((1, 0), (0, 1))]


def test_stencil_streamed_layouts(va, orc):
    """The streamed kernels (one-channel images with dword rows) against the
    oracle at widths on both sides of their 256-byte strips, heights below
    and above their row bands, every border (WRAP takes the LDS tile kernels)
    and every band height, on contiguous tensors and on dword-aligned views
    of wider buffers."""
    from visar_amd import autotune
    big = natural(80, 1100, 1, 91)
    try:
        for band in (0, 1, 2, 4, 8, 16):
            autotune.set_knob("stencil_band", band)
            sizes = [(1, 8), (2, 12), (3, 252), (5, 256), (9, 260), (17, 516), (33, 1028), (70, 64)] if band == 0 else \
                    [(1, 8), (9, 260), (70, 516)]
            for h, w in sizes:
                a = natural(h, w, 1, h * 131 + w)
                v = big[3:3 + h, 8:8 + w]
                for name, args, kw in _stencil_cases():
                    for border in range(5):
                        want = getattr(orc, name)(a, *args, border=border, **kw)
                        got = getattr(va, name)(dev(a), *args, border=border, **kw).cpu().numpy()
                        assert_same(got, want, f"{name} {args} {h}x{w} border {border} band {band}")
                        if border == 4:
                            dbig = dev(big)
                            gv = getattr(va, name)(dbig[3:3 + h, 8:8 + w], *args, border=border, **kw).cpu().numpy()
                            assert_same(gv, getattr(orc, name)(np.ascontiguousarray(v), *args, border=border, **kw),
                                        f"{name} {args} view {h}x{w}")
    finally:
        autotune.set_knob("stencil_band", 0)


def test_stencil_streamed_equals_tiles(va):
    """The streamed kernels equal the LDS tile kernels (tuning
    "stencil_direct" 1) on 1080p frames for every packed stencil and border."""
    from visar_amd import autotune, synth
    a = dev(synth.uniform_bytes(5, (1080, 1920)))
    try:
        for name, args, kw in _stencil_cases():
            for border in (0, 1, 2, 4):
                outs = []
                for knob in (0, 1):
                    autotune.set_knob("stencil_direct", knob)
                    outs.append(getattr(va, name)(a, *args, border=border, **kw))
                assert torch.equal(outs[0], outs[1]), f"{name} {args} border {border}"
    finally:
        autotune.set_knob("stencil_direct", 0)


def test_stencil_general_streamed(va, orc):
    """The general streamed GaussianBlur (one and three channels, any row
    alignment and width) against the oracle: widths from 1 pixel to both
    sides of its 240-pixel strips, heights below and above its bands, every
    border, both band heights, contiguous tensors (rows off dwords when the
    row length is not a multiple of 4) and views at odd offsets; then equal
    to the LDS tile kernels on larger frames."""
    from visar_amd import autotune
    sizes = [(1, 1), (2, 3), (5, 7), (9, 13), (3, 236), (16, 240), (7, 241), (11, 246), (4, 483), (21, 97)]
    try:
        for band in (0, 2, 4):
            autotune.set_knob("stencil_band", band)
            for cn in (1, 3):
                for h, w in (sizes if band == 0 else sizes[3::2]):
                    a = natural(h, w, cn, 7 * h + w + cn)
                    big = torch.zeros(h * (w * cn + 5) + 8, dtype=torch.uint8, device="cuda")
                    sv = big[3:3 + h * (w * cn + 5)].view(h, w * cn + 5)[:, :w * cn]
                    sv.copy_(dev(a.reshape(h, w * cn)))
                    sv = sv.view(h, w, cn) if cn == 3 else sv
                    for k in (3, 5, 7):
                        for border in range(5):
                            want = orc.gaussian_blur(a, (k, k), 0, border=border)
                            got = va.gaussian_blur(dev(a), (k, k), 0, border=border).cpu().numpy()
                            assert_same(got, want, f"k{k} cn{cn} {h}x{w} border {border} band {band}")
                        gv = va.gaussian_blur(sv, (k, k), 0, border=4).cpu().numpy()
                        assert_same(gv, orc.gaussian_blur(a, (k, k), 0, border=4), f"view k{k} cn{cn} {h}x{w}")
        autotune.set_knob("stencil_band", 0)
        for cn in (1, 3):
            a = dev(natural(270, 1001, cn, 40 + cn))
            for k in (3, 5, 7):
                for border in (0, 1, 2, 3, 4):
                    outs = []
                    for knob in (0, 1):
                        autotune.set_knob("stencil_direct", knob)
                        outs.append(va.gaussian_blur(a, (k, k), 0, border=border))
                    assert torch.equal(outs[0], outs[1]), f"k{k} cn{cn} border {border}"
    finally:
        autotune.set_knob("stencil_band", 0)
        autotune.set_knob("stencil_direct", 0)


def test_sobel_1080p_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(3, (1080, 1920))
    check(va, orc, "sobel", a, np.int16, 1, 0, ksize=3, border=4)


# --------------------------------------------------------- sepFilter2D

def test_sep_filter_2d(va, orc):
    kx = np.array([0.1, -0.7, 0.3, 1.25, 0.05], np.float32)
    ky = np.array([0.2, 0.5, -0.15], np.float32)
    sym = np.array([0.1, 0.2, 0.4, 0.2, 0.1], np.float32)
    odd = np.array([-0.5, -0.25, 0, 0.25, 0.5], np.float32)
    for sdt in (np.uint8, np.uint16, np.int16, np.float32):
        a = natural(41, 57, 3, 12, sdt)
        for ddt in (np.uint8, np.int16, np.float32, np.uint16):
            for kxx, kyy in [(kx, ky), (sym, odd), (odd, sym), (sym, sym)]:
                for border in (0, 1, 3, 4):
                    check(va, orc, "sep_filter_2d", a, ddt, kxx, kyy, border=border)


def test_sep_filter_window_lengths(va, orc):
    """Every window length the engine unrolls (odd 3..15) and the runtime
    loop on either side of them (1, 17), across and down, with plain,
    symmetric and antisymmetric taps; 300 columns so that both the folded
    body and the in-order tail of a row are used."""
    rng = np.random.default_rng(21)
    for n in range(1, 19, 2):
        m = 18 - n
        p = rng.normal(size=n).astype(np.float32)
        q = rng.normal(size=m).astype(np.float32)
        kinds_x = (p, p + p[::-1], p - p[::-1])
        kinds_y = (q, q - q[::-1], q + q[::-1])
        for sdt, ddt in ((np.uint8, np.uint8), (np.uint8, np.float32), (np.float32, np.float32), (np.uint16, np.int16)):
            a = natural(23, 300, 1, n + 40, sdt)
            for kx, ky in zip(kinds_x, kinds_y):
                check(va, orc, "sep_filter_2d", a, ddt, kx, ky, border=4)
        a = natural(23, 300, 3, n + 41)
        check(va, orc, "box_filter", a, (n, m), True, 4)
        check(va, orc, "box_filter", a, (m, n), True, 1)
        check(va, orc, "gaussian_blur", a, (n, m), 0, border=4)


# --------------------------------------------------- color / threshold

def test_cvt_gray(va, orc):
    for code, scn in [(6, 3), (7, 3), (10, 4), (11, 4)]:
        for dt in (np.uint8, np.uint16, np.float32):
            for (h, w) in [(2160, 3840), (37, 53), (1, 1), (5, 17)]:
                a = natural(h, w, scn, code + h, dt)
                check(va, orc, "cvt_color", a, code)


def test_cvt_gray_wave_blocks(va, orc):
    """The block-per-wave u8 kernel (contiguous device images of 1024 pixels
    or more): whole and partial last blocks and workgroups, 3 and 4 channels,
    both channel orders, against the oracle and the per-thread kernel
    (tuning "gray_layout" 1)."""
    from visar_amd import autotune
    try:
        for code, scn in [(6, 3), (7, 3), (10, 4), (11, 4)]:
            for (h, w) in [(32, 32), (100, 100), (1, 1024), (1, 1025), (17, 1000), (513, 16), (64, 128), (65, 128)]:
                a = rnd((h, w, scn), code * 1000 + h + w)
                want = orc.cvt_color(a, code)
                outs = []
                for knob in (0, 1):
                    autotune.set_knob("gray_layout", knob)
                    outs.append(va.cvt_color(dev(a), code).cpu().numpy())
                assert_same(outs[0], want, f"code {code} {h}x{w}")
                assert_same(outs[1], want, f"code {code} {h}x{w} per-thread")
    finally:
        autotune.set_knob("gray_layout", 0)


def test_cvt_gray_4k_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(0, (2160, 3840, 3))
    check(va, orc, "cvt_color", a, 6)


def test_threshold(va, orc):
    for dt in (np.uint8, np.uint16, np.int16, np.float32):
        a = natural(53, 71, 3, 13, dt)
        ths = {np.uint8: [127.5, -3, 300, 0, 254.9], np.uint16: [30000.2, -5, 70000], np.int16: [0.5, -40000, 1000],
               np.float32: [0.5, -1, 2]}[dt]
        for t in ths:
            for kind in range(5):
                mv = {np.uint8: 200, np.uint16: 60000, np.int16: 1234, np.float32: 1.0}[dt]
                host = va.threshold(a, t, mv, kind)[1]
                want = orc.threshold(a, t, mv, kind)[1]
                assert_same(host, want, f"threshold {dt} {t} {kind}")
                assert_same(va.threshold(dev(a), t, mv, kind)[1].cpu().numpy(), want, "threshold device")


def test_threshold_auto(va, orc):
    a = natural(120, 160, 1, 14)
    for auto in (8, 16):
        t_va, d_va = va.threshold(a, 0, 255, auto)
        t_or = orc.threshold_auto(a, auto)
        assert t_va == t_or
        assert_same(d_va, orc.threshold(a, t_or, 255, 0)[1])


# -------------------------------------------------------------- resize

def test_resize_cases(va, orc):
    for (h, w) in [(480, 640), (37, 53), (61, 97)]:
        for cn in (1, 3, 4):
            a = natural(h, w, cn, h + cn)
            for (dw, dh) in [(max(1, w // 2), max(1, h // 2)), (w * 2, h * 2), (max(1, w * 3 // 4), max(1, h * 2 // 3)),
                             (113, 71), (int(round(w * 1.5)), int(round(h * 1.5))), (w, h)]:
                for interp in (0, 1, 3):
                    check(va, orc, "resize", a, (dw, dh), interp)


def test_resize_area_cases(va, orc):
    for (sw, sh), (dw, dh) in [((640, 480), (416, 312)), ((1920, 1080), (640, 360)), ((1920, 1080), (1280, 720)),
                               ((1280, 720), (320, 180)), ((97, 61), (53, 37)), ((640, 480), (60, 45)),
                               ((640, 480), (7, 5)), ((640, 480), (640, 300)), ((640, 480), (320, 480)),
                               ((640, 480), (639, 479)), ((97, 61), (130, 80)), ((97, 61), (50, 80)),
                               ((1001, 777), (500, 388)), ((1001, 777), (501, 389))]:
        for cn in (1, 3, 4):
            a = natural(sh, sw, cn, sw + cn)
            check(va, orc, "resize", a, (dw, dh), 3)


def test_resize_generic_depths_and_channels(va, orc):
    """The generic resize kernels (resize.zig:47-60): LINEAR and AREA for
    u16, i16 and f32 at 1 and 3 channels and for u8 with 5 and 6 channels,
    at downscales, upscales, AREA-enlarging, integer factors (2x2 and
    others) and fractional shrinks; host and device images."""
    for dt, cns in [(np.uint16, (1, 3)), (np.int16, (1, 3)), (np.float32, (1, 3)), (np.uint8, (5, 6))]:
        for cn in cns:
            for (h, w) in [(61, 97), (40, 60), (7, 5)]:
                a = natural(h, w, cn, 3 * cn + h, dt)
                sizes = [(max(1, w // 2), max(1, h // 2)), (w * 2, h * 2), (max(1, w * 3 // 4), max(1, h * 2 // 3)),
                         (53, 37), (max(1, w // 3), max(1, h // 4)), (w + 13, max(1, h - 3)), (1, 1)]
                for size in sizes:
                    for interp in (1, 3):
                        check(va, orc, "resize", a, size, interp)


def test_resize_linear_scales_near_tile_limits(va, orc):
    """Linear (and area-up) scales on both sides of the LDS tile's source-box
    limits, 16-row tile bands with ragged ends, and 256-sample bands with a
    ragged last tile: the staged tiles and the direct-gather tiles alike."""
    for cn in (1, 3, 4):
        a = natural(83, 301, cn, 5 + cn)
        for (dw, dh) in [(302, 84), (451, 124), (2107, 581), (151, 42), (150, 41), (181, 50), (120, 33), (301, 250),
                         (700, 83), (3, 900), (900, 2)]:
            for interp in (1, 3):
                if interp == 3 and dw <= 301 and dh <= 83:
                    continue  # area shrink, not the linear taps
                check(va, orc, "resize", a, (dw, dh), interp)


def test_resize_misaligned_rows(va, orc):
    """Linear and area resizes with sources and destinations at every byte
    offset mod 4 and odd row strides: every output byte equals the oracle and
    the bytes around the destination view stay untouched."""
    for cn in (1, 3):
        h, w = 37, 301
        a = natural(h, w, cn, 77 + cn)
        for (dw, dh, interp) in [(452, 56, 1), (180, 22, 1), (150, 18, 3), (211, 29, 3)]:
            want = orc.resize(a, (dw, dh), interp)
            for off in range(4):
                for pad in (0, 1, 5):
                    pitch = w * cn + pad
                    sbig = torch.zeros(h * pitch + 8, dtype=torch.uint8, device="cuda")
                    sv = sbig[off:off + h * pitch].view(h, pitch)[:, :w * cn]
                    sv.copy_(dev(a.reshape(h, w * cn)))
                    sv = sv.view(h, w, cn) if cn == 3 else sv
                    dp = dw * cn + pad
                    o = (off + 1) % 4
                    dbig = torch.full((dh * dp + 8,), 77, dtype=torch.uint8, device="cuda")
                    dv = dbig[o:o + dh * dp].view(dh, dp)[:, :dw * cn]
                    dv = dv.view(dh, dw, cn) if cn == 3 else dv
                    keep = torch.ones_like(dbig, dtype=torch.bool)
                    for r in range(dh):
                        keep[o + r * dp:o + r * dp + dw * cn] = False
                    va.resize(sv, (dw, dh), interp, dst=dv)
                    assert_same(dv.cpu().numpy(), want, f"interp {interp} {w}x{h}->{dw}x{dh} cn{cn} off{off} pad{pad}")
                    assert bool((dbig[keep] == 77).all()), "resize wrote outside its view"


def test_resize_linear_streamed_layouts(va, orc):
    """The register-streamed LINEAR u8 kernel at several band heights and
    bands per wave (and the automatic choice) and the LDS tiles it replaces,
    against the oracle:
    1-4 channels, upscales (1.5x, 2x, 3.7x) and downscales down to about
    1/2, widths on both sides of its 252-sample strips, ragged row bands,
    and misaligned views of wider buffers."""
    from visar_amd import autotune
    cases = [((61, 97), (146, 92)), ((37, 253), (380, 56)), ((40, 64), (128, 80)), ((33, 70), (259, 122)),
             ((90, 300), (180, 55)), ((81, 501), (255, 41)), ((20, 1000), (1500, 30)), ((7, 5), (1, 1))]
    try:
        for stream, rows, run in [(0, 0, 0), (0, 4, 1), (0, 8, 1), (0, 4, 3), (0, 8, 16), (1, 0, 0)]:
            autotune.set_knob("lin_stream", stream)
            autotune.set_knob("lin_rows", rows)
            autotune.set_knob("lin_run", run)
            for cn in (1, 2, 3, 4):
                for (h, w), (dw, dh) in cases:
                    a = natural(h, w, cn, 17 * h + w + cn)
                    want = orc.resize(a, (dw, dh), 1)
                    assert_same(va.resize(dev(a), (dw, dh), 1).cpu().numpy(), want,
                                f"{w}x{h}->{dw}x{dh} cn{cn} stream {stream} rows {rows} run {run}")
                    pitch = w * cn + 3
                    big = torch.zeros((h + 1) * pitch, dtype=torch.uint8, device="cuda")
                    sv = big[1:1 + h * pitch].view(h, pitch)[:, :w * cn]
                    sv.copy_(dev(a.reshape(h, w * cn)))
                    sv = sv.view(h, w, cn) if cn > 1 else sv
                    assert_same(va.resize(sv, (dw, dh), 1).cpu().numpy(), want,
                                f"view {w}x{h}->{dw}x{dh} cn{cn} stream {stream} rows {rows} run {run}")
    finally:
        autotune.set_knob("lin_stream", 0)
        autotune.set_knob("lin_rows", 0)
        autotune.set_knob("lin_run", 0)


@pytest.mark.parametrize("own_stream", [False, True])
def test_queued_calls_keep_their_tables_and_temporaries(va, orc, own_stream):
    """Queued calls (no per-call wait) whose tap tables and temporaries come
    from the shared pool: each buffer is reused only after the work that
    reads it, so every output equals the oracle after one final wait. On
    HIP's default stream (a null handle) and on a stream of its own."""
    cases = []
    for i in range(24):
        cn = 1 + 2 * (i % 2)
        h, w = 40 + 7 * i, 60 + 13 * i
        a = natural(h, w, cn, 900 + i)
        if i % 3 == 0:
            cases.append((a, ("resize", (w * 3 // 2, h * 3 // 2), 1)))
        elif i % 3 == 1:
            cases.append((a, ("resize", (w // 2 + 1, h // 2 + 1), 3)))
        else:
            cases.append((a, ("morphology", 2, 0, (3, 3), 2)))
    srcs = [dev(a) for a, _ in cases]
    outs = []
    side = torch.cuda.Stream() if own_stream else None
    if side is not None:
        side.wait_stream(torch.cuda.current_stream())
        va.set_stream(side)
    va.set_sync(False)
    try:
        with torch.cuda.stream(side) if side is not None else contextlib.nullcontext():
            for sv, (_, op) in zip(srcs, cases):
                if op[0] == "resize":
                    outs.append(va.resize(sv, op[1], op[2]))
                else:
                    outs.append(va.morphology(sv, op[1], op[2], op[3], iterations=op[4]))
        va.synchronize()
    finally:
        va.set_sync(True)
        if side is not None:
            va.set_stream(torch.cuda.current_stream())
    for i, ((a, op), got) in enumerate(zip(cases, outs)):
        want = orc.resize(a, op[1], op[2]) if op[0] == "resize" else orc.morphology(a, op[1], op[2], op[3], iterations=op[4])
        assert_same(got.cpu().numpy(), want, f"queued call {i} {op}")


def test_resize_linear_4k_half_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(5, (2160, 3840))
    check(va, orc, "resize", a, (1920, 1080), 1)


# --------------------------------------------------------------- warps

def test_warp_affine(va, orc):
    from visar_amd import synth
    for cn in (1, 3, 4):
        for dt in (np.uint8, np.uint16, np.int16, np.float32):
            a = natural(90, 120, cn, cn, dt)
            ms = [synth.rotation_matrix(60, 45, 30, 0.9), synth.affine_matrix(1, 120, 90), synth.affine_matrix(2, 120, 90),
                  np.array([[1, 0, 0.5], [0, 1, -0.25]]), np.array([[2.5, 0.1, -30], [-0.2, 1.7, 4]])]
            for m in ms:
                for interp in (0, 1):
                    for border in range(5):
                        for size in [(120, 90), (77, 131)]:
                            check(va, orc, "warp_affine", a, m, size, interpolation=interp, border=border, value=37.6)


def test_warp_affine_u8_box_cases(va, orc):
    """One-channel u8 bilinear BORDER_CONSTANT (the LDS-box path): odd
    strides, offset views, tiny images, rotations up to 90 deg, scales from
    0.3 (box too wide: global taps) to 3, and maps that leave the image."""
    from visar_amd import synth
    big = natural(260, 420, 1, 91)
    views = [big[:197, :301], big[3:200, 1:306], big[5:6, 7:8], big[:2, :3], np.ascontiguousarray(big[:150, :257])]
    for a in views:
        h, w = a.shape
        ms = [synth.rotation_matrix(w / 2, h / 2, ang, sc) for ang, sc in [(8.2, 0.91), (-30, 1.2), (90, 1.0),
                                                                          (45, 0.3), (12, 3.0)]]
        ms += [np.array([[1, 0, -w * 0.6], [0, 1, h * 0.3]]), np.array([[0.5, 0.2, 3.5], [-0.1, 0.7, -2.25]])]
        for m in ms:
            for size in [(w, h), (w + 67, max(1, h - 5))]:
                want = orc.warp_affine(np.ascontiguousarray(a), m, size, 1, 0, 17.0)
                assert_same(va.warp_affine(a, m, size, 1, 0, 17.0), want, f"host {a.shape} {m.tolist()} {size}")
                dbig = dev(big)
                ys, xs = (0, 0) if a.base is None else np.unravel_index((a.__array_interface__['data'][0] -
                                                                         big.__array_interface__['data'][0]), big.shape)
                dv = dbig[ys:ys + h, xs:xs + w] if a.base is not None else dev(a)
                assert_same(va.warp_affine(dv, m, size, 1, 0, 17.0).cpu().numpy(), want, f"device {a.shape} {size}")


def test_warp_u8_layouts_agree_on_views(va, orc):
    """The one-channel u8 bilinear constant-border warps in every kernel
    layout (block layout with 8-byte tap loads issued per pixel, one or two
    blocks per wave, or after a wave-wide footprint check, and the row /
    LDS-box kernels): device views at odd offsets and strides, a view that ends at
    its buffer's last byte (the 8-byte loads must stop there), and
    perspective maps whose horizon crosses the image (w <= 0, the division's
    slow path)."""
    from visar_amd import autotune, synth
    big = natural(260, 420, 1, 93)
    dbig = dev(big)
    cases = [(0, 0, 197, 301), (3, 1, 200, 305), (5, 7, 1, 1), (0, 0, 2, 3), (60, 113, 200, 307)]
    persp = lambda w, h: ([synth.perspective_matrix(4, w, h)] if w > 1 and h > 1 else []) + [
        np.array([[1, 0.1, 2], [0.05, 1, -3], [0.004, -0.003, 1]]), np.array([[1.2, 0, 0], [0, 1.2, 0], [0, 0.01, 0.4]])]
    try:
        for layout, blocks in ((0, 0), (0, 2), (1, 0), (2, 0)):
            autotune.set_knob("warp_layout", layout)
            autotune.set_knob("warp_blocks", blocks)
            for y0, x0, h, w in cases:
                a = np.ascontiguousarray(big[y0:y0 + h, x0:x0 + w])
                dv = dbig[y0:y0 + h, x0:x0 + w]
                for m in [synth.affine_matrix(3, w, h), synth.rotation_matrix(w / 2, h / 2, -25, 0.85)]:
                    want = orc.warp_affine(a, m, (w, h), 1, 0, 5.0)
                    assert_same(va.warp_affine(dv, m, (w, h), 1, 0, 5.0).cpu().numpy(), want,
                                f"affine layout {layout} view {(y0, x0, h, w)}")
                for m in persp(w, h):
                    for size in [(w, h), (w + 33, h + 9)]:
                        want = orc.warp_perspective(a, m, size, 1, 0, 5.0)
                        assert_same(va.warp_perspective(dv, m, size, 1, 0, 5.0).cpu().numpy(), want,
                                    f"persp layout {layout} view {(y0, x0, h, w)} {m.tolist()} {size}")
            # A packed image filling its whole buffer: the last row's 8-byte loads would pass its end.
            t = dev(natural(37, 61, 1, 95))
            tn = t.cpu().numpy()
            for m in [np.array([[1, 0, 0.4], [0, 1, 0.3]]), np.array([[1, 0, 0.4], [0, 1, 0.3], [0, 0, 1]])]:
                if m.shape[0] == 2:
                    got, want = va.warp_affine(t, m, (61, 37), 1, 0, 3.0), orc.warp_affine(tn, m, (61, 37), 1, 0, 3.0)
                else:
                    got, want = va.warp_perspective(t, m, (61, 37), 1, 0, 3.0), orc.warp_perspective(tn, m, (61, 37), 1, 0, 3.0)
                assert_same(got.cpu().numpy(), want, f"layout {layout} packed {m.tolist()}")
    finally:
        autotune.set_knob("warp_layout", 0)
        autotune.set_knob("warp_blocks", 0)


def test_warp_affine_u8_c3(va, orc):
    """Three-channel u8 bilinear BORDER_CONSTANT (12-byte tap-row loads):
    device views at odd offsets and strides, a packed image filling its whole
    buffer (the last row's loads must stop at its end), tiny images, and
    maps that leave the image or shrink and enlarge it."""
    from visar_amd import synth
    big = natural(230, 410, 3, 97)
    dbig = dev(big)
    for y0, x0, h, w in [(0, 0, 200, 307), (3, 1, 197, 301), (5, 7, 1, 1), (0, 0, 2, 3), (29, 103, 200, 305)]:
        a = np.ascontiguousarray(big[y0:y0 + h, x0:x0 + w])
        dv = dbig[y0:y0 + h, x0:x0 + w]
        ms = [synth.affine_matrix(5, w, h), synth.rotation_matrix(w / 2, h / 2, -25, 0.85),
              synth.rotation_matrix(w / 2, h / 2, 60, 2.5), np.array([[1, 0, -w * 0.6], [0, 1, h * 0.3]])]
        for m in ms:
            for size in [(w, h), (w + 33, h + 9)]:
                want = orc.warp_affine(a, m, size, 1, 0, 5.0)
                assert_same(va.warp_affine(dv, m, size, 1, 0, 5.0).cpu().numpy(), want, f"view {(y0, x0, h, w)} {size}")
    t = dev(natural(37, 61, 3, 98))
    tn = t.cpu().numpy()
    for m in [np.array([[1, 0, 0.4], [0, 1, 0.3]]), synth.affine_matrix(6, 61, 37)]:
        assert_same(va.warp_affine(t, m, (61, 37), 1, 0, 3.0).cpu().numpy(), orc.warp_affine(tn, m, (61, 37), 1, 0, 3.0),
                    f"packed {m.tolist()}")
    a = synth.uniform_bytes(9, (1080, 1920, 3))
    check(va, orc, "warp_affine", a, synth.affine_matrix(0, 1920, 1080), (1920, 1080), interpolation=1, border=0)


def test_warp_affine_1080p_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(7, (1080, 1920))
    for seed in range(4):
        check(va, orc, "warp_affine", a, synth.affine_matrix(seed, 1920, 1080), (1920, 1080), interpolation=1, border=0)


def test_warp_perspective(va, orc):
    from visar_amd import synth
    for cn in (1, 3):
        for dt in (np.uint8, np.uint16, np.float32):
            a = natural(90, 120, cn, cn + 7, dt)
            ms = [synth.perspective_matrix(s, 120, 90) for s in range(3)] + [np.eye(3)]
            for m in ms:
                for interp in (0, 1):
                    for border in range(5):
                        check(va, orc, "warp_perspective", a, m, (120, 90), interpolation=interp, border=border, value=9)


def test_warp_perspective_1080p_config(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(8, (1080, 1920))
    for seed in range(3):
        check(va, orc, "warp_perspective", a, synth.perspective_matrix(seed, 1920, 1080), (1920, 1080), interpolation=1)


def undistort_maps(w, h, k1=-0.25, k2=0.08):
    fx = fy = 0.9 * w
    cx, cy = w / 2.0, h / 2.0
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    x = (xx - cx) / fx
    y = (yy - cy) / fy
    r2 = x * x + y * y
    d = 1 + k1 * r2 + k2 * r2 * r2
    return (x * d * fx + cx).astype(np.float32), (y * d * fy + cy).astype(np.float32)


def test_remap(va, orc):
    for cn in (1, 2, 3, 4):
        a = natural(72, 128, cn, 20 + cn)
        mx, my = undistort_maps(128, 72)
        for v in (0, 77.4, 300):
            check(va, orc, "remap", a, mx, my, value=v)
        rng = np.random.default_rng(cn)
        rx = rng.uniform(-3, 131, (50, 60)).astype(np.float32)
        ry = rng.uniform(-3, 75, (50, 60)).astype(np.float32)
        rx[3, 4] = np.nan
        check(va, orc, "remap", a, rx, ry, value=5)
    a = natural(720, 1280, 3, 33)
    mx, my = undistort_maps(1280, 720)
    check(va, orc, "remap", a, mx, my)


# ---------------------------------------------------------- morphology

@pytest.mark.parametrize("shape", [0, 1, 2])
def test_morphology_cases(va, orc, shape):
    for cn in (1, 3):
        a = natural(61, 83, cn, 40 + shape)
        for op in range(7):
            for ks in [(3, 3), (5, 7), (1, 1), (9, 3)]:
                check(va, orc, "morphology", a, op, shape, ks)
    g = natural(53, 41, 1, 45)
    check(va, orc, "morphology", g, 0, 0, (3, 3), iterations=3)
    check(va, orc, "morphology", g, 2, 2, (5, 5), iterations=2)
    check(va, orc, "morphology", g, 1, 1, (7, 7), iterations=0)


def test_morphology_binary_masks(va, orc):
    """{0, 255} masks (the morphology benchmark's input class), and the same
    masks with one pixel of 128: every op, element and iteration count, 1 and
    3 channels, down to one-pixel images."""
    rng = np.random.default_rng(31)
    for (h, w) in [(97, 601), (1, 1), (5, 300), (300, 5), (64, 256)]:
        for cn in (1, 3):
            shape_ = (h, w) if cn == 1 else (h, w, cn)
            m = ((rng.random(shape_) > 0.5) * 255).astype(np.uint8)
            mixed = m.copy()
            mixed.flat[(h * w * cn) // 2] = 128
            for a in (m, mixed):
                for op in range(7):
                    for shape, ks in [(0, (3, 3)), (1, (5, 5)), (2, (7, 7)), (0, (5, 3))]:
                        check(va, orc, "morphology", a, op, shape, ks)
                for it in (2, 3):
                    check(va, orc, "morphology", a, 2, 0, (3, 3), iterations=it)
                    check(va, orc, "morphology", a, 3, 1, (5, 5), iterations=it)


FUSED_ELEMENTS = [(0, (3, 3)), (1, (5, 5)), (2, (7, 7)), (2, (5, 5)), (1, (7, 3)), (0, (3, 7)), (2, (1, 5)),
                  (1, (7, 1)), (0, (1, 1)), (2, (3, 5))]


@pytest.mark.parametrize("cn", [1, 2, 3, 4])
def test_morphology_fused_chains(va, orc, cn):
    """The fused u8 pass chains against the oracle: every op, the three
    shapes at odd sizes up to 7 (square and not), 1 and 2 iterations, images
    just wider than one tile (496 bytes at a 3x3 element) and shorter than
    one, and one-pixel sides."""
    for h, w in [(45, 509 // cn + 3), (70, 1000 // cn), (1, 37), (23, 1), (2, 2)]:
        a = natural(h, w, cn, 60 + cn + h)
        for op in range(7):
            for shape, ks in FUSED_ELEMENTS:
                for it in (1, 2):
                    check(va, orc, "morphology", a, op, shape, ks, iterations=it)


def test_morphology_fused_long_chains(va, orc):
    """Chains longer than one launch takes (split into launches, the
    subtraction reading the source from memory), gradient's two chains in
    one launch and split, on binary masks and natural images."""
    rng = np.random.default_rng(77)
    m = ((rng.random((131, 777)) > 0.6) * 255).astype(np.uint8)
    g = natural(97, 530, 1, 71)
    c3 = natural(40, 300, 3, 72)
    for a in (m, g, c3):
        for op, shape, ks, it in [(2, 0, 3, 4), (5, 0, 3, 4), (6, 1, 7, 3), (5, 1, 7, 3), (4, 2, 5, 3), (4, 1, 7, 3),
                                  (4, 0, 3, 7), (0, 2, 7, 5), (1, 1, 5, 4), (3, 2, 5, 4), (6, 0, 5, 2)]:
            check(va, orc, "morphology", a, op, shape, (ks, ks), iterations=it)


MORPH_BASES = [(0, 3), (0, 5), (0, 7), (1, 3), (1, 5), (1, 7), (2, 5), (2, 7)]  # (shape, size): rect, cross, ellipse


def test_morphology_streamed_powers(va, orc):
    """The n-fold passes in one launch (the accumulator kernel, gradient's
    two chains in one launch, and the column-window kernel of morph_pow.hpp
    for the sums with too many rectangles) against the oracle's pass-by-pass
    result for every base element, op and n = 1..5 (4 runs as 2 + 2, 5 as
    3 + 2, a 7x7 cross's 3 as 2 + 1), at widths on both sides of the 496-,
    480-, 248-, 240- and 232-byte strips, heights below and above the row
    bands and the largest element's reach, and several band heights."""
    from visar_amd import autotune
    sizes = [(1, 8), (3, 12), (7, 500), (23, 496), (41, 1004), (97, 964), (13, 248), (6, 244), (30, 468)]
    try:
        for band, pow_rows in ((0, 0), (4, 8), (40, 4)):
            autotune.set_knob("morph_band", band)
            autotune.set_knob("morph_pow_rows", pow_rows)
            for shape, ks in MORPH_BASES:
                for it in (1, 2, 3, 4, 5):
                    for (h, w) in (sizes if band == 0 else sizes[2:4]):
                        a = (rnd((h, w), h * 7 + w + ks + it) > 160).astype(np.uint8) * 255 if it % 2 else rnd((h, w), h + w)
                        for op in range(7):
                            want = orc.morphology(a, op, shape, (ks, ks), it)
                            got = va.morphology(dev(a), op, shape, (ks, ks), it).cpu().numpy()
                            assert_same(got, want, f"op{op} shape{shape} k{ks} it{it} {h}x{w} band {band} {pow_rows}")
    finally:
        autotune.set_knob("morph_band", 0)
        autotune.set_knob("morph_pow_rows", 0)


def test_morphology_powers_views(va, orc):
    """The n-fold launches on dword-aligned device views (pitch and offset
    multiples of 4) with both images inside larger buffers: results equal
    the oracle's and no byte outside the destination view changes."""
    for shape, ks in MORPH_BASES:
        for it in (2, 3):
            for op in range(7):
                h, w = 37, 252
                big = dev(rnd((h + 4, w + 24), ks * 10 + it))
                src = big[2:2 + h, 8:8 + w]
                outer = torch.full((h + 6, w + 16), 77, dtype=torch.uint8, device="cuda")
                dst = outer[3:3 + h, 4:4 + w]
                va.morphology(src, op, shape, (ks, ks), it, dst=dst)
                want = orc.morphology(src.cpu().numpy(), op, shape, (ks, ks), it)
                assert_same(dst.cpu().numpy(), want, f"op{op} shape{shape} k{ks} it{it}")
                rest = outer.clone()
                rest[3:3 + h, 4:4 + w] = 77
                assert bool((rest == 77).all()), f"write outside the view op{op} shape{shape} k{ks} it{it}"


def test_morphology_band_passes(va, orc):
    """The single passes with the column window first (morph_band.hpp) against
    the oracle for every base element and op at widths on both sides of the
    248-byte strips and heights below and above the row bands, for each band
    height, and equal to morph_direct.hpp's passes (tuning "morph_pass" 1)
    on a 4096x4096 mask."""
    from visar_amd import autotune, synth
    sizes = [(1, 8), (3, 12), (9, 252), (17, 496), (40, 1000)]
    try:
        for rows in (0, 8, 16):
            autotune.set_knob("morph_pass_rows", rows)
            for shape, ks in MORPH_BASES:
                for (h, w) in (sizes if rows == 0 else sizes[2:]):
                    a = (rnd((h, w), h * 5 + w + ks) > 150).astype(np.uint8) * 255 if ks == 5 else rnd((h, w), h + w + ks)
                    for op in (0, 1, 2, 3, 5, 6):
                        want = orc.morphology(a, op, shape, (ks, ks), 1)
                        got = va.morphology(dev(a), op, shape, (ks, ks), 1).cpu().numpy()
                        assert_same(got, want, f"op{op} shape{shape} k{ks} {h}x{w} rows {rows}")
        autotune.set_knob("morph_pass_rows", 0)
        m = dev(synth.morph_mask(1, 4096))
        for shape, ks in MORPH_BASES:
            for op in (0, 1, 5, 6):
                outs = []
                for knob in (0, 1):
                    autotune.set_knob("morph_pass", knob)
                    outs.append(va.morphology(m, op, shape, (ks, ks), 1))
                assert torch.equal(outs[0], outs[1]), f"op{op} shape{shape} k{ks}"
    finally:
        autotune.set_knob("morph_pass_rows", 0)
        autotune.set_knob("morph_pass", 0)


def test_morphology_streamed_equals_per_pass(va):
    """The streamed n-fold passes equal the pass-by-pass kernels (tuning
    "morph_stream" 1) on a 1000x1400 image and a 4096x4096 mask."""
    from visar_amd import autotune, synth
    imgs = [dev(natural(1000, 1400, 1, 81)), dev(synth.morph_mask(0, 4096))]
    try:
        for x in imgs:
            for shape, ks in [(0, 3), (2, 5), (1, 7)]:
                for it in (1, 2, 3):
                    for op in range(7):
                        outs = []
                        for knob in (0, 1):
                            autotune.set_knob("morph_stream", knob)
                            outs.append(va.morphology(x, op, shape, (ks, ks), it))
                        assert torch.equal(outs[0], outs[1]), f"op{op} shape{shape} k{ks} it{it} {tuple(x.shape)}"
    finally:
        autotune.set_knob("morph_stream", 0)


def test_morphology_fused_equals_per_pass(va):
    """The fused chains equal the one-kernel-per-pass path (tuning
    "morph_chain" 1) on a 1000x1400 image, every op and element of the
    benchmark with 1-3 iterations, plus device views at odd offsets."""
    from visar_amd import autotune
    a = dev(natural(1000, 1400, 1, 73))
    big = dev(natural(300, 700, 1, 74))
    views = [big[1:, 3:613], big[5:290, 1:699]]
    try:
        for op in range(7):
            for shape, ks in [(0, 3), (2, 5), (1, 7)]:
                for it in (1, 2, 3):
                    outs = []
                    for knob in (0, 1):
                        autotune.set_knob("morph_chain", knob)
                        outs.append([va.morphology(x, op, shape, (ks, ks), it) for x in [a] + views])
                    for f, p in zip(*outs):
                        assert torch.equal(f, p), f"op{op} shape{shape} k{ks} it{it}"
    finally:
        autotune.set_knob("morph_chain", 0)


def test_morphology_depths(va, orc):
    for dt in (np.uint16, np.int16, np.float32):
        a = natural(37, 45, 3, 50, dt)
        for op in range(7):
            check(va, orc, "morphology", a, op, 2, (5, 5), iterations=2)


def test_morphology_mask_config(va, orc):
    from visar_amd import synth
    m = synth.morph_mask(0, 1024)
    for op in range(7):
        for shape, ks in [(0, (3, 3)), (2, (5, 5)), (1, (7, 7))]:
            for it in (1, 3):
                check(va, orc, "morphology", m, op, shape, ks, iterations=it)


@pytest.fixture(scope="module")
def mask4k():
    from visar_amd import synth
    return synth.morph_mask(0, 4096)


def test_morphology_4096_sample_vs_oracle(va, orc, mask4k):
    """Full-size cases of the 4096x4096 morphology workload against the oracle."""
    d = dev(mask4k)
    for op, shape, ks, it in [(2, 2, 5, 2), (5, 1, 7, 3), (4, 0, 3, 1), (1, 1, 7, 1)]:
        want = orc.morphology(mask4k, op, shape, (ks, ks), it)
        assert_same(va.morphology(d, op, shape, (ks, ks), it).cpu().numpy(), want, f"op{op} shape{shape} k{ks} it{it}")


def test_morphology_4096_all_cases_relations(va, mask4k):
    """All 63 cases at full size, tied to the single passes by the identities
    of morphologyEx (morph.zig:532-571): open = dilate^n(erode^n), close =
    erode^n(dilate^n), gradient = dilate^n - erode^n, tophat = src - open,
    blackhat = close - src (saturating), erode^n <= src <= dilate^n, and
    open/close idempotent; the single passes are checked against the oracle
    at smaller sizes."""
    d = dev(mask4k)
    for shape, ks in [(0, 3), (2, 5), (1, 7)]:
        k = (ks, ks)
        for it in (1, 2, 3):
            r = {op: va.morphology(d, op, shape, k, it) for op in range(7)}
            er, di = r[0], r[1]
            assert bool((er <= d).all()) and bool((d <= di).all())
            assert torch.equal(r[2], va.morphology(er, 1, shape, k, it))
            assert torch.equal(r[3], va.morphology(di, 0, shape, k, it))
            assert torch.equal(r[4], di - er)
            assert torch.equal(r[5], d - r[2])
            assert torch.equal(r[6], r[3] - d)
            assert bool((r[2] <= d).all()) and bool((d <= r[3]).all())
            if it == 1:
                assert torch.equal(va.morphology(r[2], 2, shape, k, 1), r[2])
                assert torch.equal(va.morphology(r[3], 3, shape, k, 1), r[3])


# ---------------------------------------------------------- background

def test_background_closed_form(va, orc):
    rng = np.random.default_rng(21)
    for (h, w) in [(240, 320), (37, 53), (1, 1), (7, 1)]:
        for shift, thr in ((4, 28), (2, 5), (8, 0), (15, 255)):
            b_host = va.Background(w, h, shift)
            b_dev = va.Background(w, h, shift)
            ref = orc.Background(w, h, shift)
            for i in range(6):
                frame = natural(h, w, 1, i)
                frame = np.clip(frame.astype(np.int32) + rng.integers(-60, 60, frame.shape), 0, 255).astype(np.uint8)
                m_host = np.empty_like(frame)
                b_host.apply(frame, m_host, thr)
                m_dev = torch.empty((h, w), dtype=torch.uint8, device="cuda")
                b_dev.apply(dev(frame), m_dev, thr)
                want = np.empty_like(frame)
                ref.apply(frame, want, thr)
                assert_same(m_host, want, "background host")
                assert_same(m_dev.cpu().numpy(), want, "background device")


# ------------------------------------------------ connected components

def test_connected_components(va, orc):
    rng = np.random.default_rng(99)
    for (h, w) in [(720, 1280), (37, 29), (1, 777), (777, 1)]:
        m = ((rng.random((h, w)) > 0.6) * 255).astype(np.uint8)
        for conn in (4, 8):
            n, lab, st, ce = va.connected_components_with_stats(m, conn)
            rn, rlab, rst, rce = orc.connected_components_with_stats(m, conn)
            assert n == rn
            assert_same(lab, rlab)
            assert_same(st, rst)
            np.testing.assert_array_equal(np.nan_to_num(ce, nan=-7), np.nan_to_num(rce, nan=-7))


def cc_masks():
    rng = np.random.default_rng(7)
    yield "noise2048", ((rng.random((2048, 2048)) > 0.55) * 255).astype(np.uint8)
    yy, xx = np.mgrid[0:301, 0:257]
    yield "checker", (((yy + xx) % 2) * 255).astype(np.uint8)
    yield "checker2", ((((yy // 2) + (xx // 2)) % 2) * 200).astype(np.uint8)
    yield "diagonals", (((xx - yy) % 5 == 0) * 255).astype(np.uint8)
    yield "antidiagonals", (((xx + yy) % 7 == 0) * 9).astype(np.uint8)
    yield "full", np.full((64, 65), 255, np.uint8)
    yield "empty", np.zeros((33, 70), np.uint8)
    yield "one", np.full((1, 1), 1, np.uint8)
    by, bx = np.mgrid[0:720, 0:1280]
    blobs = np.zeros((720, 1280), np.uint8)
    for _ in range(40):
        cy, cx, r = rng.integers(0, 720), rng.integers(0, 1280), rng.integers(3, 60)
        blobs[(by - cy) ** 2 + (bx - cx) ** 2 < r * r] = 255
    yield "blobs", blobs
    wy, wx = np.mgrid[0:200, 0:311]
    yield "waves", ((np.sin(wy / 3.0) * np.cos(wx / 5.0) > 0.2) * 255).astype(np.uint8)


def test_connected_components_device_cases(va, orc):
    """GPU labelling on device tensors (misaligned strided views included)
    against the oracle: labels, statistics rows and centroids, both
    connectivities, down to the maximum-component checkerboards."""
    for name, m in cc_masks():
        h, w = m.shape
        for conn in (4, 8):
            rn, rlab, rst, rce = orc.connected_components_with_stats(m, conn)
            big = torch.zeros((h, w + 3), dtype=torch.uint8, device="cuda")
            big[:, 1:w + 1] = dev(m)
            mv = big[:, 1:w + 1]
            lab = torch.full((h, w), -5, dtype=torch.int32, device="cuda")
            n, _, st, ce = va.connected_components_with_stats(mv, conn)
            assert n == rn, f"{name} conn {conn}: {n} labels, oracle {rn}"
            assert_same(st, rst, f"{name} conn {conn} stats")
            np.testing.assert_array_equal(np.nan_to_num(ce, nan=-7), np.nan_to_num(rce, nan=-7), f"{name} conn {conn}")
            cap = max(1, min(rn - 1, 5))
            stats = np.zeros((cap, 5), np.int32)
            cent = np.zeros((cap, 2), np.float64)
            count = va.ctypes.c_size_t(0)
            va._check(va.lib.visar_connected_components(va._ref(mv), va._ref(lab), conn,
                                                        stats.ctypes.data_as(va.ctypes.c_void_p),
                                                        cent.ctypes.data_as(va.ctypes.c_void_p), va.ctypes.c_size_t(cap),
                                                        va.ctypes.byref(count)))
            assert count.value == rn
            assert_same(lab.cpu().numpy(), rlab, f"{name} conn {conn} device labels")
            assert_same(stats, rst[:cap], f"{name} conn {conn} truncated stats")


# ------------------------------------------------------ gray downsample

def _gd_inputs(cn, k, seed):
    """The reference's shapes (videoio/frame.zig:384-470: W = 70k + 1,
    H = 9k; W = 61k + 3, H = 23k + 1) plus saturated and constant blocks."""
    rng = np.random.default_rng(seed)
    yield rng.integers(0, 256, (9 * k, 70 * k + 1, cn), dtype=np.uint8)
    yield rng.integers(0, 256, (23 * k + 1, 61 * k + 3, cn), dtype=np.uint8)
    yield np.full((3 * k + 2, 5 * k + 1, cn), 255, np.uint8)
    yield natural(4 * k + 3, 17 * k, cn, seed + 1)


@pytest.mark.parametrize("cn", [3, 4])
def test_gray_downsample(va, orc, cn):
    for k in range(1, 17):
        for a in _gd_inputs(cn, k, 100 * cn + k):
            check(va, orc, "gray_downsample", a, k)
        # Strided views (every other row of a wider buffer): host views are
        # staged; device views at unaligned offsets take the byte-load kernel,
        # 16-byte aligned ones with padded rows the vector-load kernel.
        big = rnd((6 * k, 9 * k + 5, cn), k)
        dbig = dev(big)
        for view, dview in [(big[::2, 3:], dbig[::2, 3:]), (big[1::2, 4:], dbig[1::2, 4:])]:
            want = orc.gray_downsample(np.ascontiguousarray(view), k)
            assert_same(va.gray_downsample(view, k), want, f"host view k={k}")
            assert_same(va.gray_downsample(dview, k).cpu().numpy(), want, f"device view k={k}")
        padded = rnd((5 * k + 1, 64 * k + 16, cn), k + 50)
        want = orc.gray_downsample(np.ascontiguousarray(padded[:, :61 * k + 3]), k)
        assert_same(va.gray_downsample(dev(padded)[:, :61 * k + 3], k).cpu().numpy(), want, f"padded rows k={k}")


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_gray_downsample_frames(va, orc, k):
    from visar_amd import synth
    for cn in (3, 4):
        a = synth.uniform_bytes(k + cn, (1080, 1920, cn))
        check(va, orc, "gray_downsample", a, k)


def test_gray_downsample_rect(va, orc):
    for cn in (3, 4):
        for k in (1, 2, 3, 5, 6, 8, 16):
            rng = np.random.default_rng(k * 7 + cn)
            a = rng.integers(0, 256, (23 * k + 1, 61 * k + 3, cn), dtype=np.uint8)
            whole = orc.gray_downsample(a, k)
            H, W = whole.shape
            da = dev(a)
            for _ in range(12):
                x = int(rng.integers(0, W)); y = int(rng.integers(0, H))
                w = 1 + int(rng.integers(0, W - x)); h = 1 + int(rng.integers(0, H - y))
                want = np.full((H, W), 9, np.uint8)
                want[y:y + h, x:x + w] = whole[y:y + h, x:x + w]
                assert_same(orc.gray_downsample_rect(a, k, np.full((H, W), 9, np.uint8), x, y, w, h), want, "oracle")
                got = va.gray_downsample_rect(a, k, np.full((H, W), 9, np.uint8), x, y, w, h)
                assert_same(got, want, f"rect host cn={cn} k={k} {(x, y, w, h)}")
                dd = torch.full((H, W), 9, dtype=torch.uint8, device="cuda")
                va.gray_downsample_rect(da, k, dd, x, y, w, h)
                assert_same(dd.cpu().numpy(), want, f"rect device cn={cn} k={k} {(x, y, w, h)}")
            out = np.full((H, W), 9, np.uint8)
            va.gray_downsample_rect(a, k, out, 0, 0, 0, 5)
            assert (out == 9).all()



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


@pytest.mark.parametrize("name", list(UD_DISTS))
def test_undistort(va, orc, name):
    d = UD_DISTS[name]
    for (h, w) in ((720, 1280), (37, 53)):
        K = ud_camera(w, h)
        for cn in (1, 2, 3, 4):
            check(va, orc, "undistort", natural(h, w, cn, w + cn), K, d)
        nk = K.copy()
        nk[0, 0] *= 0.8
        check(va, orc, "undistort", natural(h, w, 3, 5), K, d, new_k=nk)
    for (h, w) in ((1, 1), (3, 5), (2, 700), (4, 5000), (300, 13), (45, 300)):
        check(va, orc, "undistort", natural(h, w, 3, h + w), ud_camera(w, h), d)


def test_undistort_4k_frame(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(3, (2160, 3840, 3))
    check(va, orc, "undistort", a, ud_camera(3840, 2160), UD_DISTS["full5"])


def test_undistort_projective_and_singular_stripes(va, orc):
    h, w = 45, 300  # stripes of 4096 // 300 = 13 rows
    K = ud_camera(w, h)
    a = natural(h, w, 3, 9)
    # A new_k with a projective row: each stripe's inverse has a full third row.
    nk = K.copy()
    nk[2] = [1e-4, 2e-4, 1.0]
    check(va, orc, "undistort", a, K, UD_DISTS["full5"], new_k=nk)
    # det(new_k moved up by y0) = 1 - (27 - y0): only the stripe at y0 = 26
    # is singular and stays unwritten; the others map to huge or non-finite
    # coordinates (clamped, border 0).
    sk = np.array([[1.0, 0, 0], [0, 1, 27], [0, 1, 1]])
    want = orc.undistort(a, K, UD_DISTS["k1k2"], sk, np.full_like(a, 7))
    assert (want[26:39] == 7).all() and not (want[:26] == 7).all()
    host = va.undistort(a, K, UD_DISTS["k1k2"], sk, np.full_like(a, 7))
    assert_same(host, want, "singular stripe host")
    dd = dev(np.full_like(a, 7))
    va.undistort(dev(a), K, UD_DISTS["k1k2"], sk, dd)
    assert_same(dd.cpu().numpy(), want, "singular stripe device")
    # fx = 0: every stripe is singular, nothing is written.
    K0 = K.copy()
    K0[0, 0] = 0
    out = va.undistort(a, K, None, K0, np.full_like(a, 7))
    assert (out == 7).all()


def test_undistort_in_place(va, orc):
    a = natural(64, 96, 3, 77)
    K = ud_camera(96, 64)
    want = orc.undistort(a, K, UD_DISTS["full5"])
    da = dev(a)
    va.undistort(da, K, UD_DISTS["full5"], dst=da)
    assert_same(da.cpu().numpy(), want, "in place")


@pytest.mark.parametrize("name", list(UD_DISTS))
def test_init_undistort_rectify_map(va, orc, name):
    d = UD_DISTS[name]
    c, s = np.cos(0.02), np.sin(0.02)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    for (h, w) in ((720, 1280), (37, 53), (1, 1), (3, 5000)):
        K = ud_camera(w, h)
        nk = K.copy()
        nk[0, 0] *= 0.8
        for Rm, nkm in [(None, None), (R, nk), (None, nk), (R, None)]:
            want = orc.init_undistort_rectify_map(K, d, Rm, nkm, (w, h))
            got = va.init_undistort_rectify_map(K, d, Rm, nkm, (w, h))
            dm = (torch.empty((h, w), dtype=torch.float32, device="cuda"), torch.empty((h, w), dtype=torch.float32, device="cuda"))
            va.init_undistort_rectify_map(K, d, Rm, nkm, (w, h), out=dm)
            for g, dg, wv, ax in [(got[0], dm[0], want[0], "x"), (got[1], dm[1], want[1], "y")]:
                assert_same(g, wv, f"map{ax} host {name} {w}x{h}")
                assert_same(dg.cpu().numpy(), wv, f"map{ax} device {name} {w}x{h}")


# ------------------------------------------ pyramids, box filter, adaptive

PYR_SIZES = [(1, 1), (3, 2), (4, 5), (19, 37), (9, 70), (7, 129)]  # (h, w): pyramid.zig:615


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.int16, np.float32])
def test_pyramids(va, orc, dt):
    for (h, w) in PYR_SIZES:
        for cn in (1, 2, 3, 4, 5):
            a = natural(h, w, cn, h * 31 + w + cn, dt) if cn <= 4 else np.dstack([natural(h, w, 3, 1, dt), natural(h, w, 2, 2, dt)])
            for b in (4, 2, 1, 3):
                check(va, orc, "pyr_down", a, b)
            check(va, orc, "pyr_up", a, 4)


def test_pyramids_frames(va, orc):
    from visar_amd import synth
    for cn in (1, 3):
        a = synth.uniform_bytes(cn, (1080, 1920, cn)) if cn > 1 else synth.uniform_bytes(cn, (1080, 1920))
        check(va, orc, "pyr_down", a, 4)
        check(va, orc, "pyr_up", a[:540, :960], 4)


@pytest.mark.parametrize("cn", [1, 2, 3, 4])
def test_box_filter(va, orc, cn):
    for (h, w) in ((23, 67), (1, 1), (5, 3), (37, 53)):
        a = natural(h, w, cn, h + w + cn)
        for ks in ((3, 3), (5, 5), (7, 3), (9, 1), (15, 15), (31, 31), (1, 9), (127, 3), (3, 127), (255, 3), (3, 255),
                   (129, 129)):
            for b in range(5):
                check(va, orc, "box_filter", a, ks, True, b)
            check(va, orc, "box_filter", a, ks, False, 4)


def test_box_filter_other_depths(va, orc):
    """u16, i16 (negative samples included) and f32, windows up to 255."""
    for dt in (np.uint16, np.int16, np.float32):
        for cn in (1, 3):
            a = natural(29, 41, cn, 7, dt)
            for ks in ((3, 3), (5, 7), (21, 1), (127, 5), (255, 7), (201, 129)):
                for b in (4, 1, 0):
                    check(va, orc, "box_filter", a, ks, True, b)
                check(va, orc, "box_filter", a, ks, False, 3)


def test_box_filter_frame(va, orc):
    from visar_amd import synth
    a = synth.uniform_bytes(11, (1080, 1920))
    check(va, orc, "box_filter", a, (5, 5), True, 4)


@pytest.mark.parametrize("method", [0, 1])
def test_adaptive_threshold(va, orc, method):
    for (h, w) in ((41, 57), (1, 1), (2, 300), (200, 3)):
        a = natural(h, w, 1, h + 3 * w)
        for kind in (0, 1):
            for block in (3, 5, 7, 11, 13, 15, 31, 127):
                for delta in (0, 5, -3.5, 2.25, 300):
                    check(va, orc, "adaptive_threshold", a, 255, method, kind, block, delta)
            check(va, orc, "adaptive_threshold", a, 77.6, method, kind, 9, -1)
            if method == 0:
                for block in (129, 255):
                    check(va, orc, "adaptive_threshold", a, 255, method, kind, block, 2)
    if method == 1:
        # GaussianBlur's InvalidKernel beyond 127 taps, as the reference.
        with pytest.raises(va.VisarError) as e:
            va.adaptive_threshold(natural(20, 20, 1, 1), 255, 1, 0, 129, 0)
        assert e.value.rc == -1 and e.value.name == "InvalidKernel"
    from visar_amd import synth
    a = synth.uniform_bytes(5, (1080, 1920))
    check(va, orc, "adaptive_threshold", a, 255, method, 0, 11, 2)
    # In place on the device.
    a = natural(64, 96, 1, 8)
    want = orc.adaptive_threshold(a, 255, method, 1, 15, 3)
    da = dev(a)
    va.adaptive_threshold(da, 255, method, 1, 15, 3, dst=da)
    assert_same(da.cpu().numpy(), want, "adaptive in place")
