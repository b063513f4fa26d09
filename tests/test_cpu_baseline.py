"""The CPU baseline's row-band runner (oracle/bands.py) on the CPU: bands
partition the rows at the op's grain, and an image computed band by band on
several threads equals the same call on the whole image, for every op the
baseline splits; the baseline's decomposition of morphologyEx into banded
passes equals the oracle's own morphologyEx."""
import numpy as np
import pytest

import bench
from oracle import bands


@pytest.mark.parametrize("rows,grain,n", [(1080, 34, 16), (7, 4, 16), (100, 1, 3), (1, 8, 4), (2160, 16, 1)])
def test_bands_partition_rows(rows, grain, n):
    b = bands.bands(rows, grain, n)
    assert b[0][0] == 0 and b[-1][1] == rows and len(b) <= n
    assert all(hi > lo for lo, hi in b)
    assert all(b[k][1] == b[k + 1][0] for k in range(len(b) - 1))
    assert all(lo % grain == 0 for lo, _ in b)


def rnd(shape, seed, dtype=np.uint8):
    r = np.random.default_rng(seed)
    if dtype == np.float32:
        return r.random(shape, dtype=np.float32)
    return r.integers(0, 256, shape, dtype=np.uint8)


def run_banded(orc, ops, threads=3):
    pool = bands.BandPool(orc.lib(), threads)
    try:
        pool.run(ops)
    finally:
        pool.close()


def test_banded_ops_equal_whole_image_calls(orc):
    from visar_amd import synth
    W, H = 70, 46
    a3 = rnd((H, W, 3), 1)
    g, sm = np.empty((H, W), np.uint8), np.empty((H // 2, W // 2), np.uint8)
    sob = np.empty((H // 2, W // 2), np.int16)
    wa, wp = np.empty_like(sm), np.empty_like(sm)
    ma, mp = synth.affine_matrix(2, W // 2, H // 2), synth.perspective_matrix(2, W // 2, H // 2)
    ops = [bands.Banded(bench.oc(orc, "orc_cvt_color", a3, g, 6), H, 4),
           bands.Banded(bench.oc(orc, "orc_resize", g, sm, 1), H // 2, 2),
           bands.Banded(bench.oc(orc, "orc_sobel", sm, sob, 1, 0, 3, 1.0, 4), H // 2, 3),
           bands.Banded(bench.oc(orc, "orc_warp_affine", sm, wa, ma, 1, 0, 0.0), H // 2, 1),
           bands.Banded(bench.oc(orc, "orc_warp_perspective", sm, wp, mp, 1, 0, 0.0), H // 2, 2)]
    run_banded(orc, ops, threads=4)
    want_g = orc.cvt_color(a3, 6)
    want_s = orc.resize(want_g, (W // 2, H // 2), 1)
    np.testing.assert_array_equal(g, want_g)
    np.testing.assert_array_equal(sm, want_s)
    np.testing.assert_array_equal(sob, orc.sobel(want_s, np.int16, 1, 0, 3))
    np.testing.assert_array_equal(wa, orc.warp_affine(want_s, ma, (W // 2, H // 2), 1, 0, 0))
    np.testing.assert_array_equal(wp, orc.warp_perspective(want_s, mp, (W // 2, H // 2), 1, 0, 0))


def test_banded_filters_pyramids_and_more(orc):
    W, H = 53, 41
    b = rnd((H, W, 3), 2)
    gr = rnd((H, W), 3)
    f = rnd((H, W, 3), 4, np.float32)
    hd, wd = (H + 1) // 2, (W + 1) // 2
    outs = {k: v for k, v in [("blur", np.empty((H, W), np.uint8)), ("pd", np.empty((hd, wd, 3), np.uint8)),
                              ("pu", np.empty((2 * H, 2 * W), np.uint8)), ("pf", np.empty((hd, wd, 3), np.float32)),
                              ("box", np.empty((H, W, 3), np.uint8)), ("gd", np.empty((H // 3, W // 3), np.uint8)),
                              ("area", np.empty((17, 23, 3), np.uint8)), ("lin", np.empty((61, 79, 3), np.uint8)),
                              ("ud", np.empty((H, W, 3), np.uint8))]}
    import ctypes
    K = np.array([[40.0, 0, 26], [0, 40, 20], [0, 0, 1]])
    dist = np.array([-0.28, 0.09, 0.001, -0.0005, -0.01])
    a4 = rnd((H, W, 4), 5)
    ops = [bands.Banded(bench.oc(orc, "orc_gaussian_blur", gr, outs["blur"], 5, 5, 0.0, 0.0, 4), H, 3),
           bands.Banded(bench.oc(orc, "orc_pyr_down", b, outs["pd"], 4), hd, 2),
           bands.Banded(bench.oc(orc, "orc_pyr_up", gr, outs["pu"], 4), 2 * H, 4),
           bands.Banded(bench.oc(orc, "orc_pyr_down", f, outs["pf"], 4), hd, 2),
           bands.Banded(bench.oc(orc, "orc_box_filter", b, outs["box"], 7, 7, 1, 4), H, 3),
           bands.Banded(bench.oc(orc, "orc_gray_downsample", a4, outs["gd"], ctypes.c_uint32(3)), H // 3, 1),
           bands.Banded(bench.oc(orc, "orc_resize", b, outs["area"], 3), 17, 2),
           bands.Banded(bench.oc(orc, "orc_resize", b, outs["lin"], 1), 61, 4),
           bands.Banded(bench.oc(orc, "orc_undistort", b, outs["ud"], K, dist, ctypes.c_size_t(5), None), H, 2)]
    run_banded(orc, ops, threads=5)
    np.testing.assert_array_equal(outs["blur"], orc.gaussian_blur(gr, (5, 5), 0))
    np.testing.assert_array_equal(outs["pd"], orc.pyr_down(b))
    np.testing.assert_array_equal(outs["pu"], orc.pyr_up(gr))
    np.testing.assert_array_equal(outs["pf"], orc.pyr_down(f))
    np.testing.assert_array_equal(outs["box"], orc.box_filter(b, (7, 7)))
    np.testing.assert_array_equal(outs["gd"], orc.gray_downsample(a4, 3))
    np.testing.assert_array_equal(outs["area"], orc.resize(b, (23, 17), 3))
    np.testing.assert_array_equal(outs["lin"], orc.resize(b, (79, 61), 1))
    np.testing.assert_array_equal(outs["ud"], orc.undistort(b, K, dist))


@pytest.mark.parametrize("op", range(7))
def test_morph_decomposition_equals_morphology_ex(orc, op):
    a = np.where(rnd((57, 66), 6 + op) > 150, 255, 0).astype(np.uint8)
    for shape, ks, it in [(0, 3, 1), (2, 5, 2), (1, 7, 3)]:
        out = np.empty_like(a)
        tmp = [np.empty_like(a) for _ in range(4)]
        run_banded(orc, bench.morph_steps(orc, a, (op, shape, ks, it), out, tmp), threads=3)
        np.testing.assert_array_equal(out, orc.morphology(a, op, shape, (ks, ks), it), err_msg=f"{op} {shape} {ks} {it}")


def test_band_limits_the_rows_written(orc):
    """A call under a band writes only the band's destination rows, so the
    threads of a banded op share the work rather than repeat it."""
    import ctypes
    a = rnd((12, 9, 3), 8)
    g = np.full((12, 9), 7, np.uint8)
    f = orc.lib().orc_set_band
    f.argtypes = [ctypes.c_long, ctypes.c_long]
    f(3, 8)
    try:
        orc.cvt_color(a, 6, dst=g)
    finally:
        f(0, -1)
    want = orc.cvt_color(a, 6)
    np.testing.assert_array_equal(g[3:8], want[3:8])
    assert (g[:3] == 7).all() and (g[8:] == 7).all()
