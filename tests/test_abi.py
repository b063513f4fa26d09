"""The C ABI boundary on the CPU: the library loads, exports every symbol the
headers declare, the headers compile as C99 and C++17 (reference build.zig:91-103),
and argument validation follows the reference's order and status codes
(capi.zig:123-132 and the per-call checks) before any device work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def header_symbols():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(visar_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_hot_path():
    names = header_symbols()
    for n in ["visar_gaussian_blur", "visar_sobel", "visar_resize", "visar_warp_affine", "visar_remap",
              "visar_morphology", "visar_cvt_color", "visar_threshold", "visar_background_create",
              "visar_background_apply", "visar_background_free", "visar_warp_perspective", "visar_sep_filter_2d",
              "visar_gray_downsample", "visar_gray_downsample_rect", "visar_undistort",
              "visar_init_undistort_rectify_map", "visar_box_filter", "visar_pyr_down", "visar_pyr_up",
              "visar_adaptive_threshold", "visar_last_error", "visar_version"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    import visar_amd
    out = subprocess.run(["nm", "-D", "--defined-only", visar_amd.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    missing = [n for n in header_symbols() if n not in exported]
    assert not missing, missing
    for n in header_symbols():
        assert hasattr(visar_amd.lib, n)


@pytest.mark.parametrize("std,compiler,src", [("-std=c99", "gcc", "c"), ("-std=c++17", "g++", "cpp")])
def test_headers_compile(tmp_path, std, compiler, src):
    f = tmp_path / f"check.{src}"
    incs = "".join(f'#include "{h}"\n' for h in sorted(os.listdir(INCLUDE)) if h.endswith(".h"))
    f.write_text(incs + "int main(void) { return visar_version() == 0; }\n")
    extra = ["-pedantic"] if src == "c" else []
    subprocess.run([compiler, std, "-Wall", "-Wextra", "-Werror", *extra, "-I", INCLUDE, "-fsyntax-only", str(f)],
                   check=True)


def test_version_and_errors_without_device():
    import visar_amd as va
    assert va.version() == "0.1.0"
    a = np.zeros((4, 4), np.uint8)
    # border out of range -> UNSUPPORTED before any descriptor or device work
    with pytest.raises(va.VisarError) as e:
        va.gaussian_blur(a, (3, 3), 0, border=9)
    assert e.value.rc == -3 and e.value.name == "Unsupported"
    with pytest.raises(va.VisarError) as e:
        va.gaussian_blur(a, (-1, 3), 0)
    assert e.value.rc == -1
    # even kernel -> InvalidKernel (INVALID_ARGUMENT)
    with pytest.raises(va.VisarError) as e:
        va.gaussian_blur(a, (4, 3), 0)
    assert e.value.rc == -1 and e.value.name == "InvalidKernel"
    # shape mismatch
    with pytest.raises(va.VisarError) as e:
        va.gaussian_blur(a, (3, 3), 0, dst=np.zeros((4, 5), np.uint8))
    assert e.value.rc == -2
    # depth mismatch between src and dst
    with pytest.raises(va.VisarError) as e:
        va.gaussian_blur(a, (3, 3), 0, dst=np.zeros((4, 4), np.uint16))
    assert e.value.rc == -3
    # morphology arguments
    with pytest.raises(va.VisarError) as e:
        va.morphology(a, 0, 0, (256, 3))
    assert e.value.rc == -1
    with pytest.raises(va.VisarError) as e:
        va.morphology(a, 9, 0, (3, 3))
    assert e.value.rc == -3
    # Sobel i16 destination needs a u8 source
    with pytest.raises(va.VisarError) as e:
        va.sobel(np.zeros((4, 4), np.float32), np.int16, 1, 0)
    assert e.value.rc == -3
    # background shift range
    h = ctypes.c_void_p()
    assert va.lib.visar_background_create(4, 4, 16, ctypes.byref(h)) == -1
    # a null image
    assert va.lib.visar_cvt_color(None, None, 6) == -1
    assert va.lib.visar_last_error().decode() == "InvalidArgument"


def test_gray_downsample_errors_without_device():
    """capi.zig:176-182 and videoio/frame.zig:170-189: k <= 0 is InvalidArgument,
    a non-8U image Unsupported; k > 16, bad channels or sizes are SizeMismatch; a
    window outside dst is OutOfBounds (INVALID_ARGUMENT here); an empty window
    does nothing."""
    import visar_amd as va
    a = np.zeros((40, 60, 3), np.uint8)
    for k in (0, -3):  # capi.zig:177: k <= 0 before any image check
        with pytest.raises(va.VisarError) as e:
            va.gray_downsample(a, k, dst=np.empty((1, 1, 2), np.uint16))
        assert e.value.rc == -1 and e.value.name == "InvalidArgument"
    with pytest.raises(va.VisarError) as e:
        va.gray_downsample(a.astype(np.uint16), 2, dst=np.empty((20, 30), np.uint16))
    assert e.value.rc == -3
    for k, shape, src in [(17, (2, 3), a), (2, (20, 31), a), (2, (20, 30, 3), a),
                          (2, (20, 30), a[:, :, :1]), (3, (13, 20), a[:, :, :2])]:
        with pytest.raises(va.VisarError) as e:
            va.gray_downsample(np.ascontiguousarray(src), k, dst=np.empty(shape, np.uint8))
        assert e.value.rc == -2 and e.value.name == "SizeMismatch", (k, shape)
    d = np.full((20, 30), 7, np.uint8)
    for x, y, w, h in [(25, 0, 6, 1), (0, 19, 1, 2), (2**32 - 1, 0, 2, 1)]:
        with pytest.raises(va.VisarError) as e:
            va.gray_downsample_rect(a, 2, d, x, y, w, h)
        assert e.value.rc == -1 and e.value.name == "OutOfBounds", (x, y, w, h)
    va.gray_downsample_rect(a, 2, d, 3, 4, 0, 5)
    va.gray_downsample_rect(a, 2, d, 99, 99, 5, 0)
    assert (d == 7).all()
    with pytest.raises(va.VisarError) as e:
        va.gray_downsample_rect(a, 2, np.empty((20, 29), np.uint8), 0, 0, 0, 0)
    assert e.value.rc == -2

def test_undistort_errors_without_device():
    """capi_calib3d.zig:186-201, undistort.zig:36-41, 205-209, camera.zig:17-18."""
    import visar_amd as va
    a = np.zeros((10, 12, 3), np.uint8)
    K = np.array([[10.0, 0, 6], [0, 10, 5], [0, 0, 1]])
    cases = [((a, K, np.zeros(3)), -1, "InvalidArgument"),
             ((a, K, None, None, np.zeros((10, 13, 3), np.uint8)), -2, "SizeMismatch"),
             ((np.zeros((10, 12, 5), np.uint8), K), -3, "Unsupported"),
             ((a.astype(np.uint16), K), -3, "Unsupported")]
    for args, rc, name in cases:
        with pytest.raises(va.VisarError) as e:
            va.undistort(*args)
        assert e.value.rc == rc and e.value.name == name, args[2:]
    with pytest.raises(va.VisarError) as e:  # singular new_k * R
        va.init_undistort_rectify_map(K, None, None, np.zeros((3, 3)), (12, 10))
    assert e.value.rc == -1
    with pytest.raises(va.VisarError) as e:  # maps must be 32F
        va.init_undistort_rectify_map(K, None, None, None, (12, 10),
                                      out=(np.zeros((10, 12), np.uint8), np.zeros((10, 12), np.uint8)))
    assert e.value.rc == -3
    with pytest.raises(va.VisarError) as e:
        va.init_undistort_rectify_map(K, None, None, None, (12, 10),
                                      out=(np.zeros((10, 12), np.float32), np.zeros((10, 11), np.float32)))
    assert e.value.rc == -2

def test_box_pyramid_adaptive_errors_without_device():
    """capi.zig:208-250, 326-337 with filter.zig:617-633, pyramid.zig:65-77,
    403-408 and threshold.zig:306-310 (the Gaussian mean's 127 taps and the box
    mean's 255 give InvalidKernel)."""
    import visar_amd as va
    a = np.zeros((10, 12), np.uint8)
    cases = [(lambda: va.box_filter(a, (3, 3), True, 7), -3), (lambda: va.box_filter(a, (0, 3)), -1),
             (lambda: va.box_filter(a, (4, 3)), -1), (lambda: va.box_filter(a, (257, 3)), -1),
             (lambda: va.box_filter(a, (3, 3), dst=np.zeros((10, 13), np.uint8)), -2),
             (lambda: va.pyr_down(a, 0), -3), (lambda: va.pyr_up(a, 1), -3),
             (lambda: va.pyr_down(a, 4, np.zeros((5, 7), np.uint8)), -2),
             (lambda: va.pyr_up(a, 4, np.zeros((20, 25), np.uint8)), -2),
             (lambda: va.adaptive_threshold(a, 255, 2, 0, 3, 0), -1), (lambda: va.adaptive_threshold(a, 255, 0, 2, 3, 0), -1),
             (lambda: va.adaptive_threshold(a, 255, 0, 0, 1, 0), -1), (lambda: va.adaptive_threshold(a, 255, 0, 0, 4, 0), -1),
             (lambda: va.adaptive_threshold(a, 255, 1, 0, 129, 0), -1), (lambda: va.adaptive_threshold(a, 255, 0, 0, 257, 0), -1),
             (lambda: va.adaptive_threshold(np.zeros((10, 12, 3), np.uint8), 255, 0, 0, 3, 0), -3)]
    for i, (f, rc) in enumerate(cases):
        with pytest.raises(va.VisarError) as e:
            f()
        assert e.value.rc == rc, i

def test_last_error_is_thread_local():
    import threading

    import visar_amd as va
    a = np.zeros((4, 4), np.uint8)
    with pytest.raises(va.VisarError):
        va.gaussian_blur(a, (4, 3), 0)
    seen = []
    t = threading.Thread(target=lambda: seen.append(va.lib.visar_last_error().decode()))
    t.start()
    t.join()
    assert seen == [""]
    assert va.lib.visar_last_error().decode() == "InvalidKernel"
