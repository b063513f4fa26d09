"""visar_amd: VISAR's imgproc hot path on MI355X (gfx950) through a C ABI.

This module is the host-side driver of libvisar_amd.so, shaped like the
reference's ctypes driver (tests/py/visar_ctypes.py): the same function
names, argument meaning and error behaviour (a VisarError carrying the
status and visar_last_error()). Arrays may be NumPy arrays (host memory,
staged by the library) or torch tensors on the GPU (used in place).
Importing fails loudly when the HIP library has not been built.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("VISAR_AMD_LIB", os.path.join(HERE, "libvisar_amd.so"))

_DEPTH = {np.uint8: 0, np.int8: 1, np.uint16: 2, np.int16: 3, np.int32: 4, np.float32: 5, np.float64: 6, np.float16: 7}


class Image(ctypes.Structure):
    """visar_image (include/visar.h)."""
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("depth", ctypes.c_uint32),
        ("stride", ctypes.c_size_t),
        ("data", ctypes.c_void_p),
    ]


class VisarError(RuntimeError):
    def __init__(self, rc, name):
        super().__init__(f"{rc}: {name}")
        self.rc = rc
        self.name = name


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `make -C visar_amd/csrc`")

lib = ctypes.CDLL(LIB_PATH)
lib.visar_version.restype = ctypes.c_char_p
lib.visar_build_info.restype = ctypes.c_char_p
lib.visar_last_error.restype = ctypes.c_char_p
lib.visar_gpu_get_stream.restype = ctypes.c_void_p
lib.visar_gpu_set_stream.argtypes = [ctypes.c_void_p]


def _check(rc):
    if rc != 0:
        raise VisarError(rc, lib.visar_last_error().decode())


# ------------------------------------------------------------ descriptors

def _torch():
    import torch
    return torch


def _is_tensor(a):
    return type(a).__module__.startswith("torch")


_TORCH_DEPTH = None


def _tdepth(t):
    global _TORCH_DEPTH
    if _TORCH_DEPTH is None:
        torch = _torch()
        _TORCH_DEPTH = {torch.uint8: 0, torch.int8: 1, torch.uint16: 2, torch.int16: 3, torch.int32: 4,
                        torch.float32: 5, torch.float64: 6, torch.float16: 7}
    return _TORCH_DEPTH[t.dtype]


def _geometry(a):
    """(shape, element bytes, byte strides, depth code, address) of a NumPy
    array or a torch tensor."""
    if _is_tensor(a):
        es = a.element_size()
        return tuple(a.shape), es, tuple(k * es for k in a.stride()), _tdepth(a), a.data_ptr()
    code = _DEPTH.get(a.dtype.type)
    if code is None:
        raise ValueError(f"unsupported dtype {a.dtype}")
    return a.shape, a.itemsize, a.strides, code, a.ctypes.data


def desc(a):
    """The visar_image of an (H, W) or (H, W, C) array or tensor. Rows may be
    padded (any byte stride); within a row, pixels and channels are packed."""
    shape, es, steps, depth, addr = _geometry(a)
    if len(shape) not in (2, 3):
        raise ValueError("expected 2-D or 3-D array")
    h, w = shape[0], shape[1]
    c = shape[2] if len(shape) == 3 else 1
    # Packed: the pixel step is one pixel and the channel step one element;
    # a step along an axis of extent 1 is never taken and not checked.
    for extent, step, packed in zip(shape[1:], steps[1:], (c * es, es)):
        if extent > 1 and step != packed:
            raise ValueError("pixels must be packed")
    row = w * c * es
    return Image(w, h, c, depth, steps[0] if h > 1 else row, addr)


_tls = threading.local()


def _follow_torch(a):
    """Until the caller picks a stream with set_stream, calls that take torch
    tensors run on torch's current stream of the tensor's device, so they are
    ordered after the tensors' own fills and copies and before what torch
    does with the results."""
    if not getattr(_tls, "explicit", False):
        import torch
        lib.visar_gpu_set_stream(ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream))


def _ref(a):
    if _is_tensor(a) and a.is_cuda:
        _follow_torch(a)
    return ctypes.byref(desc(a))


def _empty(like, shape, dtype=None):
    if _is_tensor(like):
        torch = _torch()
        dt = like.dtype if dtype is None else dtype
        if not isinstance(dt, torch.dtype):
            dt = {np.uint8: torch.uint8, np.int16: torch.int16, np.uint16: torch.uint16, np.float32: torch.float32,
                  np.int32: torch.int32}[np.dtype(dt).type]
        return torch.empty(shape, dtype=dt, device=like.device)
    return np.empty(shape, dtype=like.dtype if dtype is None else dtype)


def _like(a, size=None, channels=None):
    """Shape of an output like `a`: `size` (width, height) and `channels`
    replace a's own; one channel gives a 2-D shape."""
    w, h = (a.shape[1], a.shape[0]) if size is None else size
    c = (a.shape[2] if a.ndim == 3 else 1) if channels is None else channels
    return (h, w, c) if c > 1 else (h, w)


# ------------------------------------------------------------ library

def version():
    return lib.visar_version().decode()


def build_info():
    return lib.visar_build_info().decode()


def set_num_threads(n):
    lib.visar_set_num_threads(n)


def get_num_threads():
    return lib.visar_get_num_threads()


def set_stream(stream):
    """Stream for this thread's calls: a torch.cuda.Stream or a raw hipStream_t
    int (0 is HIP's default stream); None returns to the library's own stream
    (and to following torch's current stream for calls on torch tensors)."""
    if stream is None:
        _tls.explicit = False
        _check(lib.visar_gpu_reset_stream())
        return
    if hasattr(stream, "cuda_stream"):
        stream = stream.cuda_stream
    _tls.explicit = True
    _check(lib.visar_gpu_set_stream(ctypes.c_void_p(stream)))


def get_stream():
    return lib.visar_gpu_get_stream()


def set_sync(sync):
    _check(lib.visar_gpu_set_sync(int(bool(sync))))


def synchronize():
    _check(lib.visar_gpu_synchronize())


def device_info():
    buf = ctypes.create_string_buffer(128)
    _check(lib.visar_gpu_device_info(buf, 128))
    return buf.value.decode()


# ------------------------------------------------------------ imgproc

_COLOR_CN = {6: (3, 1), 7: (3, 1), 10: (4, 1), 11: (4, 1)}


def cvt_color(a, code, dst=None):
    dst = _empty(a, _like(a, channels=_COLOR_CN[code][1])) if dst is None else dst
    _check(lib.visar_cvt_color(_ref(a), _ref(dst), code))
    return dst


def resize(a, size, interpolation=1, dst=None):
    dst = _empty(a, _like(a, size)) if dst is None else dst
    _check(lib.visar_resize(_ref(a), _ref(dst), interpolation))
    return dst


def gaussian_blur(a, ksize, sigma_x, sigma_y=0, border=4, dst=None):
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    _check(lib.visar_gaussian_blur(_ref(a), _ref(dst), ksize[0], ksize[1], ctypes.c_double(sigma_x),
                                   ctypes.c_double(sigma_y), border))
    return dst


def sobel(a, ddepth, dx, dy, ksize=3, scale=1.0, border=4, dst=None):
    dst = _empty(a, tuple(a.shape), ddepth) if dst is None else dst
    _check(lib.visar_sobel(_ref(a), _ref(dst), dx, dy, ksize, ctypes.c_double(scale), border))
    return dst


def sep_filter_2d(a, ddepth, kx, ky, border=4, dst=None):
    dst = _empty(a, tuple(a.shape), ddepth) if dst is None else dst
    kx = np.ascontiguousarray(kx, np.float32)
    ky = np.ascontiguousarray(ky, np.float32)
    _check(lib.visar_sep_filter_2d(_ref(a), _ref(dst), kx.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(kx)),
                                   ky.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(len(ky)), border))
    return dst


def threshold(a, thresh, maxval, kind, dst=None):
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    used = ctypes.c_double(0)
    _check(lib.visar_threshold(_ref(a), _ref(dst), ctypes.c_double(thresh), ctypes.c_double(maxval), kind,
                               ctypes.byref(used)))
    return used.value, dst


def morphology(a, op, shape, ksize, iterations=1, dst=None):
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    _check(lib.visar_morphology(_ref(a), _ref(dst), op, shape, ksize[0], ksize[1], iterations))
    return dst


def warp_affine(a, m, size, interpolation=1, border=0, value=0.0, dst=None):
    dst = _empty(a, _like(a, size)) if dst is None else dst
    arr = (ctypes.c_double * 6)(*np.asarray(m, np.float64).ravel()[:6])
    _check(lib.visar_warp_affine(_ref(a), _ref(dst), arr, interpolation, border, ctypes.c_double(value)))
    return dst


def warp_perspective(a, m, size, interpolation=1, border=0, value=0.0, dst=None):
    dst = _empty(a, _like(a, size)) if dst is None else dst
    arr = (ctypes.c_double * 9)(*np.asarray(m, np.float64).ravel()[:9])
    _check(lib.visar_warp_perspective(_ref(a), _ref(dst), arr, interpolation, border, ctypes.c_double(value)))
    return dst


def remap(a, mapx, mapy, value=0.0, dst=None):
    dst = _empty(a, _like(a, (mapx.shape[1], mapx.shape[0]))) if dst is None else dst
    _check(lib.visar_remap(_ref(a), _ref(dst), _ref(mapx), _ref(mapy), ctypes.c_double(value)))
    return dst


def box_filter(a, ksize, normalize=True, border=4, dst=None):
    """boxFilter / blur (filter.zig:617 boxFilter; tests/py/visar_ctypes.py:121): u8, u16, f32."""
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    _check(lib.visar_box_filter(_ref(a), _ref(dst), ksize[0], ksize[1], int(bool(normalize)), border))
    return dst


def pyr_down(a, border=4, dst=None):
    """pyramid.zig:65 pyrDown (tests/py/visar_ctypes.py:133): dst ((w + 1) // 2, (h + 1) // 2)."""
    dst = _empty(a, _like(a, ((a.shape[1] + 1) // 2, (a.shape[0] + 1) // 2))) if dst is None else dst
    _check(lib.visar_pyr_down(_ref(a), _ref(dst), border))
    return dst


def pyr_up(a, border=4, dst=None):
    """pyramid.zig:403 pyrUp (tests/py/visar_ctypes.py:139): dst (2w, 2h), reflect-101 only."""
    dst = _empty(a, _like(a, (a.shape[1] * 2, a.shape[0] * 2))) if dst is None else dst
    _check(lib.visar_pyr_up(_ref(a), _ref(dst), border))
    return dst


def adaptive_threshold(a, maxval, method, kind, block, delta, dst=None):
    """threshold.zig:306 adaptiveThreshold: u8, one channel; method 0 MEAN_C, 1 GAUSSIAN_C;
    kind 0 BINARY, 1 BINARY_INV."""
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    _check(lib.visar_adaptive_threshold(_ref(a), _ref(dst), ctypes.c_double(maxval), method, kind, block,
                                        ctypes.c_double(delta)))
    return dst


def gray_downsample(a, k, dst=None):
    """BT.601 gray of a BGR/BGRA u8 image shrunk by k (1..16) with box averaging
    (videoio/frame.zig:170 grayDownsample; tests/py/visar_ctypes.py:102); dst is (h // k, w // k)."""
    dst = _empty(a, (a.shape[0] // k, a.shape[1] // k)) if dst is None else dst
    _check(lib.visar_gray_downsample(_ref(a), _ref(dst), int(k)))
    return dst


def gray_downsample_rect(a, k, dst, x, y, w, h):
    """Only dst's (x, y, w, h) window of gray_downsample (videoio/frame.zig:182
    grayDownsampleRect); the rest of dst keeps its contents."""
    _check(lib.visar_gray_downsample_rect(_ref(a), _ref(dst), int(k), ctypes.c_uint32(x),
                                          ctypes.c_uint32(y), ctypes.c_uint32(w), ctypes.c_uint32(h)))
    return dst


def _f64_arg(m, n):
    """(array, c_void_p) of the first n float64 values of m, or (None, None)."""
    if m is None:
        return None, None
    arr = np.ascontiguousarray(np.asarray(m, np.float64).ravel()[:n])
    return arr, arr.ctypes.data_as(ctypes.c_void_p)


def undistort(a, K, dist=None, new_k=None, dst=None):
    """Lens undistortion of a u8 image, 1-4 channels (calib3d/undistort.zig:205
    undistort; tests/py/visar_calib3d.py:146); new_k defaults to K."""
    dst = _empty(a, tuple(a.shape)) if dst is None else dst
    kk, kp = _f64_arg(K, 9)
    dd, dp = _f64_arg(dist, 8)
    nk, nkp = _f64_arg(new_k, 9)
    _check(lib.visar_undistort(_ref(a), _ref(dst), kp, dp, ctypes.c_size_t(0 if dd is None else dd.size), nkp))
    return dst


def init_undistort_rectify_map(K, dist, R, new_k, size, out=None):
    """f32 maps (mapx, mapy) of size (w, h) (undistort.zig:82 initUndistortRectifyMap;
    tests/py/visar_calib3d.py:155); `out` = (mapx, mapy) arrays or tensors to fill."""
    w, h = size
    mx, my = (np.empty((h, w), np.float32), np.empty((h, w), np.float32)) if out is None else out
    kk, kp = _f64_arg(K, 9)
    dd, dp = _f64_arg(dist, 8)
    rr, rp = _f64_arg(R, 9)
    nk, nkp = _f64_arg(None if new_k is None else np.asarray(new_k, np.float64)[:3, :3], 9)
    _check(lib.visar_init_undistort_rectify_map(kp, dp, ctypes.c_size_t(0 if dd is None else dd.size), rp, nkp,
                                                _ref(mx), _ref(my)))
    return mx, my


class Background:
    """Running-average background model with its state in device memory."""

    def __init__(self, width, height, shift=4):
        self._h = ctypes.c_void_p()
        _check(lib.visar_background_create(ctypes.c_uint32(width), ctypes.c_uint32(height), shift,
                                           ctypes.byref(self._h)))

    def apply(self, frame, mask, threshold=28):
        _check(lib.visar_background_apply(self._h, _ref(frame), _ref(mask), threshold))
        return mask

    def __del__(self):
        # At interpreter exit the module globals may already be gone.
        if getattr(self, "_h", None) and self._h.value and lib is not None:
            lib.visar_background_free(self._h)
            self._h = None


def connected_components_with_stats(a, connectivity=8):
    """(count, labels i32, stats (count, 5) i32, centroids (count, 2) f64) like the reference driver
    (tests/py/visar_ctypes.py:345-361): room for 4096 labels first, then exactly the reported count."""
    h, w = a.shape[:2]
    labels = np.empty((h, w), np.int32)
    src = a.cpu().numpy() if _is_tensor(a) else a
    cap = min(h * w + 1, 4096)
    while True:
        stats = np.empty((cap, 5), np.int32)
        cent = np.empty((cap, 2), np.float64)
        count = ctypes.c_size_t(0)
        _check(lib.visar_connected_components(_ref(src), _ref(labels), connectivity,
                                              stats.ctypes.data_as(ctypes.c_void_p),
                                              cent.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(cap),
                                              ctypes.byref(count)))
        n = count.value
        if n <= cap:
            return n, labels, stats[:n].copy(), cent[:n].copy()
        cap = n
