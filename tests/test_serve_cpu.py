"""Host-side logic of the served front end (CPU): configuration defaults and
validation (rejected before any device is touched), the synthetic camera
streams, and the oracle pipeline the GPU tests check against."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def serve():
    from visar_amd import serve
    return serve


def test_config_defaults(serve):
    c = serve.default_config()
    assert (c.width, c.height, c.shift, c.threshold, c.mask_threshold) == (1280, 720, 4, 28, 127)
    assert (c.max_streams, c.max_batch, c.queue_depth, c.host_workers, c.max_inflight) == (64, 64, 4, 4, 2)
    assert c.hbm_budget == 0 and c.min_area == 250


@pytest.mark.parametrize("field,value", [("width", 0), ("height", 0), ("shift", 16), ("shift", -1),
                                         ("mask_threshold", 255), ("mask_threshold", -1), ("max_batch", 0),
                                         ("max_batch", 257), ("queue_depth", 0), ("host_workers", 0),
                                         ("max_inflight", 9), ("max_streams", 0)])
def test_invalid_config_rejected(serve, field, value):
    from visar_amd import lib
    c = serve.default_config(**{field: value})
    h = ctypes.c_void_p()
    assert lib.visar_serve_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert lib.visar_last_error() == b"InvalidArgument"
    assert not h.value


def test_null_arguments_rejected(serve):
    from visar_amd import lib
    sid = ctypes.c_int32(0)
    n = ctypes.c_size_t(0)
    assert lib.visar_serve_admit(None, ctypes.byref(sid)) == -1
    assert lib.visar_serve_submit(None, 0, None, 0, 0, None, 0) == -1
    assert lib.visar_serve_poll(None, None, 0, ctypes.byref(n), 0) == -1
    assert lib.visar_serve_create(None, None) == -1


def test_camera_streams(serve):
    cams = serve.CameraSet(3, 160, 90, seed=5, pool=4)
    f0 = cams.next_frame(0).copy()
    f1 = cams.next_frame(0).copy()
    g = cams.next_frame(1)
    assert f0.shape == (90, 160) and f0.dtype == np.uint8
    assert (f0 == 210).sum() > 500                          # five ellipses, radii >= 10
    assert 40 < f0[:, :8].mean() < 80 and 120 < f0[:, -8:].mean() < 160 or (f0[:, -8:] == 210).any()
    assert not np.array_equal(f0, f1)                       # objects move
    assert not np.array_equal(f0, g)                        # streams differ
    again = serve.CameraSet(3, 160, 90, seed=5, pool=4)
    assert np.array_equal(again.next_frame(0), f0)          # seeded


def test_oracle_pipeline_semantics(serve, orc):
    h, w = 40, 60
    bg = np.full((h, w), 100, np.uint8)
    frames = [bg.copy() for _ in range(3)]
    frames[2][10:20, 20:35] = 220
    from oracle.pipelines import serve_front_end
    masks = serve_front_end(orc, frames)
    assert not masks[0].any()                               # first frame primes the model
    assert not masks[1].any()                               # static scene
    m = masks[2]
    assert set(np.unique(m)) <= {0, 255}
    # The blurred block survives the open 3x3 (erode then dilate) as one blob.
    assert m[12:18, 22:33].all() and not m[:5].any()
    n = orc.connected_components_with_stats(m)[0]
    assert n == 2
