"""visar_threshold_noalloc, the threshold entry point that allocates no
device memory, against visar_threshold (GPU).

Both run the same kernels; the no-allocation form takes its histogram from
a device array the library holds per GPU, so two calls in a row, the second
on a smaller image, must each give what visar_threshold gives (bit-exact
outputs, the same automatic threshold)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BINARY, BINARY_INV, TRUNC, TOZERO, TOZERO_INV, OTSU, TRIANGLE = 0, 1, 2, 3, 4, 8, 16


@pytest.fixture(scope="module")
def va():
    import visar_amd
    assert torch.cuda.is_available()
    visar_amd.set_stream(torch.cuda.current_stream())
    yield visar_amd
    visar_amd.set_stream(None)


def noalloc(va, src, thresh, maxval, kind, dst=None):
    dst = torch.empty_like(src) if dst is None else dst
    used = ctypes.c_double(-1)
    rc = va.lib.visar_threshold_noalloc(va._ref(src), va._ref(dst), ctypes.c_double(thresh), ctypes.c_double(maxval),
                                        kind, ctypes.byref(used))
    va._check(rc)
    return used.value, dst


def same(a, b):
    return np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def bimodal(seed, shape, dtype=np.uint8):
    """Two overlapping clusters of samples, so Otsu and triangle pick a
    threshold between them."""
    r = np.random.default_rng(seed)
    lo = r.normal(70, 18, shape)
    hi = r.normal(180, 25, shape)
    v = np.where(r.random(shape) < 0.35, hi, lo)
    if dtype == np.float32:
        return v.astype(np.float32)
    info = np.iinfo(dtype)
    scale = 1 if dtype == np.uint8 else (info.max // 255 if dtype == np.uint16 else 64)
    return np.clip(np.rint(v * scale), info.min, info.max).astype(dtype)


SIZES = (((1080, 1920), 1), ((97, 131), 2))  # the second call is smaller than the first


@pytest.mark.parametrize("kind,thresh", [(BINARY, 100.0), (BINARY_INV, 127.5), (TRUNC, 200.0), (TOZERO, 50.0),
                                         (TOZERO_INV, 250.0), (BINARY | OTSU, 0.0), (BINARY_INV | OTSU, 0.0),
                                         (TOZERO | TRIANGLE, 0.0), (TRUNC | TRIANGLE, 0.0)])
def test_noalloc_equals_threshold_u8(va, kind, thresh):
    for shape, seed in SIZES:
        d = torch.from_numpy(bimodal(seed, shape)).cuda()
        u0, want = va.threshold(d, thresh, 255, kind)
        u1, got = noalloc(va, d, thresh, 255, kind)
        torch.cuda.synchronize()
        assert u1 == u0, f"{shape}: threshold {u1} vs {u0}"
        assert same(got, want), f"{shape}: outputs differ"


@pytest.mark.parametrize("dtype,thresh,maxval", [(np.uint16, 30000.5, 60000.0), (np.int16, 5000.0, 12000.0),
                                                 (np.float32, 123.25, 1.0)])
@pytest.mark.parametrize("kind", [BINARY, TRUNC, TOZERO_INV])
def test_noalloc_equals_threshold_other_depths(va, dtype, thresh, maxval, kind):
    for shape, seed in SIZES:
        d = torch.from_numpy(bimodal(seed, shape + (3,), dtype)).cuda()
        u0, want = va.threshold(d, thresh, maxval, kind)
        u1, got = noalloc(va, d, thresh, maxval, kind)
        torch.cuda.synchronize()
        assert u1 == u0
        assert same(got, want)


def test_noalloc_in_place_and_views(va):
    a = bimodal(7, (300, 400))
    base = torch.from_numpy(a).cuda()
    want = va.threshold(base, 0, 255, BINARY | OTSU)[1]
    x = base.clone()
    noalloc(va, x, 0, 255, BINARY | OTSU, dst=x)  # the same view in and out
    torch.cuda.synchronize()
    assert same(x, want)
    # A strided window of a larger buffer, written into another window.
    big = torch.from_numpy(bimodal(8, (320, 512))).cuda()
    src = big[5:305, 17:417]
    out = torch.full((310, 450), 9, dtype=torch.uint8, device="cuda")
    dst = out[3:303, 21:421]
    u0, w2 = va.threshold(src, 0, 200, TOZERO | TRIANGLE)
    u1, _ = noalloc(va, src, 0, 200, TOZERO | TRIANGLE, dst=dst)
    torch.cuda.synchronize()
    assert u1 == u0 and same(dst, w2)
    assert (out[:3] == 9).all() and (out[303:] == 9).all() and (out[:, :21] == 9).all() and (out[:, 421:] == 9).all()


def test_noalloc_rejects_what_it_cannot_do_without_memory(va):
    h = np.zeros((16, 16), np.uint8)
    d = torch.zeros((16, 16), dtype=torch.uint8, device="cuda")
    with pytest.raises(va.VisarError) as e:
        noalloc(va, h, 10, 255, BINARY, dst=d)
    assert e.value.rc == -1 and e.value.name == "NotDeviceMemory"
    big = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    with pytest.raises(va.VisarError) as e:
        noalloc(va, big[0:16, 0:16], 10, 255, BINARY, dst=big[1:17, 0:16])
    assert e.value.rc == -1 and e.value.name == "Overlap"
    with pytest.raises(va.VisarError) as e:
        noalloc(va, torch.zeros((8, 8, 3), dtype=torch.uint8, device="cuda"), 0, 255, BINARY | OTSU)
    assert e.value.rc == -3
