"""Batched launch vs plain per-call GPU results vs the oracle (GPU).

The seeded mixed workload (sides U{64..2048}, 1 or 3 channels, ops GaussianBlur
3/5/7, resize AREA 1/2, resize LINEAR 1.5x, warpAffine, dilate 5x5 rect) runs
as one batch; every image must equal its plain C-ABI call bit for bit, and a
sample of images must equal the oracle."""
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


def _compare_batch(va, orc, n, lo, hi, seed, oracle_every, use_graph):
    from oracle import pipelines
    from visar_amd import batch
    jobs = batch.workload(seed, n, lo, hi)
    b = batch.Batch().plan(batch.items(jobs))
    tiles, wgs = b.info()
    assert tiles > 0 and wgs > 0
    for j in jobs:
        j["dst"].fill_(7)
    b.run(use_graph=use_graph)
    b.run(use_graph=use_graph)  # a second run (graph replay) gives the same bytes
    torch.cuda.synchronize()
    for i, j in enumerate(jobs):
        plain = batch.run_plain(va, j, dst=torch.empty_like(j["dst"]))
        got = j["dst"].cpu().numpy()
        assert np.array_equal(got, plain.cpu().numpy()), f"item {i} {j['spec']} differs from the plain call"
        if i % oracle_every == 0:
            assert np.array_equal(got, pipelines.batch_item(orc, j)), f"item {i} {j['spec']} differs from the oracle"


@pytest.mark.parametrize("rows", [0, 2, 4])
def test_batch_small_mixed_graph(va, orc, rows):
    """Every stencil launch shape (tuning "batch_stencil_rows" 1, 2, 4)."""
    from visar_amd import autotune
    autotune.set_knob("batch_stencil_rows", rows)
    try:
        _compare_batch(va, orc, 96, 16, 300, 11 + rows, 1, True)
    finally:
        autotune.set_knob("batch_stencil_rows", 0)


def test_batch_small_mixed_eager(va, orc):
    _compare_batch(va, orc, 64, 2, 129, 12, 1, False)


def test_batch_config_512(va, orc):
    _compare_batch(va, orc, 512, 64, 2048, 0, 32, True)


def test_batch_rejects_bad_items(va):
    from visar_amd import batch
    a = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    o = torch.zeros((32, 32), dtype=torch.uint8, device="cuda")
    with pytest.raises(RuntimeError):
        batch.Batch().plan([batch.item(batch.GAUSSIAN_BLUR, a, o, ksize=9)])
    with pytest.raises(RuntimeError):
        batch.Batch().plan([batch.item(batch.GAUSSIAN_BLUR, a.cpu().numpy(), o, ksize=3)])


def test_batch_morphology_elements(va, orc):
    """Erode and dilate in a batch with every element kind (rectangles run in
    the streamed waves, crosses and ellipses in the fused chain tiles) at
    3/5/7, one and three channels, sizes from 1 pixel to several 240-pixel
    strips, views at odd offsets: each image equals its plain call and the
    oracle."""
    from visar_amd import batch
    rng = np.random.default_rng(5)
    jobs = []
    sizes = [(1, 1), (2, 5), (7, 3), (13, 241), (37, 500), (64, 64)]
    i = 0
    for shape in (0, 1, 2):
        for k in (3, 5, 7):
            for mop in (0, 1):
                for cn in (1, 3):
                    h, w = sizes[i % len(sizes)]
                    i += 1
                    a = torch.from_numpy(rng.integers(0, 256, (h, w, cn) if cn == 3 else (h, w), dtype=np.uint8)).cuda()
                    if i % 3 == 0:
                        big = torch.zeros(h * (w * cn + 3) + 4, dtype=torch.uint8, device="cuda")
                        v = big[1:1 + h * (w * cn + 3)].view(h, w * cn + 3)[:, :w * cn]
                        v.copy_(a.reshape(h, w * cn))
                        a = v.view(h, w, cn) if cn == 3 else v
                    dst = torch.empty((h, w, cn) if cn == 3 else (h, w), dtype=torch.uint8, device="cuda")
                    jobs.append({"spec": (w, h, cn, f"morph{mop}/{shape}/{k}"), "src": a, "dst": dst,
                                 "kw": dict(op=batch.MORPHOLOGY, ksize=k, morph_op=mop, shape=shape, iterations=1)})
    b = batch.Batch().plan(batch.items(jobs))
    b.run()
    torch.cuda.synchronize()
    for j in jobs:
        got = j["dst"].cpu().numpy()
        plain = batch.run_plain(va, j, dst=torch.empty_like(j["dst"])).cpu().numpy()
        assert np.array_equal(got, plain), f"{j['spec']} differs from the plain call"
        kw = j["kw"]
        want = orc.morphology(np.ascontiguousarray(j["src"].cpu().numpy()), kw["morph_op"], kw["shape"], (kw["ksize"], kw["ksize"]))
        assert np.array_equal(got, want), f"{j['spec']} differs from the oracle"
