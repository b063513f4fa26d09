"""C-ABI calls captured in a HIP graph with two streams in flight, as the
benchmark replays them: every replay must equal the same calls made eagerly
on one stream, bit for bit (GPU).

Each of four steps runs the 4K chain's calls (cvtColor, exact-half resize,
Sobel, warpAffine, warpPerspective) plus a box filter, pyrDown, undistort and
grayDownsample on its own frame; even steps go on the capture stream, odd
steps on a second stream forked from it inside the graph. The graph is
replayed three times, with unrelated eager work and a synchronisation
between replays, and checked after each replay.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_two_lane_graph_replays_match_eager():
    import visar_amd as va
    from visar_amd import synth
    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    H, W = 720, 1280
    K = np.array([[900.0, 0, 640], [0, 900, 360], [0, 0, 1]])
    dist = np.array([-0.28, 0.09, 0.001, -0.0005, -0.01])
    ma = synth.affine_matrix(0, W // 2, H // 2)
    mp = synth.perspective_matrix(0, W // 2, H // 2)
    frames = [torch.from_numpy(synth.uniform_bytes(40 + i, (H, W, 3))).to(dev) for i in range(4)]

    def outputs():
        return {
            "gray": torch.empty((H, W), dtype=torch.uint8, device=dev),
            "small": torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev),
            "sob": torch.empty((H // 2, W // 2), dtype=torch.int16, device=dev),
            "wa": torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev),
            "wp": torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev),
            "box": torch.empty((H, W, 3), dtype=torch.uint8, device=dev),
            "pyr": torch.empty(((H + 1) // 2, (W + 1) // 2, 3), dtype=torch.uint8, device=dev),
            "und": torch.empty((H, W, 3), dtype=torch.uint8, device=dev),
            "gd": torch.empty((H // 4, W // 4), dtype=torch.uint8, device=dev),
        }

    def step(f, o):
        sz = (W // 2, H // 2)
        va.cvt_color(f, 6, dst=o["gray"])
        va.resize(o["gray"], sz, 1, dst=o["small"])
        va.sobel(o["small"], np.int16, 1, 0, 3, dst=o["sob"])
        va.warp_affine(o["small"], ma, sz, 1, 0, 0, dst=o["wa"])
        va.warp_perspective(o["small"], mp, sz, 1, 0, 0, dst=o["wp"])
        va.box_filter(f, (5, 5), dst=o["box"])
        va.pyr_down(f, dst=o["pyr"])
        va.undistort(f, K, dist, dst=o["und"])
        va.gray_downsample(f, 4, dst=o["gd"])

    main = torch.cuda.current_stream()
    va.set_stream(main)
    want = []
    for i in range(4):
        o = outputs()
        step(frames[i], o)
        want.append(o)
    torch.cuda.synchronize()
    want = [{k: v.cpu().numpy() for k, v in o.items()} for o in want]

    try:
        got = [outputs() for _ in range(4)]
        va.set_sync(False)  # a captured call must not wait for its stream
        cap = torch.cuda.Stream()
        side = torch.cuda.Stream()
        cap.wait_stream(main)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(cap):
            with torch.cuda.graph(g, stream=cap):
                fork = torch.cuda.Event()
                fork.record(cap)
                side.wait_event(fork)
                for i in range(4):
                    st = side if i % 2 else cap
                    va.set_stream(st)
                    with torch.cuda.stream(st):
                        step(frames[i], got[i])
                va.set_stream(cap)
                cap.wait_stream(side)
        va.set_stream(main)
        main.wait_stream(cap)
        torch.cuda.synchronize()
        noise = torch.empty((H, W), dtype=torch.uint8, device=dev)
        for rep in range(3):
            for o in got:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for i in range(4):
                for k, v in got[i].items():
                    assert np.array_equal(v.cpu().numpy(), want[i][k]), (rep, i, k)
            va.gaussian_blur(torch.from_numpy(want[rep]["gray"]).to(dev), (5, 5), 0, dst=noise)
            torch.cuda.synchronize()
    finally:
        va.set_sync(True)
        va.set_stream(None)
