"""Compositions of oracle calls that the checks of the batched launch and of
the served front end compare with.

TEST INFRASTRUCTURE ONLY, like the rest of this package: the GPU tests and
the CPU-baseline leg of bench.py use these as the checker; the product
package never imports them.
"""
import numpy as np

# Op codes of a batch job (visar_amd.batch job dicts).
GAUSSIAN_BLUR, RESIZE, WARP_AFFINE, MORPHOLOGY = 0, 1, 2, 3


def batch_item(orc, job):
    """The oracle's output for one batch job (its source a torch tensor or a
    NumPy array): the plain call the job stands for."""
    kw, src = job["kw"], job["src"]
    src = src.cpu().numpy() if hasattr(src, "cpu") else src
    op = kw["op"]
    dh, dw = job["dst"].shape[:2]
    if op == GAUSSIAN_BLUR:
        return orc.gaussian_blur(src, (kw["ksize"], kw["ksize"]), kw["sigma"], border=kw["border"])
    if op == RESIZE:
        return orc.resize(src, (dw, dh), kw["interpolation"])
    if op == WARP_AFFINE:
        return orc.warp_affine(src, kw["m"], (dw, dh), kw["interpolation"], kw["border"], kw["value"])
    return orc.morphology(src, kw["morph_op"], kw["shape"], (kw["ksize"], kw["ksize"]), kw["iterations"])


def serve_front_end(orc, frames, shift=4, threshold=28, mask_threshold=127):
    """The detector front end (demos/track_lib.zig:533-544) through the
    oracle, frame by frame: GaussianBlur 5x5 -> Background.apply ->
    threshold -> open 3x3 rect. Returns the masks."""
    h, w = frames[0].shape
    bg = orc.Background(w, h, shift)
    out = []
    for f in frames:
        b = orc.gaussian_blur(f, (5, 5), 0)
        m = np.empty_like(f)
        bg.apply(b, m, threshold)
        t = orc.threshold(m, mask_threshold, 255, 0)[1]
        out.append(orc.morphology(t, 2, 0, (3, 3)))
    return out
