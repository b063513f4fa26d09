"""Seeded synthetic inputs for the imgproc workloads (NumPy only).

Uniform bytes come from a splitmix64 stream (8 little-endian bytes per
output word) so the same seed gives the same image on any machine without
depending on NumPy's generator versions; matrices and batch specs use
numpy.random.default_rng (PCG64), which is stable across NumPy releases.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """The first n outputs of splitmix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = (x ^ (x >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform_bytes(seed, shape):
    """Uniform-random u8 image of `shape` from the splitmix64 stream."""
    n = int(np.prod(shape))
    words = splitmix64(seed, (n + 7) // 8)
    return words.view(np.uint8)[:n].reshape(shape).copy()


def rotation_matrix(cx, cy, angle_deg, scale):
    """getRotationMatrix2D (reference src/imgproc/geometry.zig:676-682)."""
    a = angle_deg * np.pi / 180.0
    alpha = np.cos(a) * scale
    beta = np.sin(a) * scale
    return np.array([[alpha, beta, (1 - alpha) * cx - beta * cy], [-beta, alpha, beta * cx + (1 - alpha) * cy]])


def affine_matrix(seed, w, h):
    """Seeded warp: rotation U(-30, 30) deg about the centre, scale U(0.8, 1.2),
    translation U(-10 %, 10 %) of the size."""
    rng = np.random.default_rng(seed)
    angle = rng.uniform(-30, 30)
    scale = rng.uniform(0.8, 1.2)
    tx = rng.uniform(-0.1, 0.1) * w
    ty = rng.uniform(-0.1, 0.1) * h
    m = rotation_matrix(w / 2.0, h / 2.0, angle, scale)
    m[0, 2] += tx
    m[1, 2] += ty
    return m


def perspective_transform(src_pts, dst_pts):
    """3x3 matrix mapping four points onto four points (getPerspectiveTransform)."""
    a = np.zeros((8, 8))
    b = np.zeros(8)
    for i, ((x, y), (u, v)) in enumerate(zip(src_pts, dst_pts)):
        a[2 * i] = [x, y, 1, 0, 0, 0, -u * x, -u * y]
        a[2 * i + 1] = [0, 0, 0, x, y, 1, -v * x, -v * y]
        b[2 * i] = u
        b[2 * i + 1] = v
    h = np.linalg.solve(a, b)
    return np.append(h, 1.0).reshape(3, 3)


def perspective_matrix(seed, w, h):
    """The image corners each moved by U(-5 %, 5 %) of the size."""
    rng = np.random.default_rng(seed)
    src = np.array([[0, 0], [w - 1, 0], [w - 1, h - 1], [0, h - 1]], np.float64)
    dst = src + rng.uniform(-0.05, 0.05, (4, 2)) * np.array([w, h])
    return perspective_transform(src, dst)


def _reflect101_pad(a, r):
    return np.pad(a, r, mode="reflect")


def blur5_u8(a):
    """Exact 5x5 binomial blur of a 1-channel u8 image with reflect-101 borders
    ([1 4 6 4 1] per axis, (sum + 128) >> 8: VISAR's GaussianBlur 5x5, sigma 0)."""
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = _reflect101_pad(a.astype(np.int64), 2)
    h, w = a.shape
    v = sum(k[j] * p[j:j + h, :] for j in range(5))
    o = sum(k[i] * v[:, i:i + w] for i in range(5))
    return np.minimum((o + 128) >> 8, 255).astype(np.uint8)


def morph_mask(seed, size=4096, p=0.3):
    """Binary {0, 255} mask: Bernoulli(p) noise, 5x5 blur, threshold at 128."""
    rng = np.random.default_rng(seed)
    noise = (rng.random((size, size)) < p).astype(np.uint8) * 255
    return np.where(blur5_u8(noise) > 128, 255, 0).astype(np.uint8)


BATCH_OPS = ("gauss3", "gauss5", "gauss7", "area_half", "linear_1p5", "warp_affine", "dilate5")


def batch_specs(seed, n=512, lo=64, hi=2048):
    """Per image: (width, height, channels, op, content seed, op seed)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        w = int(rng.integers(lo, hi + 1))
        h = int(rng.integers(lo, hi + 1))
        cn = 1 if rng.random() < 0.5 else 3
        op = BATCH_OPS[int(rng.integers(0, len(BATCH_OPS)))]
        specs.append((w, h, cn, op, int(rng.integers(0, 2**31)), int(rng.integers(0, 2**31))))
    return specs


class StreamScene:
    """One synthetic camera: gradient background 60 -> 140 plus N(0, 4^2) noise
    per frame, and five ellipses of intensity 210 with radii U(10, 40) moving
    at constant velocities of up to 6 px/frame, bouncing off the edges."""

    def __init__(self, seed, w=1280, h=720, n_objects=5):
        self.rng = np.random.default_rng(seed)
        self.w, self.h = w, h
        self.base = np.tile(np.linspace(60, 140, w), (h, 1))
        self.obj = []
        for _ in range(n_objects):
            rx, ry = self.rng.uniform(10, 40, 2)
            x = self.rng.uniform(rx, w - rx)
            y = self.rng.uniform(ry, h - ry)
            vx, vy = self.rng.uniform(-6, 6, 2)
            self.obj.append([x, y, vx, vy, rx, ry])
        self.yy, self.xx = np.mgrid[0:h, 0:w]

    def next_frame(self):
        img = self.base + self.rng.normal(0, 4, (self.h, self.w))
        for o in self.obj:
            x, y, vx, vy, rx, ry = o
            x0, x1 = int(max(0, x - rx)), int(min(self.w, x + rx + 1))
            y0, y1 = int(max(0, y - ry)), int(min(self.h, y + ry + 1))
            sub = ((self.xx[y0:y1, x0:x1] - x) / rx) ** 2 + ((self.yy[y0:y1, x0:x1] - y) / ry) ** 2 <= 1
            img[y0:y1, x0:x1][sub] = 210
            x, y = x + vx, y + vy
            if x < rx or x > self.w - rx:
                vx = -vx
            if y < ry or y > self.h - ry:
                vy = -vy
            o[:] = [x, y, vx, vy, rx, ry]
        return np.clip(np.rint(img), 0, 255).astype(np.uint8)
