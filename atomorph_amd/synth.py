"""Synthetic RGBA8 key frames for the morph workloads (no datasets needed).

All randomness comes from a SplitMix64 stream, so the frames are identical on
every machine.  Frames are numpy uint8 arrays of shape (H, W, 4), transparent
background (alpha 0), opaque blobs (alpha 255).

Workloads:
  ellipse_frames   two 64x64 frames of 6 filled axis-aligned ellipses each
                   (centres uniform, radii 4-16 px, RGB uniform) -- the
                   CPU-runnable parity case;
  noise_blob_frames  thresholded fBm gradient noise (octaves 4, persistence
                   0.5), the N largest 4-connected components kept, one
                   random colour per blob -- the 512^2 / 1024^2 workloads;
  powerlaw_blob_frame  N separate blobs (rotated filled ellipses) whose areas
                   follow a power law p(s) ~ s^-alpha on [smin, smax] -- the
                   2048^2 workloads: on [8, 60000] (about 0.2 M blob pixels
                   per frame), and on [200, 60000] (about 2 M per frame, 4 M
                   over the pair).
"""
import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """Portable 64-bit generator (Steele, Lea & Flood 2014)."""

    def __init__(self, seed):
        self.state = seed & _M64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Integer in [lo, hi] inclusive."""
        return lo + self.next_u64() % (hi - lo + 1)

    def permutation(self, n):
        p = list(range(n))
        for i in range(n - 1, 0, -1):
            j = self.next_u64() % (i + 1)
            p[i], p[j] = p[j], p[i]
        return np.array(p, dtype=np.int64)


def ellipse_frames(seed=1, size=64, count=6, nframes=2):
    """Key frames of `count` filled ellipses each, painted in order."""
    rng = SplitMix64(seed)
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    frames = []
    for _ in range(nframes):
        img = np.zeros((size, size, 4), np.uint8)
        for _ in range(count):
            cx, cy = rng.uniform(0, size), rng.uniform(0, size)
            rx, ry = rng.uniform(4, 16), rng.uniform(4, 16)
            col = [rng.randint(0, 255) for _ in range(3)]
            inside = ((xx - cx) / rx) ** 2 + ((yy - cy) / ry) ** 2 <= 1.0
            img[inside, 0:3] = col
            img[inside, 3] = 255
        frames.append(img)
    return frames


def _gradient_noise(h, w, period, rng):
    """2-D gradient (Perlin-style) noise with `period`-pixel cells."""
    gx_n, gy_n = w // period + 2, h // period + 2
    ang = np.array([rng.uniform(0, 2 * np.pi) for _ in range(gx_n * gy_n)]).reshape(gy_n, gx_n)
    gvx, gvy = np.cos(ang), np.sin(ang)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    fx, fy = xs / period, ys / period
    x0, y0 = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
    tx, ty = fx - x0, fy - y0

    def dot(ix, iy, dx, dy):
        return gvx[iy, ix] * dx + gvy[iy, ix] * dy

    n00 = dot(x0, y0, tx, ty)
    n10 = dot(x0 + 1, y0, tx - 1, ty)
    n01 = dot(x0, y0 + 1, tx, ty - 1)
    n11 = dot(x0 + 1, y0 + 1, tx - 1, ty - 1)
    sx = tx * tx * tx * (tx * (tx * 6 - 15) + 10)
    sy = ty * ty * ty * (ty * (ty * 6 - 15) + 10)
    a = n00 + sx * (n10 - n00)
    b = n01 + sx * (n11 - n01)
    return a + sy * (b - a)


def fbm(h, w, base_period, rng, octaves=4, persistence=0.5):
    total = np.zeros((h, w))
    amp, period = 1.0, float(base_period)
    for _ in range(octaves):
        total += amp * _gradient_noise(h, w, max(int(round(period)), 2), rng)
        amp *= persistence
        period /= 2.0
    return total


def noise_blob_frame(size, nblobs, coverage, seed, base_period=None):
    """One frame: the `nblobs` largest components of {fbm > t}, where t is the
    (1 - coverage) quantile of the field.  Returns (frame, blob sizes)."""
    from scipy import ndimage

    rng = SplitMix64(seed)
    if base_period is None:
        base_period = max(8, int(size / (np.sqrt(nblobs) * 1.5)))
    field = fbm(size, size, base_period, rng)
    t = np.quantile(field, 1.0 - coverage)
    lab, n = ndimage.label(field > t, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    sizes = np.bincount(lab.ravel())[1:]
    keep = np.argsort(-sizes, kind="stable")[:nblobs] + 1
    img = np.zeros((size, size, 4), np.uint8)
    for k in keep:
        m = lab == k
        img[m, 0:3] = [rng.randint(0, 255) for _ in range(3)]
        img[m, 3] = 255
    return img, np.sort(sizes[keep - 1])[::-1]


def noise_blob_frames(size, nframes, nblobs, coverage, seed):
    return [noise_blob_frame(size, nblobs, coverage, seed * 1000003 + f)[0] for f in range(nframes)]


def powerlaw_sizes(n, alpha, smin, smax, rng):
    """n areas from p(s) ~ s^-alpha on [smin, smax] by inverting the CDF."""
    e = 1.0 - alpha
    lo, hi = smin ** e, smax ** e
    return np.array([int(round((lo + rng.uniform() * (hi - lo)) ** (1.0 / e))) for _ in range(n)], np.int64)


def _ellipse_mask(area, aspect, theta):
    """Filled ellipse of about `area` pixels, its largest 4-connected part."""
    from scipy import ndimage
    a = np.sqrt(area * aspect / np.pi)
    b = np.sqrt(area / (aspect * np.pi))
    r = int(np.ceil(a)) + 1
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1].astype(np.float64)
    c, s = np.cos(theta), np.sin(theta)
    u, v = xx * c + yy * s, -xx * s + yy * c
    m = (u / a) ** 2 + (v / b) ** 2 <= 1.0
    if not m.any():
        m[r, r] = True
    lab, n = ndimage.label(m, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    if n > 1:
        m = lab == (np.argmax(np.bincount(lab.ravel())[1:]) + 1)
    ys, xs = np.nonzero(m)
    return m[ys.min():ys.max() + 1, xs.min():xs.max() + 1]


def powerlaw_blob_frame(size, nblobs, alpha, smin, smax, seed, tries=200):
    """One frame of `nblobs` blobs that never touch (a free pixel between
    any two), areas from powerlaw_sizes, placed largest first at uniform
    random positions, each with a random colour.  Returns (frame, areas)."""
    from scipy import ndimage
    rng = SplitMix64(seed)
    areas = np.sort(powerlaw_sizes(nblobs, alpha, smin, smax, rng))[::-1]
    taken = np.zeros((size, size), bool)        # blob pixels and their 4-neighbours
    img = np.zeros((size, size, 4), np.uint8)
    plus = np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], bool)
    placed = []
    for area in areas:
        m = _ellipse_mask(float(area), rng.uniform(1.0, 3.0), rng.uniform(0.0, np.pi))
        mh, mw = m.shape
        if mh > size or mw > size:
            raise ValueError("blob larger than the frame")
        for _ in range(tries):
            y0, x0 = rng.randint(0, size - mh), rng.randint(0, size - mw)
            if not (taken[y0:y0 + mh, x0:x0 + mw] & m).any():
                break
        else:
            raise RuntimeError("no free place for a blob; lower the count or the sizes")
        grown = ndimage.binary_dilation(np.pad(m, 1), structure=plus)
        ty0, tx0 = y0 - 1, x0 - 1
        gy0, gx0 = max(ty0, 0), max(tx0, 0)
        gy1, gx1 = min(ty0 + mh + 2, size), min(tx0 + mw + 2, size)
        taken[gy0:gy1, gx0:gx1] |= grown[gy0 - ty0:gy1 - ty0, gx0 - tx0:gx1 - tx0]
        region = img[y0:y0 + mh, x0:x0 + mw]
        region[m, 0:3] = [rng.randint(0, 255) for _ in range(3)]
        region[m, 3] = 255
        placed.append(int(m.sum()))
    return img, np.array(placed, np.int64)


# The named workloads.
def noisy_ellipse_frames(rng, n, W, H, nell=5):
    """n RGBA8 frames of filled ellipses with per-pixel colour noise,
    quantised to steps of 16 so that equal colour distances (ties) occur:
    inputs for blob detection below the default colour threshold."""
    out = []
    yy, xx = np.mgrid[0:H, 0:W]
    for _ in range(n):
        img = np.zeros((H, W, 4), np.uint8)
        for _ in range(nell):
            cx, cy = rng.uniform(0, W), rng.uniform(0, H)
            rx, ry = rng.uniform(3, W / 3), rng.uniform(3, H / 3)
            m = ((xx - cx) / rx) ** 2 + ((yy - cy) / ry) ** 2 <= 1
            col = np.clip(rng.integers(0, 256, 3) + rng.integers(-30, 31, (H, W, 3)), 0, 255)
            img[m, :3] = col[m]
            img[m, 3] = 255
        img[..., :3] = (img[..., :3] // 16) * 16
        out.append(img)
    return out


def workload(name):
    if name == "ellipses64":      # 2 x 64^2, 6 ellipses, seed 1
        return ellipse_frames(seed=1, size=64, count=6, nframes=2)
    if name == "blobs512":        # 2 x 512^2, 40 blobs per frame
        return noise_blob_frames(512, 2, 40, 0.36, seed=2)
    if name == "blobs1024":       # 4 x 1024^2, 150 blobs per frame
        return noise_blob_frames(1024, 4, 150, 0.18, seed=3)
    if name == "powerlaw2048":    # 2 x 2048^2, 1200 blobs, areas ~ s^-1.8 on [8, 60000]
        return [powerlaw_blob_frame(2048, 1200, 1.8, 8, 60000, seed=5 * 1000003 + f)[0] for f in range(2)]
    if name == "powerlaw2048_4m":  # the same law on [200, 60000]: about 2 M blob pixels per frame
        return [powerlaw_blob_frame(2048, 1200, 1.8, 200, 60000, seed=7 * 1000003 + f)[0] for f in range(2)]
    raise KeyError(name)


def to_words(frame):
    """(H, W, 4) uint8 -> (H*W,) uint32 packed r | g<<8 | b<<16 | a<<24."""
    f = np.ascontiguousarray(frame, dtype=np.uint8)
    return f.view(np.uint32).reshape(-1)
