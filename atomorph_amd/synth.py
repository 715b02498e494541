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
                   random colour per blob -- the 512^2 / 1024^2 workloads.
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


# The named workloads.
def workload(name):
    if name == "ellipses64":      # 2 x 64^2, 6 ellipses, seed 1
        return ellipse_frames(seed=1, size=64, count=6, nframes=2)
    if name == "blobs512":        # 2 x 512^2, 40 blobs per frame
        return noise_blob_frames(512, 2, 40, 0.36, seed=2)
    if name == "blobs1024":       # 4 x 1024^2, 150 blobs per frame
        return noise_blob_frames(1024, 4, 150, 0.18, seed=3)
    raise KeyError(name)


def to_words(frame):
    """(H, W, 4) uint8 -> (H*W,) uint32 packed r | g<<8 | b<<16 | a<<24."""
    f = np.ascontiguousarray(frame, dtype=np.uint8)
    return f.view(np.uint32).reshape(-1)
