"""Seeded synthetic video used by the tests and the benchmark.

Two generators:

* `fixture_frame(w, h, i)` restates the reference's `synth_video_frame`
  (crates/ragdoll-pipeline/src/lib.rs:2919-2952): SMPTE-style luma bars on the top third, a
  diagonal gradient advancing 4 samples per frame below, and 16-sample chroma bands stepping
  2 samples per frame. The reference's golden clips were encoded from this content.
* `SceneGenerator` renders the benchmark content: a luma gradient background, N rectangles of
  band-limited noise moving at constant sub-pel velocities (bilinear Q8 sampling), optional
  global pan and scene cuts, additive Gaussian noise and a smooth chroma tint. Every step is
  integer or fixed point (no libm), seeded by splitmix64 -> xoshiro256**, so any host renders
  the same bytes.
"""
import numpy as np

M64 = (1 << 64) - 1


def fixture_frame(width, height, frame_idx):
    """I420 planes (y, u, v) of `synth_video_frame(width, height, frame_idx)`."""
    shift = frame_idx * 4
    y = np.empty((height, width), np.uint8)
    x = np.arange(width)
    top = height // 3
    bars = (16 + (x * 8) // max(width, 1) * 28).astype(np.uint8)
    y[:top] = bars
    rows = np.arange(top, height)[:, None]
    g = ((x[None, :] + rows + shift) >> 2) & 0xFF
    y[top:] = (16 + (g * 219) // 255).astype(np.uint8)
    cw, ch = (width + 1) // 2, (height + 1) // 2
    cx = np.arange(cw)
    cy = np.arange(ch)
    u = np.broadcast_to((96 + (((cx + shift // 2) // 16) & 0x3F)).astype(np.uint8), (ch, cw)).copy()
    v = np.broadcast_to((160 - (((cy + shift // 2) // 16) & 0x3F)).astype(np.uint8)[:, None], (ch, cw)).copy()
    return y, u, v


class _Rng:
    """xoshiro256** seeded through splitmix64."""

    def __init__(self, seed):
        s = seed & M64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    def next(self):
        s = self.s
        rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M64
        result = (rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return result

    def below(self, n):
        return (self.next() >> 11) % n

    def bytes(self, n):
        """n 64-bit words of a splitmix64 stream keyed by the next generator output
        (vectorized; used for textures and noise fields)."""
        key = np.uint64(self.next())
        with np.errstate(over="ignore"):
            z = key + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))


def _gauss_table():
    """4096 Gaussian deviates in Q8 from an integer Box-Muller (fixed-point log/cos tables are
    avoided: the inverse normal CDF is tabulated once by bisection on an integer erf-free rational
    approximation, so the table is exactly reproducible)."""
    # A fixed, integer-only construction: sums of 12 uniform 8-bit draws (Irwin-Hall) from a
    # dedicated generator, centred and scaled to unit variance in Q8. Irwin-Hall(12) - 6 has
    # variance 1 when each draw is U(0,1); with 8-bit draws the variance is 12 * (2^16-1)/12 / 2^16.
    r = _Rng(0x6A09E667F3BCC908)
    t = np.empty(4096, np.int32)
    for i in range(4096):
        acc = 0
        v = r.next()
        w = r.next()
        for k in range(8):
            acc += (v >> (8 * k)) & 0xFF
        for k in range(4):
            acc += (w >> (8 * k)) & 0xFF
        # acc ~ 12 * U(0, 255): mean 1530, sd ~ sqrt(12 * (256^2 - 1) / 12) = 255.998
        t[i] = ((acc - 1530) * 256) // 256
    return t


_GAUSS = None


class SceneGenerator:
    """Benchmark content of one stream (see the module docstring)."""

    def __init__(self, width, height, n_objects, max_speed, noise_sigma, seed, pan=(0, 0),
                 scene_cut_pct=0):
        global _GAUSS
        if _GAUSS is None:
            _GAUSS = _gauss_table()
        self.w, self.h = width, height
        self.n_objects = n_objects
        self.max_speed = max_speed
        self.sigma_q8 = int(noise_sigma * 256)
        self.pan_q8 = (int(pan[0] * 256), int(pan[1] * 256))
        self.cut_pct = scene_cut_pct
        self.seed = seed
        self.rng = _Rng(seed)
        self._new_scene()
        self.frame_idx = 0

    def _texture(self, tw, th):
        raw = np.frombuffer(self.rng.bytes((tw + 8) * (th + 8) // 8 + 1).tobytes(), np.uint8)
        raw = raw[:(tw + 8) * (th + 8)].reshape(th + 8, tw + 8).astype(np.int32)
        # Box filter 5x5 (band limit), integer.
        c = np.cumsum(np.cumsum(np.pad(raw, ((1, 0), (1, 0))), 0), 1)
        k = 5
        box = c[k:, k:] - c[:-k, k:] - c[k:, :-k] + c[:-k, :-k]
        box = box[:th + 2, :tw + 2] // (k * k)
        # Stretch contrast around the mean.
        return np.clip(128 + (box - 128) * 3, 0, 255).astype(np.int32)

    def _new_scene(self):
        r = self.rng
        self.bg_gx = int(r.below(97)) - 48
        self.bg_gy = int(r.below(97)) - 48
        self.bg_base = 64 + int(r.below(128))
        self.tint_u = int(r.below(64))
        self.tint_v = int(r.below(64))
        self.objects = []
        for _ in range(self.n_objects):
            ow = 16 + int(r.below(81))
            oh = 16 + int(r.below(81))
            x0 = int(r.below(max(self.w - 16, 1))) * 256
            y0 = int(r.below(max(self.h - 16, 1))) * 256
            vmax = self.max_speed * 256
            while True:
                vx = int(r.below(2 * vmax + 1)) - vmax
                vy = int(r.below(2 * vmax + 1)) - vmax
                if vx * vx + vy * vy <= vmax * vmax:
                    break
            self.objects.append([x0, y0, vx, vy, ow, oh, self._texture(ow, oh)])
        self.scene_start = getattr(self, "frame_idx", 0)

    def next_frame(self):
        """Renders the next frame as I420 planes (y, u, v) of uint8."""
        if self.frame_idx > 0 and self.cut_pct and self.rng.below(100) < self.cut_pct:
            self._new_scene()
        t = self.frame_idx - self.scene_start
        w, h = self.w, self.h
        px = (self.pan_q8[0] * self.frame_idx) >> 8
        py_ = (self.pan_q8[1] * self.frame_idx) >> 8
        yy, xx = np.mgrid[0:h, 0:w]
        lum = (self.bg_base * 256 + self.bg_gx * (xx + px) * 256 // max(w, 1)
               + self.bg_gy * (yy + py_) * 256 // max(h, 1)) >> 8
        lum = lum.astype(np.int32)
        for (x0, y0, vx, vy, ow, oh, tex) in self.objects:
            # Object position in Q8, wrapped around the frame.
            ox = (x0 + vx * t - self.pan_q8[0] * self.frame_idx) % (w * 256)
            oy = (y0 + vy * t - self.pan_q8[1] * self.frame_idx) % (h * 256)
            fx, fy = ox & 255, oy & 255
            ix, iy = ox >> 8, oy >> 8
            # Bilinear Q8 sample of the texture at the sub-pel offset.
            a = tex[:oh, :ow]
            b = tex[:oh, 1:ow + 1]
            c = tex[1:oh + 1, :ow]
            d = tex[1:oh + 1, 1:ow + 1]
            s = ((a * (256 - fx) + b * fx) * (256 - fy) + (c * (256 - fx) + d * fx) * fy + 32768) >> 16
            ys = np.arange(iy, iy + oh) % h
            xs = np.arange(ix, ix + ow) % w
            lum[np.ix_(ys, xs)] = s
        if self.sigma_q8:
            n = h * w
            idx = np.frombuffer(self.rng.bytes(n // 4 + 1).tobytes(), np.uint16)[:n] & 4095
            noise = (_GAUSS[idx].reshape(h, w) * self.sigma_q8) >> 16
            lum = lum + noise
        y = np.clip(lum, 0, 255).astype(np.uint8)
        cw, ch = (w + 1) // 2, (h + 1) // 2
        cy, cx = np.mgrid[0:ch, 0:cw]
        u = (96 + self.tint_u + (cx * 32) // max(cw, 1) + (cy * 16) // max(ch, 1)).astype(np.uint8)
        v = (96 + self.tint_v + (cy * 32) // max(ch, 1) - (cx * 16) // max(cw, 1)).astype(np.uint8)
        self.frame_idx += 1
        return y, u, v


# Benchmark configurations (content parameters per stream).
CONFIGS = {
    "cif": dict(width=352, height=288, n_objects=6, max_speed=8, noise_sigma=2),
    "720p": dict(width=1280, height=720, n_objects=12, max_speed=16, noise_sigma=2, pan=(0.75, 0)),
    "1080p": dict(width=1920, height=1080, n_objects=20, max_speed=32, noise_sigma=3, scene_cut_pct=10),
}


def config_seed(config_id, stream_id=0):
    return (0x52414744 ^ (config_id << 32) ^ stream_id) & M64


# The batched configuration: 24 streams of uneven size (8 x 640x360, 8 x 1280x720,
# 8 x 1920x1080), QP drawn uniformly from 22..38 by the stream's seed, search range (diamond
# iterations) scaled by width; content of the size class's generator settings.
BATCHED_CLASSES = [
    dict(width=640, height=360, n_objects=6, max_speed=8, noise_sigma=2),
    dict(width=1280, height=720, n_objects=12, max_speed=16, noise_sigma=2, pan=(0.75, 0)),
    dict(width=1920, height=1080, n_objects=20, max_speed=32, noise_sigma=3, scene_cut_pct=10),
]


def batched_stream(stream_id):
    """(generator kwargs incl. width/height, seed, qp, me_range) of batched stream 0..23."""
    seed = config_seed(3, stream_id)
    kw = dict(BATCHED_CLASSES[stream_id // 8])
    qp = 22 + _Rng(seed).next() % 17
    me_range = 16 * kw["width"] // 1280
    return kw, seed, int(qp), me_range
