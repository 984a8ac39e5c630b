"""Seeded synthetic 8-bit 4:2:0 content and the encoder's frame layout.

Content (per the benchmark configurations): a background that is the sum of
`n_sines` random-phase 2-D sinusoids (periods 8-64 px, random orientation)
translated by a global pan per frame, `n_objects` opaque rectangles
(16-96 px) moving at velocities that are multiples of 1/8 px (|v| <= max_vel
px/frame per axis, wrapping around the frame), and additive Gaussian noise,
clipped to [0, 255].

Chroma (the configurations leave it open): each chroma plane has its own
set of 3 sinusoids at half resolution (amplitude about +-30 around 128),
panned by half the luma pan; each rectangle carries its own (U, V) colour;
chroma noise has half the luma sigma. Rectangle edges are anti-aliased by
exact box coverage, so 1/8-pel motion is visible as sub-pel motion.

`scene_cut` re-seeds the background (new sinusoids) from that frame on.

Determinism: numpy's PCG64 generator drives every random draw (integer
draws only), and every sum is integer fixed-point arithmetic over a sine
table whose entries are provably the same on any host (see _sine_table), so
a seed gives the same bytes everywhere. tests/golden/synth_sha256.json holds
the SHA-256 of the benchmark configurations' frames; the CPU tests and the
benchmark check them. Noise is the Irwin-Hall approximation of a Gaussian
(sum of 12 uniforms).

Frame layout (`pad_frame`): the encoder's padded planar buffer, luma stride
sb_cols * 64 and height sb_rows * 64, chroma stride sb_cols * 32 and height
sb_rows * 32, right and bottom edges replicated (src/encoder.c:93-120,
297-310).
"""
import math

import numpy as np


class Geometry:
    """Frame geometry of the encoder (src/encoder.c:93-120)."""

    def __init__(self, width, height, tile_cols_log2=2):
        if width < 16 or height < 16 or width > 65536 or height > 65536:
            raise ValueError(f"unsupported frame size {width}x{height}")
        self.w, self.h = width, height
        self.mi_cols = 2 * ((width + 7) >> 3)
        self.mi_rows = 2 * ((height + 7) >> 3)
        self.sb_cols = (self.mi_cols + 15) >> 4
        self.sb_rows = (self.mi_rows + 15) >> 4
        self.num_sbs = self.sb_cols * self.sb_rows
        k = 0
        while ((4096 >> 6) << k) < self.sb_cols:
            k += 1
        lo = k
        max_cols = min(self.sb_cols, 64)
        k = 0
        while (1 << k) < max_cols:
            k += 1
        hi = k
        l = min(max(tile_cols_log2, lo), hi)
        self.tile_cols_log2 = l
        self.tile_w_sb = (self.sb_cols + (1 << l) - 1) >> l
        self.tile_cols = (self.sb_cols + self.tile_w_sb - 1) // self.tile_w_sb
        self.stride = [self.sb_cols * 64, self.sb_cols * 32, self.sb_cols * 32]
        self.plane_h = [self.sb_rows * 64, self.sb_rows * 32, self.sb_rows * 32]
        self.offset = [0, self.stride[0] * self.plane_h[0], 0]
        self.offset[2] = self.offset[1] + self.stride[1] * self.plane_h[1]
        self.frame_bytes = self.offset[2] + self.stride[2] * self.plane_h[2]
        self.cw, self.ch = (width + 1) >> 1, (height + 1) >> 1


def pad_frame(geo, planes):
    """I420 planes (Y h x w, U and V ch x cw, uint8) -> padded frame bytes."""
    out = np.empty(geo.frame_bytes, np.uint8)
    dims = [(geo.h, geo.w), (geo.ch, geo.cw), (geo.ch, geo.cw)]
    for p in range(3):
        ph, pw = dims[p]
        src = np.asarray(planes[p], np.uint8).reshape(ph, pw)
        aw, ah = geo.stride[p], geo.plane_h[p]
        rows = np.minimum(np.arange(ah), ph - 1)
        cols = np.minimum(np.arange(aw), pw - 1)
        out[geo.offset[p]:geo.offset[p] + aw * ah] = src[rows][:, cols].reshape(-1)
    return out


def unpad_frame(geo, buf):
    """Padded frame bytes -> (Y, U, V) visible planes."""
    buf = np.asarray(buf, np.uint8)
    dims = [(geo.h, geo.w), (geo.ch, geo.cw), (geo.ch, geo.cw)]
    res = []
    for p in range(3):
        ph, pw = dims[p]
        pl = buf[geo.offset[p]:geo.offset[p] + geo.stride[p] * geo.plane_h[p]].reshape(geo.plane_h[p], geo.stride[p])
        res.append(pl[:ph, :pw].copy())
    return res


# Fixed-point arithmetic throughout, so every host generates the same bytes.
# Phases are in units of 2^-24 turn; the sine table has 4096 entries in Q14.
_PH_BITS = 24
_LUT_BITS = 12


def _sine_table():
    """round(16384 * sin(2 pi k / 4096)). Every entry lies at least 1e-6 away
    from a rounding boundary (checked), so any libm whose sin is accurate to
    far better than that gives the same table."""
    vals = []
    for k in range(1 << _LUT_BITS):
        x = 16384.0 * math.sin(2.0 * math.pi * k / (1 << _LUT_BITS))
        frac = abs(x - math.floor(x) - 0.5)
        if frac < 1e-6:
            raise AssertionError("sine table entry too close to a rounding boundary")
        vals.append(int(math.floor(x + 0.5)))
    return np.asarray(vals, np.int64)


_SIN = _sine_table()


def _sines(rng, n, amp_lo, amp_hi):
    """n sinusoids: integer wave vector (turns per pixel in Q24), phase (Q24)
    and amplitude (Q8). Periods 8-64 px (Q8), direction one of 4096 angles
    over a half turn."""
    s = []
    for _ in range(n):
        period_q8 = int(rng.integers(8 * 256, 64 * 256 + 1))
        t = int(rng.integers(0, 1 << (_LUT_BITS - 1)))  # angle t / 4096 turn, in [0, 1/2)
        c, sn = int(_SIN[(t + 1024) & 4095]), int(_SIN[t])  # cos, sin in Q14
        kx = (c << (_PH_BITS + 8 - 14)) // period_q8
        ky = (sn << (_PH_BITS + 8 - 14)) // period_q8
        phase = int(rng.integers(0, 1 << _PH_BITS))
        amp_q8 = int(rng.integers(int(amp_lo * 256), int(amp_hi * 256) + 1))
        s.append((kx, ky, phase, amp_q8))
    return s


def _field(sines, h, w, pan8, t, scale, base):
    """base + sum of the sinusoids in Q8 at pixel (x, y) shifted by the pan
    (pan8 in 1/8 px per frame at luma scale; scale 1 luma, 2 chroma)."""
    q = 3 + (scale - 1)  # positions in 1/8 (luma) or 1/16 (chroma) px
    xx = (np.arange(w, dtype=np.int64) << q) - pan8[0] * t
    yy = (np.arange(h, dtype=np.int64) << q) - pan8[1] * t
    v = np.full((h, w), base << 8, np.int32)
    m = (1 << (_PH_BITS + q)) - 1
    for kx, ky, ph, a in sines:
        # phase = floor((x kx + y ky) / 2^q) + ph modulo 2^24: the row and
        # column terms reduced modulo 2^(24+q) (which commutes with the
        # floor), the constant folded into the row term, int32 throughout
        px = ((xx * kx) & m).astype(np.int32)
        py = ((yy * ky + (ph << q)) & m).astype(np.int32)
        idx = ((py[:, None] + px[None, :]) >> (q + _PH_BITS - _LUT_BITS)) & ((1 << _LUT_BITS) - 1)
        v += ((a * _SIN) >> 14).astype(np.int32)[idx]
    return v.astype(np.int64)


def _coverage(pos, size, n, unit):
    """Coverage (0..unit) of [pos, pos + size) on pixels [j, j + 1), all in
    1/unit px."""
    j = np.arange(n, dtype=np.int64) * unit
    return np.clip(np.minimum(j + unit, pos + size) - np.maximum(j, pos), 0, unit)


def _noise(rng, shape, sigma_q8):
    """Gaussian noise in Q8 by the Irwin-Hall sum of 12 uniform integers in
    [0, 4096) (mean 12 * 2047.5, standard deviation 4096): sample i takes
    the low 12 bits of each 16-bit field of the raw 64-bit PCG64 outputs i,
    n + i and 2n + i (summed lane-wise, then across the four lanes)."""
    n = int(np.prod(shape))
    raw = rng.bit_generator.random_raw(3 * n).reshape(3, n)
    mask = np.uint64(0x0FFF0FFF0FFF0FFF)
    lanes = (raw[0] & mask) + (raw[1] & mask) + (raw[2] & mask)  # each 16-bit lane < 3 * 4096
    m16 = np.uint64(0xFFFF)
    u = (lanes & m16) + ((lanes >> np.uint64(16)) & m16) + ((lanes >> np.uint64(32)) & m16) + (lanes >> np.uint64(48))
    u = u.astype(np.int64) * 2 - 12 * 4095
    return ((u * sigma_q8) // 8192).reshape(shape)


class Synth:
    def __init__(self, width, height, seed, n_objects=4, max_vel=8, noise_sigma=2.0, pan=(1.25, -0.5),
                 n_sines=6, scene_cut=None):
        self.w, self.h = width, height
        self.cw, self.ch = (width + 1) >> 1, (height + 1) >> 1
        self.seed = seed
        self.pan8 = (int(round(pan[0] * 8)), int(round(pan[1] * 8)))
        self.sigma_q8 = int(round(noise_sigma * 256))
        self.scene_cut = scene_cut
        rng = np.random.default_rng(seed)
        self.bg = self._background(rng, n_sines)
        self.bg2 = self._background(np.random.default_rng(seed ^ 0x5EC7), n_sines) if scene_cut else None
        self.objs = []
        for _ in range(n_objects):
            ow, oh = int(rng.integers(16, 97)), int(rng.integers(16, 97))
            x0, y0 = int(rng.integers(0, width)), int(rng.integers(0, height))
            lim = 8 * max_vel
            vx8, vy8 = int(rng.integers(-lim, lim + 1)), int(rng.integers(-lim, lim + 1))
            col = (int(rng.integers(20, 236)), int(rng.integers(40, 217)), int(rng.integers(40, 217)))
            self.objs.append((ow, oh, x0, y0, vx8, vy8, col))
        self.noise_seed = seed * 7919 + 17

    @staticmethod
    def _background(rng, n_sines):
        return (_sines(rng, n_sines, 8.0, 22.0), _sines(rng, 3, 6.0, 12.0), _sines(rng, 3, 6.0, 12.0))

    def frame(self, t):
        """Frame t as (Y, U, V) uint8 arrays."""
        bg = self.bg2 if (self.scene_cut is not None and t >= self.scene_cut) else self.bg
        y = _field(bg[0], self.h, self.w, self.pan8, t, 1, 128)
        u = _field(bg[1], self.ch, self.cw, self.pan8, t, 2, 128)
        v = _field(bg[2], self.ch, self.cw, self.pan8, t, 2, 128)
        for ow, oh, x0, y0, vx8, vy8, col in self.objs:
            # top-left corner in 1/8 px, wrapping around the frame
            xs = (8 * x0 + vx8 * t) % (8 * (self.w + ow)) - 8 * ow
            ys = (8 * y0 + vy8 * t) % (8 * (self.h + oh)) - 8 * oh
            cov = _coverage(ys, 8 * oh, self.h, 8)[:, None] * _coverage(xs, 8 * ow, self.w, 8)[None, :]  # 0..64
            y = (y * (64 - cov) + (col[0] << 8) * cov + 32) >> 6
            # chroma: the same rectangle at half scale, in 1/16 chroma px
            ccov = _coverage(ys, 8 * oh, self.ch, 16)[:, None] * _coverage(xs, 8 * ow, self.cw, 16)[None, :]
            u = (u * (256 - ccov) + (col[1] << 8) * ccov + 128) >> 8
            v = (v * (256 - ccov) + (col[2] << 8) * ccov + 128) >> 8
        rng = np.random.default_rng(self.noise_seed + t)
        y = y + _noise(rng, y.shape, self.sigma_q8)
        u = u + _noise(rng, u.shape, self.sigma_q8 // 2)
        v = v + _noise(rng, v.shape, self.sigma_q8 // 2)
        q = lambda a: np.clip((a + 128) >> 8, 0, 255).astype(np.uint8)
        return q(y), q(u), q(v)


# The benchmark configurations' generators.
def config_synth(cfg, seed):
    if cfg == 1:
        return Synth(352, 288, seed, n_objects=4, max_vel=8, noise_sigma=2.0)
    if cfg == 2:
        return Synth(1920, 1080, seed, n_objects=12, max_vel=48, noise_sigma=3.0)
    if cfg == 3:
        return Synth(1920, 1080, seed, n_objects=16, max_vel=48, noise_sigma=3.0, scene_cut=32)
    raise ValueError(cfg)


# The batched workload: 16 streams of uneven sizes, each from its own seed
# (the generator of the 352x288 stream: 6 sinusoids, 4 rectangles moving up
# to +-8 px/frame over the global pan, noise sigma 2).
BATCH_STREAMS = [(1920, 1080)] * 4 + [(1280, 720)] * 4 + [(854, 480)] * 8


def batch_synth(i, seed_base=400):
    w, h = BATCH_STREAMS[i]
    return Synth(w, h, seed_base + i, n_objects=4, max_vel=8, noise_sigma=2.0)

