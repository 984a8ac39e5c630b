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

Determinism: numpy's PCG64 generator drives every random draw. The sums use
float64 sin/cos, whose last-ulp results can differ between CPUs; tests that
need fixed bytes store the frames they use instead of regenerating them.

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


def _sines(rng, n, amp_lo, amp_hi):
    s = []
    for _ in range(n):
        period = rng.uniform(8.0, 64.0)
        theta = rng.uniform(0.0, math.pi)
        phase = rng.uniform(0.0, 2 * math.pi)
        amp = rng.uniform(amp_lo, amp_hi)
        s.append((2 * math.pi / period * math.cos(theta), 2 * math.pi / period * math.sin(theta), phase, amp))
    return s


def _field(sines, h, w, ox, oy, base):
    yy = np.arange(h, dtype=np.float64)[:, None] - oy
    xx = np.arange(w, dtype=np.float64)[None, :] - ox
    v = np.full((h, w), base, np.float64)
    for kx, ky, ph, a in sines:
        v += a * np.sin(xx * kx + yy * ky + ph)
    return v


def _coverage(pos, size, n):
    """Per-pixel coverage of [pos, pos + size) on pixels [j, j + 1)."""
    j = np.arange(n, dtype=np.float64)
    return np.clip(np.minimum(j + 1, pos + size) - np.maximum(j, pos), 0.0, 1.0)


class Synth:
    def __init__(self, width, height, seed, n_objects=4, max_vel=8, noise_sigma=2.0, pan=(1.25, -0.5),
                 n_sines=6, scene_cut=None):
        self.w, self.h = width, height
        self.cw, self.ch = (width + 1) >> 1, (height + 1) >> 1
        self.seed = seed
        self.pan = pan
        self.sigma = noise_sigma
        self.scene_cut = scene_cut
        rng = np.random.default_rng(seed)
        self.bg = self._background(rng, n_sines)
        self.bg2 = self._background(np.random.default_rng(seed ^ 0x5EC7), n_sines) if scene_cut else None
        self.objs = []
        for _ in range(n_objects):
            ow, oh = int(rng.integers(16, 97)), int(rng.integers(16, 97))
            x0, y0 = int(rng.integers(0, width)), int(rng.integers(0, height))
            lim = 8 * max_vel
            vx, vy = int(rng.integers(-lim, lim + 1)) / 8.0, int(rng.integers(-lim, lim + 1)) / 8.0
            col = (float(rng.integers(20, 236)), float(rng.integers(40, 217)), float(rng.integers(40, 217)))
            self.objs.append((ow, oh, x0, y0, vx, vy, col))
        self.noise_seed = seed * 7919 + 17

    @staticmethod
    def _background(rng, n_sines):
        return (_sines(rng, n_sines, 8.0, 22.0), _sines(rng, 3, 6.0, 12.0), _sines(rng, 3, 6.0, 12.0))

    def frame(self, t):
        """Frame t as (Y, U, V) uint8 arrays."""
        bg = self.bg2 if (self.scene_cut is not None and t >= self.scene_cut) else self.bg
        px, py = self.pan[0] * t, self.pan[1] * t
        y = _field(bg[0], self.h, self.w, px, py, 128.0)
        u = _field(bg[1], self.ch, self.cw, px / 2, py / 2, 128.0)
        v = _field(bg[2], self.ch, self.cw, px / 2, py / 2, 128.0)
        for ow, oh, x0, y0, vx, vy, col in self.objs:
            xs = (x0 + vx * t) % (self.w + ow) - ow
            ys = (y0 + vy * t) % (self.h + oh) - oh
            cov = _coverage(ys, oh, self.h)[:, None] * _coverage(xs, ow, self.w)[None, :]
            y = y * (1 - cov) + col[0] * cov
            ccov = _coverage(ys / 2, oh / 2, self.ch)[:, None] * _coverage(xs / 2, ow / 2, self.cw)[None, :]
            u = u * (1 - ccov) + col[1] * ccov
            v = v * (1 - ccov) + col[2] * ccov
        rng = np.random.default_rng(self.noise_seed + t)
        y = y + rng.normal(0.0, self.sigma, y.shape)
        u = u + rng.normal(0.0, self.sigma / 2, u.shape)
        v = v + rng.normal(0.0, self.sigma / 2, v.shape)
        q = lambda a: np.clip(np.rint(a), 0, 255).astype(np.uint8)
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
