"""The synthetic generator makes the same bytes as when the committed
hashes were written (tests/golden/synth_sha256.json), so every host codes
the same frames; its content matches the configurations' description."""
import hashlib
import json
import os

import numpy as np

import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "synth_sha256.json")


def _sha(planes):
    h = hashlib.sha256()
    for p in planes:
        h.update(p.tobytes())
    return h.hexdigest()


def test_frames_match_committed_hashes():
    ref = json.load(open(GOLDEN))
    for cfg, seed, frames in [(1, 1, [0, 1, 9]), (2, 2024, [0, 49]), (3, 3, [31, 32])]:
        g = synth.config_synth(cfg, seed)
        want = ref[f"config{cfg}_seed{seed}"]
        assert (want["width"], want["height"]) == (g.w, g.h)
        for t in frames:
            assert _sha(g.frame(t)) == want["frames"][t], f"config {cfg} frame {t} differs from its committed hash"


def test_sine_table():
    s = synth._SIN
    assert len(s) == 4096 and s[0] == 0 and s[1024] == 16384 and s[3072] == -16384
    # odd symmetry and a half-turn sign flip, as sin has
    assert np.array_equal(s[1:2048], s[4095:2048:-1] * -1)
    assert np.array_equal(s[:2048], -s[2048:])


def test_noise_statistics():
    rng = np.random.default_rng(3)
    n = synth._noise(rng, (512, 512), 512)  # sigma 2 in Q8
    assert abs(n.mean() / 256) < 0.02
    assert abs(n.std() / 256 - 2.0) < 0.02


def test_content_moves_as_described():
    """Rectangles of constant colour move by their velocity: frame t+1
    shifted back by the pan reproduces the background away from the
    rectangles (no noise, one sinusoid)."""
    g = synth.Synth(128, 96, 11, n_objects=0, noise_sigma=0.0, pan=(2.0, -1.0), n_sines=1)
    y0 = g.frame(0)[0].astype(int)
    y1 = g.frame(1)[0].astype(int)
    # integer pan (2, -1): y1[y, x] == y0[y + 1, x - 2]
    assert np.array_equal(y1[:-1, 2:], y0[1:, :-2])
