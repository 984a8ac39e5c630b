"""Writes tests/golden/synth_sha256.json: the SHA-256 of every frame (Y, U
and V planes concatenated) the synthetic generator makes for the benchmark
configurations, with the seeds the benchmark and tests use. The generator
is integer fixed-point arithmetic, so these hashes hold on every host;
tests/test_synth.py and bench.py check frames against them.

    python tests/golden/make_synth_sha256.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "mav1-encoder_amd"))

import synth  # noqa: E402

# (configuration, seed, frames): config 1 (352x288, 10 frames), config 2
# (1080p, 50 frames, the benchmark's stream), config 3 (1080p, 64 frames with
# the scene cut at 32)
STREAMS = [(1, 1, 10), (2, 2024, 50), (3, 3, 64)]


def frame_sha256(planes):
    h = hashlib.sha256()
    for p in planes:
        h.update(p.tobytes())
    return h.hexdigest()


def main():
    out = {}
    for cfg, seed, n in STREAMS:
        g = synth.config_synth(cfg, seed)
        out[f"config{cfg}_seed{seed}"] = {"width": g.w, "height": g.h, "frames": [frame_sha256(g.frame(t))
                                                                                   for t in range(n)]}
        print(f"config {cfg}: {n} frames", flush=True)
    with open(os.path.join(HERE, "synth_sha256.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
