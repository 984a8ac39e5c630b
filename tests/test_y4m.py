"""The y4m reader and PSNR of the front end (tools/mav1enc.c:22-34, 67-77)
as library calls: frames read back equal the frames written, headers other
than 4:2:0 and truncated frames are refused, and the PSNR equals its closed
form."""
import math

import numpy as np
import pytest

import mav1hip
import synth


def _write_y4m(path, frames, w, h, header_extra="C420jpeg"):
    with open(path, "wb") as f:
        f.write(f"YUV4MPEG2 W{w} H{h} F25:1 Ip A1:1 {header_extra}\n".encode())
        for y, u, v in frames:
            f.write(b"FRAME\n")
            for p in (y, u, v):
                f.write(np.ascontiguousarray(p).tobytes())


def test_roundtrip_odd_size(tmp_path):
    w, h = 37, 21
    g = synth.Synth(w, h, 4, n_objects=2)
    frames = [g.frame(t) for t in range(3)]
    p = str(tmp_path / "a.y4m")
    _write_y4m(p, frames, w, h)
    r = mav1hip.Y4mReader(p)
    assert (r.width, r.height, r.fps) == (w, h, (25, 1))
    for fr in frames:
        got = r.read()
        assert all(np.array_equal(a, b) for a, b in zip(got, fr))
    assert r.read() is None
    r.close()


def test_refuses_non_420_and_truncated(tmp_path):
    p = str(tmp_path / "b.y4m")
    _write_y4m(p, [], 32, 32, header_extra="C444")
    with pytest.raises(mav1hip.HipError, match="4:2:0"):
        mav1hip.Y4mReader(p)
    g = synth.Synth(32, 32, 1, n_objects=1)
    _write_y4m(p, [g.frame(0)], 32, 32)
    with open(p, "r+b") as f:
        f.truncate(f.seek(0, 2) - 10)
    r = mav1hip.Y4mReader(p)
    with pytest.raises(mav1hip.HipError):
        r.read()
    r.close()
    with open(p, "wb") as f:
        f.write(b"P6\n")
    with pytest.raises(mav1hip.HipError, match="not y4m"):
        mav1hip.Y4mReader(p)


def test_psnr_closed_form():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, 10007, dtype=np.uint8)
    b = a.copy()
    assert mav1hip.psnr(a, b) == 99.0
    b[::3] = np.clip(b[::3].astype(int) + rng.integers(-9, 10, b[::3].size), 0, 255)
    sse = float(((a.astype(np.int64) - b) ** 2).sum())
    assert mav1hip.psnr(a, b) == pytest.approx(10 * math.log10(255.0 ** 2 * a.size / sse), rel=1e-12)
