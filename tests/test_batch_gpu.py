"""Batched streams (mav1hip_batch_*): several sequences of different sizes
coded together, every stage one launch over the round's frames. Each
stream's coded frames must equal what its own sequence driver codes from the
same frames (every field, bit-exact), and the small full-RD case is also
checked against the oracle's serial driver directly."""
import os
import sys

import numpy as np
import pytest

import mav1hip
import synth

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_seq_gpu import compare, run_gpu  # noqa: E402

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import seq as oracle_seq  # noqa: E402

pytestmark = pytest.mark.gpu


def run_batch(cfgs, planes, lr_test=0, device_frames=None):
    """planes[i]: stream i's frames; streams fed in step (frame t to every
    stream that has one), then ended; returns the coded frames per stream."""
    b = mav1hip.Batch(cfgs, lr_test=lr_test, dump_decisions=True)
    out = [[] for _ in cfgs]
    try:
        T = max(len(p) for p in planes)
        for t in range(T):
            for i, pl in enumerate(planes):
                if t >= len(pl):
                    continue
                while True:
                    ok = (b.send_device(i, device_frames[i][t].data_ptr()) if device_frames is not None
                          else b.send(i, pl[t]))
                    if ok:
                        break
                    out[i].append(b.receive(i, wait=True))
            for i, pl in enumerate(planes):
                if t == len(pl) - 1:
                    while not b.send(i, None):
                        out[i].append(b.receive(i, wait=True))
        for i in range(len(cfgs)):
            while True:
                r = b.receive(i, wait=True)
                if r is None:
                    break
                out[i].append(r)
    finally:
        b.close()
    return out


def check_against_sequences(cfgs, planes, lr_test=0, device=False):
    dev = None
    if device:
        import torch
        dev = []
        for cfg, pl in zip(cfgs, planes):
            geo = synth.Geometry(cfg.width, cfg.height, cfg.tile_cols_log2)
            dev.append([torch.from_numpy(synth.pad_frame(geo, p)).to("cuda:0") for p in pl])
        torch.cuda.synchronize()
    got = run_batch(cfgs, planes, lr_test, dev)
    for i, (cfg, pl) in enumerate(zip(cfgs, planes)):
        ref = run_gpu(cfg, pl, lr_test, dev[i] if dev is not None else None)
        assert len(got[i]) == len(ref) == len(pl), f"stream {i}: {len(got[i])} coded frames, expected {len(pl)}"
        for g, r in zip(got[i], ref):
            try:
                compare(g, r)
            except AssertionError as e:
                raise AssertionError(f"stream {i} ({cfg.width}x{cfg.height}): {e}") from None
        print(f"stream {i} {cfg.width}x{cfg.height}: {len(ref)} frames equal", flush=True)


def test_batch_uneven_sizes_turbo_restoration(gpu):
    """1080p, 720p and two 480p streams (different content), turbo q125 with
    restoration: key frame, ALTREF / leaf mini-GOPs, a closing regular frame."""
    print("\n[batch] 1920x1080 + 1280x720 + 2 x 854x480 turbo q125 restoration, 6 frames", flush=True)
    sizes = [(1920, 1080, 1), (1280, 720, 2), (854, 480, 3), (854, 480, 4)]
    cfgs = [mav1hip.default_config(w, h, turbo=1, qindex=125, restoration=1, keyint=60) for w, h, _ in sizes]
    planes = [[synth.Synth(w, h, seed, n_objects=12, max_vel=48, noise_sigma=3.0).frame(t) for t in range(6)]
              for w, h, seed in sizes]
    check_against_sequences(cfgs, planes)


def test_batch_streams_of_different_lengths(gpu):
    """Streams ending at different frames: partial rounds, a shorter stream's
    last mini-GOP and the other streams' rounds without it; lr_test 1."""
    print("\n[batch] 352x288 (9 frames) + 204x116 (4) + 200x136 (7), turbo restoration lr_test 1", flush=True)
    sizes = [(352, 288, 9, 11), (204, 116, 4, 12), (200, 136, 7, 13)]
    cfgs = [mav1hip.default_config(w, h, turbo=1, qindex=90, restoration=1, keyint=0) for w, h, _, _ in sizes]
    planes = [[synth.Synth(w, h, seed, n_objects=4, max_vel=8, noise_sigma=2.0).frame(t) for t in range(n)]
              for w, h, n, seed in sizes]
    check_against_sequences(cfgs, planes, lr_test=1)


def test_batch_device_frames(gpu):
    """Frames sent from device memory."""
    print("\n[batch] 2 x 640x360 turbo, device frames", flush=True)
    cfgs = [mav1hip.default_config(640, 360, turbo=1, qindex=170, restoration=1, keyint=0) for _ in range(2)]
    planes = [[synth.Synth(640, 360, 40 + i, n_objects=6, max_vel=16, noise_sigma=3.0).frame(t) for t in range(5)]
              for i in range(2)]
    check_against_sequences(cfgs, planes, device=True)


def test_batch_full_rd_pyramid_vs_oracle(gpu):
    """Full RD, altref 8 with leaf pyramids and a key frame every 6 frames,
    two tile columns, two streams of different sizes: each stream's coded
    frames against the oracle's serial driver."""
    print("\n[batch] 352x288 + 200x136 full RD pyramid altref 8 keyint 6", flush=True)
    sizes = [(352, 288, 21), (200, 136, 22)]
    cfgs = [mav1hip.default_config(w, h, turbo=0, qindex=125, altref=8, leaf=1, keyint=6, tile_cols_log2=1)
            for w, h, _ in sizes]
    planes = [[synth.Synth(w, h, seed, n_objects=6, max_vel=12, noise_sigma=2.0).frame(t) for t in range(11)]
              for w, h, seed in sizes]
    got = run_batch(cfgs, planes)
    for i, (cfg, pl) in enumerate(zip(cfgs, planes)):
        geo = synth.Geometry(cfg.width, cfg.height, cfg.tile_cols_log2)
        orc = oracle_seq.OracleSequence(mav1hip, cfg, 0)
        for p in pl:
            orc.send(synth.pad_frame(geo, p))
        orc.end()
        assert len(got[i]) == len(orc.coded) == len(pl)
        assert 3 in [c["frame_kind"] for c in orc.coded]
        for g, o in zip(got[i], orc.coded):
            compare(g, o)
        print(f"stream {i}: {len(pl)} frames equal to the oracle", flush=True)


def test_batch_rejects_too_many_streams(gpu):
    cfg = mav1hip.default_config(204, 116, turbo=1, qindex=90, keyint=0)
    with pytest.raises(mav1hip.HipError, match="1 to 32 streams"):
        mav1hip.Batch([cfg] * 33)


def luma_psnr(cfg, planes, coded):
    geo = synth.Geometry(cfg.width, cfg.height, cfg.tile_cols_log2)
    out = []
    for c in coded:
        y = synth.unpad_frame(geo, c["recon"])[0]
        out.append(mav1hip.psnr(y, planes[c["disp"]][0]))
    return out


def test_batch_pruned_analysis(gpu):
    """The learned pruner in the batched analysis (trained weights, confidence
    0.9): every stream codes, the coding is repeatable, and its luma PSNR stays
    within 0.5 dB of the exhaustive search's on every frame; the pruner off
    again gives the exhaustive search's frames exactly."""
    import train_pruner
    w = train_pruner.load_weights()
    assert w is not None and w.size == mav1hip.lib().mav1hip_pruner_num_params()
    print("\n[batch] 2 x 854x480 + 640x360 turbo q125, pruner at tau 0.1", flush=True)
    sizes = [(854, 480, 31), (854, 480, 32), (640, 360, 33)]
    cfgs = [mav1hip.default_config(sw, sh, turbo=1, qindex=125, keyint=0) for sw, sh, _ in sizes]
    planes = [[synth.Synth(sw, sh, seed, n_objects=6, max_vel=12, noise_sigma=2.5).frame(t) for t in range(6)]
              for sw, sh, seed in sizes]

    def run(tau):
        b = mav1hip.Batch(cfgs, dump_decisions=True)
        b.set_pruner(w, tau)
        out = [[] for _ in cfgs]
        try:
            for t in range(6):
                for i in range(len(cfgs)):
                    while not b.send(i, planes[i][t]):
                        out[i].append(b.receive(i, True))
            for i in range(len(cfgs)):
                while not b.send(i, None):
                    out[i].append(b.receive(i, True))
            for i in range(len(cfgs)):
                while True:
                    r = b.receive(i, True)
                    if r is None:
                        break
                    out[i].append(r)
        finally:
            b.close()
        return out

    plain = run_batch(cfgs, planes)
    a, bb, off = run(0.1), run(0.1), run(0.0)
    changed = 0
    for i, cfg in enumerate(cfgs):
        assert len(a[i]) == 6
        for x, y in zip(a[i], bb[i]):
            compare(x, y)
        for x, y in zip(off[i], plain[i]):
            compare(x, y)
        pp, pa = luma_psnr(cfg, planes[i], plain[i]), luma_psnr(cfg, planes[i], a[i])
        changed += sum(not np.array_equal(x["blocks"], y["blocks"]) for x, y in zip(a[i], plain[i]))
        print(f"stream {i}: luma PSNR exhaustive {np.mean(pp):.3f} dB, pruned {np.mean(pa):.3f} dB", flush=True)
        assert min(p - q for p, q in zip(pa, pp)) > -0.5
    print(f"{changed} frames with other decisions", flush=True)
