"""The sequence driver (mav1hip_seq_*) in closed loop against the oracle's
serial driver (oracle/seq.py): every coded frame's state, parameters,
block records, levels, CDEF choices, restoration units, final picture and
decision record are bit-exact, with each frame predicting from the
pictures the driver itself reconstructed. The 1080p case is the turbo
stream with restoration on, on content with fast motion (12 objects up to
+-48 px/frame, noise sigma 3); the small cases cover the full-RD pyramid
(hidden ALTREF, pyramid middle frames, leaf frames), a key frame inside
the stream, two tile columns, a regular frame closing the stream, the
forced restoration alternation, and frames sent from device memory."""
import os
import sys

import numpy as np
import pytest

import mav1hip
import synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import seq as oracle_seq  # noqa: E402

pytestmark = pytest.mark.gpu

FIELDS = ["coded_index", "disp", "is_key", "frame_kind", "nonref", "hidden", "order_hint", "refresh_frame_flags",
          "last_slot", "arf_slot", "last_hint", "arf_hint", "engine"]


def run_gpu(cfg, planes, lr_test, device_frames=None):
    seq = mav1hip.Sequence(cfg, lr_test=lr_test, dump_decisions=True)
    out = []
    try:
        for i, pl in enumerate(planes):
            while True:
                ok = seq.send_device(device_frames[i].data_ptr()) if device_frames is not None else seq.send(pl)
                if ok:
                    break
                out.append(seq.receive(wait=True))
        while not seq.send(None):
            out.append(seq.receive(wait=True))
        while True:
            r = seq.receive(wait=True)
            if r is None:
                break
            out.append(r)
    finally:
        seq.close()
    return out


def compare(g, o):
    tag = f"coded frame {o['coded_index']} (display {o['disp']}, kind {o['frame_kind']})"
    for f in FIELDS:
        assert g[f] == o[f], f"{tag}: {f} {g[f]} != {o[f]}"
    assert bytes(g["fp"]) == bytes(o["fp"]), f"{tag}: frame parameters differ"
    assert bytes(g["plan"]) == bytes(o["plan"]), f"{tag}: filter plan differs"
    fp = o["fp"]
    assert np.array_equal(g["sbs"], o["sbs"]), f"{tag}: superblock records differ"
    for s in range(fp.sb_cols * fp.sb_rows):
        n = o["sbs"][s, 1]
        assert np.array_equal(g["blocks"][s * 64:s * 64 + n], o["blocks"][s * 64:s * 64 + n]), \
            f"{tag}: blocks of superblock {s} differ"
    bad = np.nonzero(g["coefs"] != o["coefs"])[0]
    assert bad.size == 0, f"{tag}: levels differ at {bad.size} positions, first {bad[:6]}"
    plan = o["plan"]
    if plan.cdef_select:
        assert np.array_equal(g["cdef_choice"], o["cdef_choice"]), f"{tag}: CDEF choices differ"
    if plan.restoration:
        bad = np.nonzero((g["lr_units"] != o["lr_units"]).any(axis=1))[0]
        assert bad.size == 0, f"{tag}: restoration units differ at {bad[:8]}"
    bad = np.nonzero(g["recon"] != o["recon"])[0]
    assert bad.size == 0, f"{tag}: final picture differs at {bad.size} bytes, first {bad[:6]}"
    assert np.array_equal(g["decision"], o["decision"]), f"{tag}: decision record differs"


def run_case(cfg, gen, nframes, lr_test, device=False):
    geo = synth.Geometry(cfg.width, cfg.height, cfg.tile_cols_log2)
    planes = [gen.frame(t) for t in range(nframes)]
    padded = [synth.pad_frame(geo, p) for p in planes]
    orc = oracle_seq.OracleSequence(mav1hip, cfg, lr_test)
    for f in padded:
        orc.send(f)
    orc.end()
    dev = None
    if device:
        import torch
        dev = [torch.from_numpy(f).to("cuda:0") for f in padded]
        torch.cuda.synchronize()
    got = run_gpu(cfg, planes, lr_test, dev)
    assert len(got) == len(orc.coded) == nframes
    kinds = [o["frame_kind"] + 10 * o["is_key"] for o in orc.coded]
    print(f"coded {nframes} frames, kinds {kinds}", flush=True)
    for g, o in zip(got, orc.coded):
        compare(g, o)
    return orc.coded


def test_seq_1080p_turbo_fast_motion_restoration(gpu):
    """Key frame, three ALTREF / leaf mini-GOPs and a regular frame closing
    the stream at 1920x1080, --turbo -q 125 with restoration on."""
    print("\n[seq] 1920x1080 turbo q125 restoration, config-2 content, 8 frames", flush=True)
    cfg = mav1hip.default_config(1920, 1080, turbo=1, qindex=125, restoration=1, keyint=50)
    coded = run_case(cfg, synth.config_synth(2, seed=2024), 8, 0)
    assert [c["frame_kind"] for c in coded] == [0, 1, 2, 1, 2, 1, 2, 0]
    assert sum(c["plan"].restoration for c in coded) >= 4


@pytest.mark.parametrize("lr_test", [0, 1])
def test_seq_full_rd_pyramid(gpu, lr_test):
    """Full RD, altref 8 with leaf pyramids, a key frame every 6 frames, two
    tile columns: the coding order holds pyramid middle frames whose slots
    the second engine writes and reads."""
    print(f"\n[seq] 352x288 full RD pyramid altref 8 keyint 6 lr_test {lr_test}", flush=True)
    cfg = mav1hip.default_config(352, 288, turbo=0, qindex=125, altref=8, leaf=1, keyint=6, tile_cols_log2=1)
    coded = run_case(cfg, synth.Synth(352, 288, 77 + lr_test, n_objects=6, max_vel=12, noise_sigma=2.0), 11, lr_test)
    assert 3 in [c["frame_kind"] for c in coded] and sum(c["is_key"] for c in coded) == 2


def test_seq_full_rd_1080p_frame_kinds(gpu):
    """The full-RD configuration at 1920x1080 (tx search, tx depth,
    switchable dual filter, 4 CDEF presets, restoration, two tile columns)
    over a key frame and a mini-GOP of 4 with leaf frames: the hidden
    ALTREF (kind 1), a pyramid middle frame (kind 3) and two leaf frames
    (kind 2), on content with 16 objects moving up to +-48 px/frame."""
    print("\n[seq] 1920x1080 full RD q125 altref 4 leaf tiles 2", flush=True)
    cfg = mav1hip.default_config(1920, 1080, turbo=0, qindex=125, altref=4, leaf=1, tile_cols_log2=1)
    coded = run_case(cfg, synth.config_synth(3, seed=3), 5, 0)
    assert sorted(c["frame_kind"] for c in coded) == [0, 1, 2, 2, 3]


@pytest.mark.parametrize("q", [80, 170, 215])
def test_seq_full_rd_720p_q_ladder(gpu, q):
    """The full-RD configuration at 1280x720 over the rest of the q ladder."""
    print(f"\n[seq] 1280x720 full RD q{q} altref 4 leaf tiles 2", flush=True)
    cfg = mav1hip.default_config(1280, 720, turbo=0, qindex=q, altref=4, leaf=1, tile_cols_log2=1)
    run_case(cfg, synth.Synth(1280, 720, 300 + q, n_objects=16, max_vel=48, noise_sigma=3.0), 5, 0)


def test_seq_full_rd_altref16_pyramid(gpu):
    """altref 16: a 16-frame pyramid (middle frames at depths 0-3) after the
    key frame, at 352x288."""
    print("\n[seq] 352x288 full RD altref 16", flush=True)
    cfg = mav1hip.default_config(352, 288, turbo=0, qindex=125, altref=16, leaf=1, tile_cols_log2=1)
    coded = run_case(cfg, synth.Synth(352, 288, 16, n_objects=8, max_vel=16, noise_sigma=3.0), 17, 0)
    assert sum(c["frame_kind"] == 3 for c in coded) == 7


def test_seq_device_frames_odd_size(gpu):
    """Frames sent from device memory, at a size whose width is not a
    multiple of 8 (the restored picture keeps the reconstruction outside
    the visible area)."""
    print("\n[seq] 204x116 turbo restoration, device frames", flush=True)
    cfg = mav1hip.default_config(204, 116, turbo=1, qindex=90, restoration=1, keyint=0)
    run_case(cfg, synth.Synth(204, 116, 5, n_objects=4, max_vel=6, noise_sigma=2.0), 7, 0, device=True)
