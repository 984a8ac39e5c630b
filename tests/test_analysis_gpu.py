"""Analysis (pass 1 table + pass 2 partition / blocks) on the GPU against
the oracle, bit-exact: pass-1 costs of every launched item, the decision of
every evaluated item, every emitted block and every superblock record."""
import numpy as np
import pytest

import libs
import mav1hip

pytestmark = pytest.mark.gpu


def _params(w, h, kind, turbo=1, **kw):
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    if kind in ("key", "arf", "leaf"):
        fr = mav1hip.turbo_frame(kind, 1)
    else:
        fr = mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, _ = mav1hip.setup_params(cfg, fr)
    return cfg, fp


def launched_mask(fp):
    """Table items pass 1 evaluates (slots per level, as level_slots)."""
    sq = (fp.turbo & 1) and (not fp.frame_is_intra or (fp.turbo & 256))
    nsb = fp.sb_cols * fp.sb_rows
    m = np.zeros((nsb, 85, 13), bool)
    for level in range(4):
        ns = 1 if (level == 3 or sq) else (5 if (level == 2 or fp.fast) else 13)
        b = ((1 << (2 * level)) - 1) // 3
        m[:, b:b + (1 << (2 * level)), :ns] = True
    return m.reshape(-1)


def run_case(cfg, fp, frames):
    o = libs.Impl("oracle")
    mvf = mvg = None
    if not fp.frame_is_intra:
        mvf, mvg = o.motion_search(fp, frames[1], frames[0], frames[2])
    if mvf is None:
        mvf = np.zeros(fp.sb_cols * fp.sb_rows * 170, np.int32)
    if mvg is None:
        mvg = np.zeros_like(mvf)
    ctx = mav1hip.Context(cfg)
    g = ctx.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    ctx.close()
    a = o.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    mask = launched_mask(fp)
    gc, gi, gb, gs = g
    oc, oi, ob, os_ = a
    bad = np.nonzero((gc != oc) & mask)[0]
    assert bad.size == 0, f"pass-1 costs differ at items {bad[:6]}: gpu {gc[bad[:6]]} oracle {oc[bad[:6]]}"
    ev = (oc >= 0) & mask
    assert np.array_equal(gi[ev][:, libs.INFO_COLS], oi[ev][:, libs.INFO_COLS]), "pass-1 decisions differ"
    assert np.array_equal(gs, os_), "superblock records differ"
    for s in range(fp.sb_cols * fp.sb_rows):
        n = os_[s, 1]
        assert np.array_equal(gb[s * 64:s * 64 + n], ob[s * 64:s * 64 + n]), f"blocks of SB {s} differ"


@pytest.mark.parametrize("kind", ["key", "arf", "leaf"])
@pytest.mark.parametrize("w,h,seed", [(352, 288, 31), (200, 136, 8), (854, 480, 3), (1920, 1080, 4)])
def test_analysis_turbo(gpu, kind, w, h, seed):
    print(f"\n[analysis] {w}x{h} seed {seed} {kind}", flush=True)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=12)
    cfg, fp = _params(w, h, kind)
    run_case(cfg, fp, frames)


@pytest.mark.parametrize("variant", ["full_rd", "fast_multiref", "key_full", "q_low", "usub_off"])
def test_analysis_other_presets(gpu, variant):
    w, h = 136, 104
    print(f"\n[analysis] preset {variant}", flush=True)
    geo, frames = libs.make_sequence(w, h, 3, 17, n_objects=5, max_vel=6)
    if variant == "full_rd":
        cfg, fp = _params(w, h, "inter", turbo=0, multi_ref=1)
    elif variant == "fast_multiref":
        cfg, fp = _params(w, h, "inter", turbo=0, multi_ref=1, fast=1)
    elif variant == "key_full":
        cfg, fp = _params(w, h, "key", turbo=0)
    elif variant == "q_low":
        cfg, fp = _params(w, h, "leaf", q=20)
    else:
        cfg, fp = _params(w, h, "arf", knob_turbo_mask=0x47CBF1 & ~0x400000 & ~0x1)
    run_case(cfg, fp, frames)


def test_device_analysis_sees_upload_after_motion_search(gpu):
    """On the device path analysis overlaps the frame's motion search (a
    second stream) only when it directly follows it. A source upload enqueued
    in between must be visible to the analysis: it equals the oracle's on
    the new source with the motion field the GPU searched on the old one."""
    w, h = 352, 288
    cfg, fp = _params(w, h, "arf")
    geo, frames = libs.make_sequence(w, h, 3, 17, n_objects=8, max_vel=10)
    nsb = fp.sb_cols * fp.sb_rows
    ctx = mav1hip.Context(cfg)
    try:
        ctx.set_params(fp)
        ctx.upload(mav1hip.BUF_SRC, frames[1])
        ctx.upload(mav1hip.BUF_REF, frames[0])
        ctx.launch_motion_search()
        ctx.upload(mav1hip.BUF_SRC, frames[2])
        ctx.launch_analyze()
        ctx.sync()
        mvf = ctx.download(mav1hip.BUF_MVFIELD, np.int32, nsb * 170)
        blocks = ctx.download(mav1hip.BUF_BLOCKS, np.int32, nsb * 64 * 27).reshape(-1, 27)
        sbs = ctx.download(mav1hip.BUF_SBS, np.int32, nsb * 89).reshape(-1, 89)
    finally:
        ctx.close()
    o = libs.Impl("oracle")
    _, _, ob, os_ = o.analyze(fp, frames[2], frames[0], frames[2], mvf, np.zeros_like(mvf))
    assert np.array_equal(sbs, os_), "superblock records differ"
    for s in range(nsb):
        n = os_[s, 1]
        assert np.array_equal(blocks[s * 64:s * 64 + n], ob[s * 64:s * 64 + n]), f"blocks of SB {s} differ"
