"""Motion search on the GPU against the oracle: quarter-pel / decimated
planes byte for byte, and the motion fields of every node of every level
for LAST and the second reference, bit-exact."""
import numpy as np
import pytest

import libs
import mav1hip

pytestmark = pytest.mark.gpu

CASES = [(352, 288, 11), (136, 72, 5), (854, 480, 4), (1920, 1080, 21)]


@pytest.mark.parametrize("w,h,seed", CASES)
@pytest.mark.parametrize("kind", ["arf", "leaf"])
def test_motion_search_matches_oracle(gpu, w, h, seed, kind):
    print(f"\n[me] {w}x{h} seed {seed} {kind}", flush=True)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=6, max_vel=12)
    cfg, fp = libs.turbo_params(w, h, kind)
    ctx = mav1hip.Context(cfg)
    mv, mvg = ctx.motion_search(fp, frames[1], frames[0], frames[2])
    mo, go = libs.Impl("oracle").motion_search(fp, frames[1], frames[0], frames[2])
    bad = np.nonzero(mv != mo)[0]
    assert bad.size == 0, f"LAST field differs at {bad[:8]} (gpu {mv[bad[:8]]}, oracle {mo[bad[:8]]})"
    if fp.golden:
        bad = np.nonzero(mvg != go)[0]
        assert bad.size == 0, f"second-reference field differs at {bad[:8]}"
    ctx.close()


def test_planes_match_oracle(gpu):
    w, h = 854, 480
    geo, frames = libs.make_sequence(w, h, 3, 2)
    cfg, fp = libs.turbo_params(w, h, "leaf")
    ctx = mav1hip.Context(cfg)
    ctx.motion_search(fp, frames[1], frames[0], frames[2])
    planes = ctx.download(mav1hip.BUF_PLANES)
    eb = mav1hip.lib().mav1hip_planes_entry_bytes(ctx.h)  # entry stride (256-byte aligned)
    n16 = 16 * w * h
    nd = ((w + 3) >> 2) * ((h + 3) >> 2)
    o = libs.Impl("oracle")
    for r, f in enumerate((frames[0], frames[2])):
        assert np.array_equal(planes[r * eb:r * eb + n16], o.qpel_build(fp, f)), f"quarter-pel planes of ref {r}"
        assert np.array_equal(planes[r * eb + n16:r * eb + n16 + nd], o.dec_build(fp, f)), f"decimated plane of ref {r}"
    assert np.array_equal(planes[2 * eb:2 * eb + nd], o.dec_build(fp, frames[1])), "decimated source"
    ctx.close()


@pytest.mark.parametrize("mask", [0x47CBF1 & ~0x20000, 0x47CBF1 & ~0x200 | 0x2000, 0x47CBF1 | 0x8 | 0x1000,
                                  0x47CBF1 & ~(0x40 | 0x8000 | 0x10000), 0x47CBF1 | 0x2 | 0x4])
def test_motion_search_turbo_bits(gpu, mask):
    w, h = 200, 136
    print(f"\n[me] mask {mask:#x}", flush=True)
    geo, frames = libs.make_sequence(w, h, 3, 9, n_objects=5, max_vel=10)
    cfg, fp = libs.turbo_params(w, h, "leaf", knob_turbo_mask=mask)
    ctx = mav1hip.Context(cfg)
    mv, mvg = ctx.motion_search(fp, frames[1], frames[0], frames[2])
    mo, go = libs.Impl("oracle").motion_search(fp, frames[1], frames[0], frames[2])
    assert np.array_equal(mv, mo)
    assert np.array_equal(mvg, go)
    ctx.close()


def test_motion_search_device_resident_pool(gpu):
    """The pool / bind form gives the same fields as the host-buffer call, and
    reuses state correctly over consecutive frames of different kinds."""
    w, h = 352, 288
    geo, frames = libs.make_sequence(w, h, 5, 13)
    cfg, fp_arf = libs.turbo_params(w, h, "arf", order_hint=2)
    _, fp_leaf = libs.turbo_params(w, h, "leaf", order_hint=1)
    ctx = mav1hip.Context(cfg)
    ctx.pool_alloc(len(frames))
    for i, f in enumerate(frames):
        ctx.pool_upload(i, f)
    o = libs.Impl("oracle")
    for (fp, s, r, g) in [(fp_arf, 2, 0, -1), (fp_leaf, 1, 0, 2), (fp_arf, 4, 2, -1), (fp_leaf, 3, 2, 4)]:
        ctx.bind_frames(s, r, g)
        ctx.set_params(fp)
        ctx.launch_motion_search()
        ctx.sync()
        mv = ctx.download(mav1hip.BUF_MVFIELD, np.int32)
        mo, go = o.motion_search(fp, frames[s], frames[r], frames[g] if g >= 0 else None)
        assert np.array_equal(mv, mo)
        if fp.golden:
            assert np.array_equal(ctx.download(mav1hip.BUF_MVGOLD, np.int32), go)
    ctx.close()


@pytest.mark.parametrize("variant", ["full_rd", "full_rd_352", "mask_no_hpel"])
def test_motion_search_without_planes(gpu, variant):
    """The non-turbo search and a turbo mask without the plane bits: sub-pel
    costs from EIGHTTAP block predictions."""
    if variant == "full_rd":
        w, h, seed, kind, kw = 136, 104, 17, "inter", {"turbo": 0, "multi_ref": 1}
    elif variant == "full_rd_352":
        w, h, seed, kind, kw = 352, 288, 4, "inter", {"turbo": 0, "multi_ref": 1}
    else:
        w, h, seed, kind, kw = 200, 136, 9, "leaf", {"knob_turbo_mask": 0x47CBF1 & ~0x30}
    print(f"\n[me] no planes {variant}", flush=True)
    turbo = kw.pop("turbo", 1)
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=125, **kw)
    fr = mav1hip.turbo_frame(kind, 1) if kind != "inter" else mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, _ = mav1hip.setup_params(cfg, fr)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=5, max_vel=10)
    ctx = mav1hip.Context(cfg)
    try:
        mv, mvg = ctx.motion_search(fp, frames[1], frames[0], frames[2])
    finally:
        ctx.close()
    mo, go = libs.Impl("oracle").motion_search(fp, frames[1], frames[0], frames[2])
    assert np.array_equal(mv, mo)
    if fp.golden:
        assert np.array_equal(mvg, go)
