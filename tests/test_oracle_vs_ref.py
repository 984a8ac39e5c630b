"""The oracle restatement against the reference's own kernel code
(oracle/_ref, built from the reference sources in this container): both run
on the same seeded synthetic frames and must agree bit for bit."""
import numpy as np
import pytest

import libs

pytestmark = pytest.mark.skipif(not libs.ref_available(), reason="oracle/_ref not built (needs the reference sources)")


@pytest.mark.parametrize("w,h,seed", [(352, 288, 1), (200, 136, 7), (64, 48, 3)])
def test_qpel_and_dec_planes(w, h, seed):
    geo, frames = libs.make_sequence(w, h, 2, seed)
    _, fp = libs.turbo_params(w, h, "arf")
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    assert np.array_equal(o.qpel_build(fp, frames[0]), r.qpel_build(fp, frames[0]))
    assert np.array_equal(o.dec_build(fp, frames[1]), r.dec_build(fp, frames[1]))


@pytest.mark.parametrize("kind", ["arf", "leaf"])
@pytest.mark.parametrize("w,h,seed", [(352, 288, 11), (136, 72, 5), (1920, 1080, 21)])
def test_motion_search_turbo(kind, w, h, seed):
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=6, max_vel=12)
    _, fp = libs.turbo_params(w, h, kind)
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    mo, go = o.motion_search(fp, frames[1], frames[0], frames[2])
    mr, gr = r.motion_search(fp, frames[1], frames[0], frames[2])
    assert np.array_equal(mo, mr)
    if fp.golden:
        assert np.array_equal(go, gr)
    assert np.any(mo != 0)  # the content moves: the search found vectors


@pytest.mark.parametrize("mask", [0x47CBF1 & ~0x20000, 0x47CBF1 & ~0x200 | 0x2000, 0x47CBF1 | 0x8 | 0x1000,
                                  0x47CBF1 & ~(0x40 | 0x8000 | 0x10000), 0x47CBF1 | 0x2 | 0x4])
def test_motion_search_turbo_bits(mask):
    """Other speed-bit combinations exercise the remaining branches
    (no HME3, CROSS, KEEP + L3SUB, CR 4 without HALF01 / RSUB, CR1 + L3COPY)."""
    w, h = 200, 136
    geo, frames = libs.make_sequence(w, h, 3, 9, n_objects=5, max_vel=10)
    _, fp = libs.turbo_params(w, h, "leaf", knob_turbo_mask=mask)
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    mo, go = o.motion_search(fp, frames[1], frames[0], frames[2])
    mr, gr = r.motion_search(fp, frames[1], frames[0], frames[2])
    assert np.array_equal(mo, mr)
    assert np.array_equal(go, gr)


def _frame_params(w, h, kind, turbo=1, **kw):
    import mav1hip
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    if kind in ("key", "arf", "leaf"):
        fr = mav1hip.turbo_frame(kind, 1)
    else:  # regular inter frame with a GOLDEN reference (multi_ref)
        fr = mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, _ = mav1hip.setup_params(cfg, fr)
    return fp


def _analysis_pair(fp, frames):
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    mvf = mvg = None
    if not fp.frame_is_intra:
        mvf, mvg = r.motion_search(fp, frames[1], frames[0], frames[2])
    if mvg is None:
        mvg = np.zeros_like(mvf) if mvf is not None else None
    a = o.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    b = r.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    return a, b


@pytest.mark.parametrize("kind", ["key", "arf", "leaf"])
@pytest.mark.parametrize("w,h,seed", [(352, 288, 31), (200, 136, 8), (1920, 1080, 4)])
def test_analysis_turbo(kind, w, h, seed):
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=12)
    fp = _frame_params(w, h, kind)
    a, b = _analysis_pair(fp, frames)
    libs.compare_analysis(a, b, geo.num_sbs)
    assert a[3][:, 1].sum() > geo.num_sbs  # some superblocks split


@pytest.mark.parametrize("variant", ["full_rd", "fast_multiref", "key_full", "q_low", "usub_off"])
def test_analysis_other_presets(variant):
    """Non-turbo analysis: all 13 slots with A/B and 4-way partitions, angle
    deltas, CfL; the fast preset's 5 slots; low q; turbo without unit
    subsampling."""
    w, h = 136, 104
    geo, frames = libs.make_sequence(w, h, 3, 17, n_objects=5, max_vel=6)
    if variant == "full_rd":
        fp = _frame_params(w, h, "inter", turbo=0, multi_ref=1)
    elif variant == "fast_multiref":
        fp = _frame_params(w, h, "inter", turbo=0, multi_ref=1, fast=1)
    elif variant == "key_full":
        fp = _frame_params(w, h, "key", turbo=0)
    elif variant == "q_low":
        fp = _frame_params(w, h, "leaf", q=20)
    else:
        fp = _frame_params(w, h, "arf", knob_turbo_mask=0x47CBF1 & ~0x400000 & ~0x1)
    a, b = _analysis_pair(fp, frames)
    libs.compare_analysis(a, b, geo.num_sbs)


def test_predict_intra_all_modes():
    """predict_intra with every mode, angle delta, size, edge filter setting
    and neighbour availability, on random pixels."""
    rng = np.random.default_rng(5)
    S = 320
    px = rng.integers(0, 256, (S, S), dtype=np.uint8)
    px[100:200, 100:200] = np.clip(px[100:200, 100:200] // 8 + 120, 0, 255)  # smooth area
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    sizes = [(2, 2), (3, 3), (4, 4), (5, 5), (6, 6), (2, 3), (3, 2), (3, 4), (4, 3), (4, 5), (5, 4), (5, 6), (6, 5),
             (2, 4), (4, 2), (3, 5), (5, 3), (4, 6), (6, 4)]
    n = 0
    for lw, lh in sizes:
        for mode in range(13):
            deltas = range(-3, 4) if 1 <= mode <= 8 else [0]
            for d in deltas:
                for ef in (0, 1):
                    ft = int(rng.integers(0, 2))
                    x, y = int(rng.integers(1, 3)) * 64 + int(rng.integers(0, 4)) * 8, int(rng.integers(1, 3)) * 64
                    hl, ha, har, hbl = (int(v) for v in rng.integers(0, 2, 4))
                    maxx = x + int(rng.integers(1, 2 << lw))
                    maxy = y + int(rng.integers(1, 2 << lh))
                    args = [hl, ha, har, hbl, mode, d, lw, lh, maxx, maxy, ef, ft]
                    a = o.predict_intra(px, S, x, y, args, 1 << lw, 1 << lh)
                    b = r.predict_intra(px, S, x, y, args, 1 << lw, 1 << lh)
                    assert np.array_equal(a, b), f"mode {mode} delta {d} size {lw},{lh} args {args}"
                    n += 1
    assert n > 1000


@pytest.mark.parametrize("variant", ["full_rd", "full_rd_352", "mask_no_hpel"])
def test_motion_search_without_quarter_pel_planes(variant):
    """The non-turbo search (level 0 on a step-4 grid over +-32, +-3
    refinement, +-4 below, sub-pel from the EIGHTTAP prediction of the
    block) and a turbo mask without the plane bits (16, 32)."""
    if variant == "full_rd":
        w, h, seed = 136, 104, 17
        fp = _frame_params(w, h, "inter", turbo=0, multi_ref=1)
    elif variant == "full_rd_352":
        w, h, seed = 352, 288, 4
        fp = _frame_params(w, h, "inter", turbo=0, multi_ref=1)
    else:
        w, h, seed = 200, 136, 9
        fp = _frame_params(w, h, "leaf", knob_turbo_mask=0x47CBF1 & ~0x30)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=5, max_vel=10)
    assert not (fp.turbo & 0x10)
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    mo, go = o.motion_search(fp, frames[1], frames[0], frames[2])
    mr, gr = r.motion_search(fp, frames[1], frames[0], frames[2])
    assert np.array_equal(mo, mr)
    if fp.golden:
        assert np.array_equal(go, gr)
    assert np.any(mo % 4 != 0), "no quarter-pel vector: the sub-pel path was not exercised"
