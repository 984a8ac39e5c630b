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
