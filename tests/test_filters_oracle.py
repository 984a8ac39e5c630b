"""The oracle's in-loop filter stages against the reference's own kernels
(deblocking, CDEF, loop-restoration apply and parameter search), on frames
coded by the reference's own recon, plus the pinning checks for the two
driver functions whose reference source does not build here (filter info,
CDEF selection)."""
import numpy as np
import pytest

import libs
import mav1hip

needs_ref = pytest.mark.skipif(not libs.ref_available(), reason="oracle/_ref not built (no reference sources)")


def coded_frame(w, h, kind, seed, turbo=1, **kw):
    """A frame reconstructed by the reference (ME, analysis, recon), the
    oracle's filter info of it, and the source."""
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    fr = mav1hip.turbo_frame(kind, 1) if kind in ("key", "arf", "leaf") else mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, npresets = mav1hip.setup_params(cfg, fr)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=10)
    r = libs.Impl("ref")
    n = fp.sb_cols * fp.sb_rows * 170
    mvf, mvg = np.zeros(n, np.int32), np.zeros(n, np.int32)
    if not fp.frame_is_intra:
        mvf, g = r.motion_search(fp, frames[1], frames[0], frames[2])
        mvg = mvg if g is None else g
    _, _, blocks, sbs = r.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    if cfg.intra_edge_filter:
        blocks = libs.edge_types(libs.Impl("oracle"), fp, blocks, sbs)
    rec, _, blocks = r.recon(fp, blocks, sbs, frames[1], frames[0], frames[2], libs.ref_mats())
    info = libs.filter_info(fp, blocks, sbs)
    return cfg, fp, npresets, frames[1], rec, blocks, sbs, info


FRAMES = {
    "turbo_arf_352x288": (352, 288, "arf", 3, {}),
    "turbo_key_200x136": (200, 136, "key", 9, {}),
    "full_rd_136x104": (136, 104, "inter", 7, {"turbo": 0, "multi_ref": 1}),
    "key_full_204x116": (204, 116, "key", 5, {"turbo": 0}),
    "q230_arf_136x104": (136, 104, "arf", 11, {"q": 230}),
    "q40_arf_352x288": (352, 288, "arf", 3, {"q": 40}),
}

_cache = {}


def frame(name):
    if name not in _cache:
        w, h, kind, seed, kw = FRAMES[name]
        _cache[name] = coded_frame(w, h, kind, seed, **dict(kw))
    return _cache[name]


@needs_ref
@pytest.mark.parametrize("name", list(FRAMES))
def test_deblock_equals_reference(name):
    cfg, fp, _, src, rec, blocks, sbs, info = frame(name)
    assert max(fp.lf_level) > 0
    a = libs.Filters("oracle").loop_filter(fp, rec, info)
    b = libs.Filters("ref").loop_filter(fp, rec, info)
    assert not np.array_equal(a, rec), "nothing was filtered"
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("sharp", [0, 1, 3, 5, 7])
def test_deblock_levels_and_sharpness(sharp):
    cfg, fp, _, src, rec, blocks, sbs, info = frame("full_rd_136x104")
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    rng = np.random.default_rng(sharp)
    f.lf_sharpness = sharp
    for i in range(4):
        f.lf_level[i] = int(rng.integers(0, 64))
    a = libs.Filters("oracle").loop_filter(f, rec, info)
    b = libs.Filters("ref").loop_filter(f, rec, info)
    assert np.array_equal(a, b)


def _deblocked(name):
    cfg, fp, npresets, src, rec, blocks, sbs, info = frame(name)
    return fp, src, libs.Filters("ref").loop_filter(fp, rec, info) if libs.ref_available() else None, info


@needs_ref
@pytest.mark.parametrize("name", list(FRAMES))
def test_cdef_equals_reference(name):
    fp, src, dbk, info = _deblocked(name)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    presets = [(0, 0, 0, 0), (2, 1, 1, 0), (4, 2, 2, 1), (8, 4, 4, 2)]
    for k, (yp, ys, up, us) in enumerate(presets):
        f.cdef_y_pri[k], f.cdef_y_sec[k], f.cdef_uv_pri[k], f.cdef_uv_sec[k] = yp, ys, up, us
    a = libs.Filters("oracle").cdef(f, dbk, info, 4)
    b = libs.Filters("ref").cdef(f, dbk, info, 4)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[3], dbk), "nothing was filtered"


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cdef_random_strengths_and_skips(seed):
    fp, src, dbk, info = _deblocked("turbo_key_200x136")
    rng = np.random.default_rng(seed)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    f.cdef_damping = int(rng.integers(3, 7))
    for k in range(8):
        f.cdef_y_pri[k] = int(rng.integers(0, 16))
        f.cdef_y_sec[k] = int(rng.choice([0, 1, 2, 4]))
        f.cdef_uv_pri[k] = int(rng.integers(0, 16))
        f.cdef_uv_sec[k] = int(rng.choice([0, 1, 2, 4]))
    inf = info.copy()
    flip = rng.random(inf.size) < 0.3
    inf[flip] ^= 32
    a = libs.Filters("oracle").cdef(f, dbk, inf, 8)
    b = libs.Filters("ref").cdef(f, dbk, inf, 8)
    assert np.array_equal(a, b)


def test_cdef_select_properties():
    """The chosen set has the smallest visible-area SSE of its superblock
    (first one on ties), rec holds it over the mi-aligned area and is
    unchanged elsewhere."""
    fp, src, dbk, info = _deblocked("turbo_arf_352x288")
    if dbk is None:
        pytest.skip("needs a reconstructed frame")
    rng = np.random.default_rng(4)
    n = 3
    buf = np.stack([dbk, dbk.copy(), dbk.copy()])
    buf[1] = np.clip(buf[1].astype(int) + rng.integers(-2, 3, buf[1].shape), 0, 255)
    buf[2] = src  # zero SSE: wins wherever set 0 is not exact as well
    rec0 = np.full_like(dbk, 7)
    out, choice = libs.cdef_select(fp, src, buf, rec0)
    fw, fh = fp.frame_w, fp.frame_h
    for sr in range(fp.sb_rows):
        for sc in range(fp.sb_cols):
            sse = []
            for k in range(n):
                s = 0
                for p in range(3):
                    sub = p > 0
                    m = 64 >> sub
                    pw, ph = ((fw + 1) >> 1, (fh + 1) >> 1) if p else (fw, fh)
                    ys, xs = slice(sr * m, min((sr + 1) * m, ph)), slice(sc * m, min((sc + 1) * m, pw))
                    plane = lambda a: a[fp.offset[p]:fp.offset[p] + fp.stride[p] * fp.plane_h[p]].reshape(
                        fp.plane_h[p], fp.stride[p])
                    d = plane(buf[k])[ys, xs].astype(np.int64) - plane(src)[ys, xs]
                    s += int((d * d).sum())
                sse.append(s)
            k = int(choice[sr * fp.sb_cols + sc])
            assert sse[k] == min(sse) and all(sse[j] > sse[k] for j in range(k))
            for p in range(3):
                sub = p > 0
                m = 64 >> sub
                mw, mh = (fp.mi_cols * 4) >> sub, (fp.mi_rows * 4) >> sub
                plane = lambda a: a[fp.offset[p]:fp.offset[p] + fp.stride[p] * fp.plane_h[p]].reshape(
                    fp.plane_h[p], fp.stride[p])
                ys, xs = slice(sr * m, min((sr + 1) * m, mh)), slice(sc * m, min((sc + 1) * m, mw))
                assert np.array_equal(plane(out)[ys, xs], plane(buf[k])[ys, xs])
    covered = np.zeros_like(dbk, bool)
    for p in range(3):
        sub = p > 0
        mw, mh = (fp.mi_cols * 4) >> sub, (fp.mi_rows * 4) >> sub
        v = covered[fp.offset[p]:fp.offset[p] + fp.stride[p] * fp.plane_h[p]].reshape(fp.plane_h[p], fp.stride[p])
        v[:mh, :mw] = True
    assert np.all(out[~covered] == 7)
    # a tie: two identical sets, the first one is kept everywhere
    out2, choice2 = libs.cdef_select(fp, src, np.stack([buf[1], buf[1]]), rec0)
    assert np.all(choice2 == 0)


def test_filter_info_known_layout():
    """Two blocks of a hand-built superblock: bsize | skip << 5 | depth << 6
    over each block's mi area, cropped at the frame edge."""
    cfg = mav1hip.default_config(72, 40, turbo=1, qindex=125)
    fp, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("key", 0))
    nsb = fp.sb_cols * fp.sb_rows
    blocks = np.zeros((nsb * 64, 27), np.int32)
    sbs = np.zeros((nsb, 89), np.int32)
    sbs[0, 1] = 2
    sbs[1, 0], sbs[1, 1] = 64, 1
    # block 0: 32x32 (bsize 9) at mi (0,0), eob (0,0,0) -> skip, depth 1
    blocks[0, 0:3] = (0, 0, 9)
    blocks[0, 19] = 1
    # block 1: 16x16 (bsize 6) at mi (8, 4), eob[1] set, depth 0
    blocks[1, 0:3] = (8, 4, 6)
    blocks[1, 12] = 1
    # superblock 1's block: 64x64 (bsize 12) at mi (0, 16), cropped to the frame
    blocks[64, 0:3] = (0, 16, 12)
    info = libs.filter_info(fp, blocks, sbs).reshape(fp.mi_rows, fp.mi_cols)
    assert fp.mi_cols == 18 and fp.mi_rows == 10
    assert np.all(info[0:8, 0:8] == (9 | 32 | 64))
    assert np.all(info[8:10, 4:8] == 6)
    assert np.all(info[0:10, 16:18] == (12 | 32))
    assert np.all(info[0:8, 8:16] == 0) and np.all(info[8:10, 0:4] == 0)


def _random_units(fp, rng):
    n = sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
    u = np.zeros((n, 12), np.int32)
    lo, hi = [-5, -23, -17], [10, 8, 46]
    for i in range(n):
        t = int(rng.integers(0, 3))
        u[i, 0] = t
        if t == 1:
            for ps in range(2):
                for j in range(3):
                    u[i, 4 + ps * 3 + j] = int(rng.integers(lo[j], hi[j] + 1))
        elif t == 2:
            u[i, 1] = int(rng.integers(0, 16))
            u[i, 2] = int(rng.integers(-96, 32))
            u[i, 3] = int(rng.integers(-32, 96))
    return u


@needs_ref
@pytest.mark.parametrize("name", ["turbo_arf_352x288", "key_full_204x116", "q230_arf_136x104"])
def test_lr_apply_equals_reference(name):
    fp, src, dbk, info = _deblocked(name)
    cd = libs.Filters("ref").cdef(fp, dbk, info, 4)[2]
    rng = np.random.default_rng(len(name))
    units = _random_units(fp, rng)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    for p in range(3):
        f.lr_type[p] = 3
    a = libs.Filters("oracle").lr_apply(f, cd, dbk, units)
    b = libs.Filters("ref").lr_apply(f, cd, dbk, units)
    assert np.array_equal(a, b)
    # every restoration type occurs
    assert set(units[:, 0]) == {0, 1, 2}


@needs_ref
@pytest.mark.parametrize("name,lr_test", [("q40_arf_352x288", 0), ("full_rd_136x104", 1), ("key_full_204x116", 1),
                                          ("q230_arf_136x104", 1), ("turbo_key_200x136", 0)])
def test_lr_select_equals_reference(name, lr_test):
    """Units, types and taps equal; at q 40 the RD search restores some
    units, lr_test forces Wiener and self-guided units."""
    fp, src, dbk, info = _deblocked(name)
    cd = libs.Filters("ref").cdef(fp, dbk, info, 4)[2]
    ua, ta = libs.Filters("oracle").lr_select(fp, src, cd, dbk, lr_test)
    ub, tb = libs.Filters("ref").lr_select(fp, src, cd, dbk, lr_test)
    assert ta == tb
    assert np.array_equal(ua, ub)
    if lr_test or name.startswith("q40"):
        assert (ua[:, 0] == 1).any() and (ua[:, 0] == 2).any(), "both filters occur"


def test_filter_plan_turbo_gop():
    """run_frame_core's decisions for the turbo mini-GOP (src/encoder.c:
    618-640): the ARF is deblocked, CDEF'd with two sets and selected, and
    restored only when restoration is on; leaf frames filter nothing; a key
    frame with CDEF off still deblocks."""
    for lr in (0, 1):
        cfg = mav1hip.default_config(1920, 1080, turbo=1, qindex=125, restoration=lr)
        for kind in ("key", "arf", "leaf"):
            fr = mav1hip.turbo_frame(kind, 1)
            fp, n = mav1hip.setup_params(cfg, fr)
            plan = mav1hip.filter_plan(cfg, fr, fp, n)
            if kind == "leaf":
                assert (plan.filter_info, plan.deblock, plan.cdef_sets, plan.cdef_select, plan.restoration) == (0, 0, 0, 0, 0)
            else:
                assert plan.filter_info and plan.deblock and plan.cdef_sets == 2 and plan.cdef_select
                assert plan.restoration == lr
    cfg = mav1hip.default_config(352, 288, turbo=0, qindex=125, cdef=0)
    fr = mav1hip.turbo_frame("key", 0)
    fp, n = mav1hip.setup_params(cfg, fr)
    plan = mav1hip.filter_plan(cfg, fr, fp, n)
    assert n == 0 and plan.filter_info and plan.deblock and plan.cdef_sets == 0 and not plan.cdef_select
    cfg = mav1hip.default_config(352, 288, turbo=0, qindex=0)
    fp, n = mav1hip.setup_params(cfg, fr)
    plan = mav1hip.filter_plan(cfg, fr, fp, n)
    assert not plan.deblock and not plan.cdef_select and not plan.restoration


def test_decision_grid_known_layout():
    """The dump_decisions cells of a hand-built superblock: a 32x32 intra
    block, a compound 16x16 block, four 4x4 blocks sharing one 8x8 cell (the
    last in decode order wins), a cell no block covers stays zero."""
    import ctypes as C
    cfg = mav1hip.default_config(64, 64, turbo=0, qindex=125)
    fp, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("key", 0))
    blocks = np.zeros((64, 27), np.int32)
    sbs = np.zeros((1, 89), np.int32)
    rows = [
        (0, 0, 9, 0, 0, 0, 0, 0),        # mi_row, mi_col, bsize, is_inter, ref_frame, ref_frame1, mv_row, mv_col
        (0, 8, 6, 1, 1, 7, -12, 34),
        (8, 8, 0, 1, 1, 0, 5, -6),
        (8, 9, 0, 1, 2, 0, 7, -8),
        (9, 8, 0, 0, 0, 0, 0, 0),
        (9, 9, 0, 1, 1, 0, -3, 300),
    ]
    for k, (r, c, bs, inter, rf, rf1, mr, mc) in enumerate(rows):
        blocks[k, 0:4] = (r, c, bs, inter)
        blocks[k, 8:10] = (mr, mc)
        blocks[k, 23:25] = (rf, rf1)
    sbs[0, 1] = len(rows)
    lib = libs.oracle()
    lib.oracle_decision_grid.argtypes = [C.c_void_p] * 4
    grid = np.full(8 * 8 * 8, 0xAB, np.uint8)
    lib.oracle_decision_grid(C.addressof(fp), mav1hip.ptr(blocks), mav1hip.ptr(sbs), mav1hip.ptr(grid))
    cells = grid.reshape(8, 8, 8)

    def cell(bs, mode, rf, mr, mc):
        return np.frombuffer(np.array([bs, mode, rf, 0], np.uint8).tobytes() + np.array([mr, mc], np.int16).tobytes(),
                             np.uint8)

    assert all(np.array_equal(cells[r, c], cell(9, 0, 0, 0, 0)) for r in range(4) for c in range(4))
    assert all(np.array_equal(cells[r, c], cell(6, 2, 1, -12, 34)) for r in range(2) for c in range(4, 6))
    assert np.array_equal(cells[4, 4], cell(0, 1, 1, -3, 300))
    assert not cells[4, 5].any() and not cells[7, 7].any() and not cells[0, 6].any()
