"""In-loop filters on the GPU against the oracle, bit-exact: filter info,
deblocking, CDEF sets, the per-superblock CDEF choice and loop-restoration
apply, per stage through the host-buffer entry points and as the
device-resident chain the benchmark runs."""
import numpy as np
import pytest

import libs
import mav1hip

pytestmark = pytest.mark.gpu


def coded(w, h, kind, seed, turbo=1, **kw):
    """A frame reconstructed by the oracle pipeline."""
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    fr = mav1hip.turbo_frame(kind, 1) if kind in ("key", "arf", "leaf") else mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, npresets = mav1hip.setup_params(cfg, fr)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=10)
    o = libs.Impl("oracle")
    n = fp.sb_cols * fp.sb_rows * 170
    mvf, mvg = np.zeros(n, np.int32), np.zeros(n, np.int32)
    if not fp.frame_is_intra:
        mvf, g = o.motion_search(fp, frames[1], frames[0], frames[2])
        mvg = mvg if g is None else g
    _, _, blocks, sbs = o.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    if cfg.intra_edge_filter:
        blocks = libs.edge_types(o, fp, blocks, sbs)
    rec, _, blocks = o.recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mav1hip.forward_matrices())
    return cfg, fr, fp, npresets, frames, rec, blocks, sbs


CASES = {
    "turbo_arf_352x288": (352, 288, "arf", 3, {}),
    "turbo_key_200x136": (200, 136, "key", 9, {}),
    "full_rd_136x104": (136, 104, "inter", 7, {"turbo": 0, "multi_ref": 1}),
    "key_full_204x116": (204, 116, "key", 5, {"turbo": 0}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_filter_stages_equal_oracle(gpu, name):
    w, h, kind, seed, kw = CASES[name]
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(w, h, kind, seed, **dict(kw))
    src = frames[1]
    ctx = mav1hip.Context(cfg)
    try:
        info = ctx.filter_info(fp, blocks, sbs)
        assert np.array_equal(info, libs.filter_info(fp, blocks, sbs)), "filter info differs"
        f = mav1hip.FrameParams.from_buffer_copy(fp)
        presets = [(0, 0, 0, 0), (2, 1, 1, 0), (4, 2, 2, 1), (8, 4, 4, 2)]
        for k, (yp, ys, up, us) in enumerate(presets):
            f.cdef_y_pri[k], f.cdef_y_sec[k], f.cdef_uv_pri[k], f.cdef_uv_sec[k] = yp, ys, up, us
        dg, cg = ctx.filter(f, info, rec, 4)
        of = libs.Filters("oracle")
        do = of.loop_filter(f, rec, info)
        assert np.array_equal(dg, do), "deblocking differs"
        co = of.cdef(f, do, info, 4)
        assert np.array_equal(cg, co), "CDEF differs"
        base = np.full_like(rec, 3)
        sg, chg = ctx.cdef_select(f, src, co, base)
        so, cho = libs.cdef_select(f, src, co, base)
        assert np.array_equal(chg, cho), "CDEF choice differs"
        assert np.array_equal(sg, so), "CDEF selection output differs"
        rng = np.random.default_rng(seed)
        units = _random_units(fp, rng)
        for p in range(3):
            f.lr_type[p] = 3
        lg = ctx.restore(f, so, do, units)
        lo = of.lr_apply(f, so, do, units)
        assert np.array_equal(lg, lo), "loop restoration differs"
    finally:
        ctx.close()


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


@pytest.mark.parametrize("sharp,seed", [(0, 1), (3, 2), (7, 3)])
def test_deblock_levels_and_cdef_strengths(gpu, sharp, seed):
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(136, 104, "inter", 7, turbo=0, multi_ref=1)
    info = libs.filter_info(fp, blocks, sbs)
    rng = np.random.default_rng(seed)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    f.lf_sharpness = sharp
    for i in range(4):
        f.lf_level[i] = int(rng.integers(0, 64))
    f.cdef_damping = int(rng.integers(3, 7))
    for k in range(8):
        f.cdef_y_pri[k] = int(rng.integers(0, 16))
        f.cdef_y_sec[k] = int(rng.choice([0, 1, 2, 4]))
        f.cdef_uv_pri[k] = int(rng.integers(0, 16))
        f.cdef_uv_sec[k] = int(rng.choice([0, 1, 2, 4]))
    inf = info.copy()
    inf[rng.random(inf.size) < 0.3] ^= 32
    ctx = mav1hip.Context(cfg)
    try:
        dg, cg = ctx.filter(f, inf, rec, 8)
    finally:
        ctx.close()
    of = libs.Filters("oracle")
    do = of.loop_filter(f, rec, inf)
    assert np.array_equal(dg, do)
    assert np.array_equal(cg, of.cdef(f, do, inf, 8))


def test_device_chain_1080p_turbo_arf(gpu):
    """The benchmark's device-resident filter chain on a 1080p turbo ARF
    frame with restoration on: filter info, deblocking, CDEF sets,
    deblocked copy, CDEF choice, restoration with the units given."""
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(1920, 1080, "arf", 21, restoration=1)
    plan = mav1hip.filter_plan(cfg, fr, fp, npresets)
    assert plan.filter_info and plan.deblock and plan.cdef_sets == 2 and plan.cdef_select and plan.restoration
    src = frames[1]
    ctx = mav1hip.Context(cfg)
    try:
        ctx.set_params(fp)
        ctx.upload(mav1hip.BUF_SRC, src)
        ctx.upload(mav1hip.BUF_REC, rec)
        ctx.upload(mav1hip.BUF_BLOCKS, blocks)
        ctx.upload(mav1hip.BUF_SBS, sbs)
        units = _random_units(fp, np.random.default_rng(5))
        f = mav1hip.FrameParams.from_buffer_copy(fp)
        for p in range(3):
            f.lr_type[p] = 3
        ctx.upload(mav1hip.BUF_LRUNITS, units)
        ctx.launch_filter_info()
        ctx.launch_deblock()
        ctx.launch_cdef(plan.cdef_sets)
        ctx.launch_save_deblocked()
        ctx.launch_cdef_select(plan.cdef_sets)
        ctx.set_params(f)
        ctx.launch_restore()
        fb = ctx.frame_bytes
        out = ctx.download(mav1hip.BUF_LROUT)
        cdef_rec = ctx.download(mav1hip.BUF_REC)
        choice = ctx.download(mav1hip.BUF_CDEF_CHOICE, np.int8)
    finally:
        ctx.close()
    of = libs.Filters("oracle")
    info = libs.filter_info(fp, blocks, sbs)
    do = of.loop_filter(fp, rec, info)
    co = of.cdef(fp, do, info, plan.cdef_sets)
    so, cho = libs.cdef_select(fp, src, co, do)
    assert np.array_equal(choice, cho)
    assert np.array_equal(cdef_rec, so)
    lo = of.lr_apply(f, so, do, units)
    # outside the planes' visible areas the output buffer is not written
    for p in range(3):
        sub = p > 0
        pw, ph = (fp.frame_w + sub) >> sub, (fp.frame_h + sub) >> sub
        view = lambda a: a[fp.offset[p]:fp.offset[p] + fp.stride[p] * fp.plane_h[p]].reshape(fp.plane_h[p], fp.stride[p])
        assert np.array_equal(view(out[:fb])[:ph, :pw], view(lo)[:ph, :pw]), f"plane {p} differs"


LR_CASES = {
    "q40_arf_352x288": ((352, 288, "arf", 3, {"q": 40}), 0),
    "q125_arf_352x288_test": ((352, 288, "arf", 3, {}), 1),
    "full_rd_136x104_test": ((136, 104, "inter", 7, {"turbo": 0, "multi_ref": 1}), 1),
    "key_204x116": ((204, 116, "key", 5, {"turbo": 0, "q": 60}), 0),
    "q80_arf_640x360": ((640, 360, "arf", 13, {"q": 80}), 0),
    "q125_arf_1920x1080": ((1920, 1080, "arf", 21, {}), 0),
    "q170_arf_1920x1080_test": ((1920, 1080, "arf", 22, {"q": 170}), 1),
}


@pytest.mark.parametrize("name", list(LR_CASES))
def test_lr_select_equals_oracle(gpu, name):
    """Restoration parameters per unit and the per-plane types equal the
    oracle's search (greedy Wiener refinement with the running references
    included), then the restored frame with them."""
    (w, h, kind, seed, kw), lr_test = LR_CASES[name]
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(w, h, kind, seed, **dict(kw))
    src = frames[1]
    info = libs.filter_info(fp, blocks, sbs)
    of = libs.Filters("oracle")
    dbk = of.loop_filter(fp, rec, info)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    for k, (yp, ys, up, us) in enumerate([(0, 0, 0, 0), (4, 2, 2, 1)]):
        f.cdef_y_pri[k], f.cdef_y_sec[k], f.cdef_uv_pri[k], f.cdef_uv_sec[k] = yp, ys, up, us
    cd, _ = libs.cdef_select(f, src, of.cdef(f, dbk, info, 2), dbk)
    uo, to = of.lr_select(fp, src, cd, dbk, lr_test)
    ctx = mav1hip.Context(cfg)
    try:
        ug, tg, reruns = ctx.lr_select(fp, src, cd, dbk, lr_test)
        assert tg == to, "restoration types differ"
        bad = np.nonzero((ug != uo).any(axis=1))[0]
        assert bad.size == 0, f"units differ at {bad[:8]}: gpu {ug[bad[:2]]} oracle {uo[bad[:2]]}"
        f2 = mav1hip.FrameParams.from_buffer_copy(fp)
        for p in range(3):
            f2.lr_type[p] = to[p]
        lg = ctx.restore(f2, cd, dbk, ug)
        assert np.array_equal(lg, of.lr_apply(f2, cd, dbk, uo)), "restored frame differs"
    finally:
        ctx.close()
    print(name, "units", len(uo), "types", np.bincount(uo[:, 0], minlength=3), "reruns", reruns)


@pytest.mark.parametrize("name", ["turbo_arf_352x288", "key_full_204x116", "full_rd_136x104"])
def test_decision_record_equals_oracle(gpu, name):
    """The dump_decisions record: header fields and the per-8x8 cells of the
    coded blocks equal the oracle's grid."""
    import ctypes as C
    w, h, kind, seed, kw = CASES[name]
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(w, h, kind, seed, **dict(kw))
    hdr7 = [5, 4, 6, 1, 2, 1, 12345]
    ctx = mav1hip.Context(cfg)
    try:
        rec_bytes = ctx.decision_record(fp, blocks, sbs, hdr7)
    finally:
        ctx.close()
    rows8, cols8 = (fp.mi_rows + 1) >> 1, (fp.mi_cols + 1) >> 1
    hdr = np.frombuffer(rec_bytes[:44].tobytes(), np.int32)
    assert list(hdr) == [0x4345444D, 5, 4, 6, 1, 2, 1, fp.base_q_idx, rows8, cols8, 12345]
    lib = libs.oracle()
    lib.oracle_decision_grid.argtypes = [C.c_void_p] * 4
    grid = np.zeros(rows8 * cols8 * 8, np.uint8)
    lib.oracle_decision_grid(C.addressof(fp), mav1hip.ptr(np.ascontiguousarray(blocks, np.int32)),
                             mav1hip.ptr(np.ascontiguousarray(sbs, np.int32)), mav1hip.ptr(grid))
    assert np.array_equal(rec_bytes[44:], grid)
