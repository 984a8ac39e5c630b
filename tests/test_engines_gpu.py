"""Frame engines as separate contexts (encoder.c:1150-1224 runs a stream's
mini-GOP on engine encoders, each with its own backend): a leaf frame coded
on a second context from references imported with mav1hip_pool_import
equals the same leaf coded on the context that coded its references, and
the import is ordered on both streams (the source slot is overwritten right
after the import is enqueued)."""
import numpy as np
import pytest

import libs
import mav1hip

pytestmark = pytest.mark.gpu


def run_stages(ctx, plan, kind):
    if kind != "key":
        ctx.launch_motion_search()
    ctx.launch_analyze()
    ctx.launch_edge_types()
    ctx.launch_recon()
    if plan.filter_info:
        ctx.launch_filter_info()
        if plan.deblock:
            ctx.launch_deblock()
        ctx.launch_cdef(plan.cdef_sets)
        if plan.restoration:
            ctx.launch_save_deblocked()
        if plan.cdef_select:
            ctx.launch_cdef_select(plan.cdef_sets)
    if plan.restoration:
        ctx.launch_lr_select(0)
        ctx.launch_restore()
    return mav1hip.BUF_LROUT if plan.restoration else mav1hip.BUF_REC


def used_blocks(ctx, fp):
    """The block records the superblocks list (the buffer's unused tail
    keeps whatever earlier frames left there)."""
    nsb = fp.sb_cols * fp.sb_rows
    blk = ctx.download(mav1hip.BUF_BLOCKS, np.int32).reshape(-1, 27)
    sbs = ctx.download(mav1hip.BUF_SBS, np.int32, nsb * 89).reshape(-1, 89)
    return np.concatenate([blk[f:f + n] for f, n in zip(sbs[:, 0], sbs[:, 1])])


@pytest.mark.parametrize("w,h,seed", [(352, 288, 31), (200, 136, 32)])
def test_leaf_on_second_engine_equals_one_engine(gpu, w, h, seed):
    cfg = mav1hip.default_config(w, h, turbo=1, qindex=125, restoration=1)
    geo, frames = libs.make_sequence(w, h, 5, seed, n_objects=8, max_vel=10)
    params, plans = {}, {}
    for kind in ("key", "arf", "leaf"):
        fr = mav1hip.turbo_frame(kind, 0 if kind == "key" else 1)
        fp, npre = mav1hip.setup_params(cfg, fr)
        params[kind], plans[kind] = fp, mav1hip.filter_plan(cfg, fr, fp, npre)
    n = len(frames)
    a, b = mav1hip.Context(cfg), mav1hip.Context(cfg)
    try:
        for ctx in (a, b):
            ctx.pool_alloc(n + 2)
            for i, f in enumerate(frames):
                ctx.pool_upload(i, f)

        def code(ctx, kind, src, ref, gold):
            ctx.bind_frames(src, ref, gold)
            ctx.set_params(params[kind])
            return run_stages(ctx, plans[kind], kind)

        # engine A: key frame 0 into slot n, ALTREF 2 (LAST = 0) into slot n + 1
        a.pool_store(n, code(a, "key", 0, -1, -1))
        a.pool_store(n + 1, code(a, "arf", 2, n, -1))
        # one engine: leaf 1 on A from its own slots
        fin = code(a, "leaf", 1, n, n + 1)
        a.sync()
        ref_out = a.download(fin)
        ref_blocks = used_blocks(a, params["leaf"])
        # two engines: B imports both references, then A overwrites its slot n
        # at once (the import must have read it first)
        b.bind_frames(1, -1, -1)
        b.pool_import(n, a, n)
        b.pool_import(n + 1, a, n + 1)
        a.bind_frames(0, -1, -1)
        a.pool_upload(n, np.zeros_like(frames[0]))
        a.pool_store(n, code(a, "arf", 4, n + 1, -1))
        out = code(b, "leaf", 1, n, n + 1)
        b.sync()
        a.sync()
        assert np.array_equal(b.download(out), ref_out)
        assert np.array_equal(used_blocks(b, params["leaf"]), ref_blocks)
    finally:
        a.close()
        b.close()
