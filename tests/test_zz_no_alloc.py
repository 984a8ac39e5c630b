"""mav1hip_lr_select_static (the restoration search with the library's
static scratch, no allocation) equals mav1hip_lr_select (context buffers)
on two calls in a row, the second on a smaller frame; a frame with more
units than the scratch holds is refused."""
import numpy as np
import pytest

import libs
import mav1hip
from test_filters_gpu import coded

pytestmark = pytest.mark.gpu


def _inputs(w, h, kind, seed, q):
    cfg, fr, fp, npresets, frames, rec, blocks, sbs = coded(w, h, kind, seed, q=q)
    src = frames[1]
    info = libs.filter_info(fp, blocks, sbs)
    of = libs.Filters("oracle")
    dbk = of.loop_filter(fp, rec, info)
    f = mav1hip.FrameParams.from_buffer_copy(fp)
    for k, (yp, ys, up, us) in enumerate([(0, 0, 0, 0), (4, 2, 2, 1)]):
        f.cdef_y_pri[k], f.cdef_y_sec[k], f.cdef_uv_pri[k], f.cdef_uv_sec[k] = yp, ys, up, us
    cd, _ = libs.cdef_select(f, src, of.cdef(f, dbk, info, 2), dbk)
    return cfg, fp, src, cd, dbk


def test_lr_select_static_equals_context_entry_point(gpu):
    import torch
    cases = [(640, 360, "arf", 13, 80, 0), (352, 288, "arf", 3, 125, 1)]  # the second call is smaller
    dev = torch.device("cuda:0")
    for w, h, kind, seed, q, lr_test in cases:
        print(f"lr_select_static {w}x{h} q{q} lr_test {lr_test}", flush=True)
        cfg, fp, src, cd, dbk = _inputs(w, h, kind, seed, q)
        ctx = mav1hip.Context(cfg)
        try:
            ue, te, _ = ctx.lr_select(fp, src, cd, dbk, lr_test)
        finally:
            ctx.close()
        n = sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
        d_src, d_cd, d_dbk = (torch.from_numpy(a).to(dev) for a in (src, cd, dbk))
        # outputs pre-filled with garbage: the search must write every record
        d_units = torch.full((n, 12), -7, dtype=torch.int32, device=dev)
        d_type = torch.full((3,), -7, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev)
        mav1hip.lr_select_static(fp, d_src.data_ptr(), d_cd.data_ptr(), d_dbk.data_ptr(), lr_test,
                                 d_units.data_ptr(), d_type.data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize(dev)
        us, ts = d_units.cpu().numpy(), [int(v) for v in d_type.cpu()]
        assert ts == te, f"restoration types differ: static {ts} context {te}"
        bad = np.nonzero((us != ue).any(axis=1))[0]
        assert bad.size == 0, f"units differ at {bad[:8]}"


def test_lr_select_static_refuses_too_many_units(gpu):
    # 4096x2304: 64 x 36 luma units and 2 x 32 x 18 chroma units = 3456 > 3072
    cfg = mav1hip.default_config(4096, 2304, turbo=1, qindex=125, restoration=1)
    fp, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("arf", 1))
    n = sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
    assert n > mav1hip.STATIC_LR_UNITS
    with pytest.raises(RuntimeError, match="static scratch holds"):
        # refused before any pointer is used
        mav1hip.lr_select_static(fp, 1, 1, 1, 0, 1, 1)
