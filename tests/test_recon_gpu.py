"""Edge filter types and reconstruction on the GPU against the oracle,
bit-exact: reconstructed frame, quantised levels and every block's
reconstruction outputs."""
import numpy as np
import pytest

import libs
import mav1hip

pytestmark = pytest.mark.gpu


def _setup(w, h, kind, seed, turbo=1, **kw):
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    fr = mav1hip.turbo_frame(kind, 1) if kind in ("key", "arf", "leaf") else mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, _ = mav1hip.setup_params(cfg, fr)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=10)
    o = libs.Impl("oracle")
    mvf = mvg = None
    if not fp.frame_is_intra:
        mvf, mvg = o.motion_search(fp, frames[1], frames[0], frames[2])
    n = fp.sb_cols * fp.sb_rows * 170
    mvf = np.zeros(n, np.int32) if mvf is None else mvf
    mvg = np.zeros(n, np.int32) if mvg is None else mvg
    _, _, blocks, sbs = o.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    return cfg, fp, frames, blocks, sbs


def _check(cfg, fp, frames, blocks, sbs):
    o = libs.Impl("oracle")
    ctx = mav1hip.Context(cfg)
    if cfg.intra_edge_filter:
        eg = ctx.edge_types(fp, blocks, sbs)
        eo = libs.edge_types(o, fp, blocks, sbs)
        assert np.array_equal(eg, eo), "edge filter types differ"
        blocks = eo
    mats = mav1hip.forward_matrices()
    rg, cg, bg = ctx.recon(fp, blocks, sbs, frames[1], frames[0], frames[2])
    ro, co, bo = o.recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    ctx.close()
    bad = np.nonzero(rg != ro)[0]
    assert bad.size == 0, f"reconstruction differs at {bad.size} bytes, first {bad[:6]}"
    assert np.array_equal(cg, co), f"levels differ at {np.nonzero(cg != co)[0][:6]}"
    for s in range(fp.sb_cols * fp.sb_rows):
        n = sbs[s, 1]
        assert np.array_equal(bg[s * 64:s * 64 + n], bo[s * 64:s * 64 + n]), f"block outputs of SB {s} differ"


@pytest.mark.parametrize("kind", ["key", "arf", "leaf"])
@pytest.mark.parametrize("w,h,seed", [(352, 288, 3), (200, 136, 9), (854, 480, 6), (1920, 1080, 5)])
def test_recon_turbo(gpu, kind, w, h, seed):
    print(f"\n[recon] {w}x{h} seed {seed} {kind}", flush=True)
    _check(*_setup(w, h, kind, seed))


@pytest.mark.parametrize("variant", ["full_rd", "key_full", "lossless", "q_low", "no_edge", "full_rd_q200"])
def test_recon_other_presets(gpu, variant):
    w, h = 136, 104
    print(f"\n[recon] preset {variant}", flush=True)
    if variant == "full_rd":
        args = _setup(w, h, "inter", 7, turbo=0, multi_ref=1)
    elif variant == "full_rd_q200":
        args = _setup(w, h, "inter", 8, turbo=0, multi_ref=1, q=200)
    elif variant == "key_full":
        args = _setup(w, h, "key", 7, turbo=0)
    elif variant == "lossless":
        args = _setup(w, h, "key", 7, turbo=0, q=0)
    elif variant == "q_low":
        args = _setup(w, h, "arf", 7, q=30)
    else:
        args = _setup(w, h, "key", 7, turbo=0, intra_edge_filter=0)
    _check(*args)


@pytest.mark.parametrize("kind,w,h,seed", [("key", 352, 288, 11), ("key", 200, 136, 12), ("arf", 352, 288, 13)])
def test_recon_forced_intra_modes(gpu, kind, w, h, seed):
    """Every intra block gets a random luma mode and angle, and a random
    chroma mode (CfL on blocks up to 32x32) and angle: all prediction paths,
    the edge filter and upsampling cases, CfL's alpha search (U, then V
    with alpha 0 excluded when U took 0) against the oracle."""
    print(f"\n[recon] forced modes {kind} {w}x{h} seed {seed}", flush=True)
    cfg, fp, frames, blocks, sbs = _setup(w, h, kind, seed)
    blocks, n_cfl, n_dir = libs.force_intra_modes(fp, blocks, sbs, seed)
    print(f"  intra blocks with CfL {n_cfl}, directional luma {n_dir}", flush=True)
    assert n_cfl > 0 and n_dir > 0
    _check(cfg, fp, frames, blocks, sbs)
