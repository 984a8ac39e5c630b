"""Transforms, quantiser and full-frame reconstruction of the oracle
against the reference's own kernels (oracle/_ref), plus closed-form checks
that hold without the reference."""
import ctypes as C

import numpy as np
import pytest

import libs
import mav1hip

needs_ref = pytest.mark.skipif(not libs.ref_available(), reason="oracle/_ref not built (needs the reference sources)")


@needs_ref
def test_inverse_transforms_all_sizes_and_types():
    rng = np.random.default_rng(1)
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    n = 0
    for tx in range(19):
        for t in range(16):
            if not libs.valid_tx(tx, t):
                continue
            for trial in range(6):
                dq = np.zeros((64, 64), np.int32)
                tw, th = min(32, libs.TXW[tx]), min(32, libs.TXH[tx])
                amp = [50, 400, 3000, 32767][trial % 4]
                blk = rng.integers(-amp, amp + 1, (th, tw)).astype(np.int32)
                blk[rng.random((th, tw)) < 0.6] = 0
                dq[:th, :tw] = blk
                a = np.zeros((64, 64), np.int32)
                b = np.zeros((64, 64), np.int32)
                o.fn("itx_2d")(mav1hip.ptr(dq), mav1hip.ptr(a), tx, t, 0)
                r.fn("itx_2d")(mav1hip.ptr(dq.copy()), mav1hip.ptr(b), tx, t, 0)
                assert np.array_equal(a, b), f"itx size {tx} type {t}"
                n += 1
    # lossless WHT
    for trial in range(20):
        dq = np.zeros((64, 64), np.int32)
        dq[:4, :4] = rng.integers(-500, 500, (4, 4))
        a = np.zeros((64, 64), np.int32)
        b = np.zeros((64, 64), np.int32)
        o.fn("itx_2d")(mav1hip.ptr(dq), mav1hip.ptr(a), 0, 0, 1)
        r.fn("itx_2d")(mav1hip.ptr(dq.copy()), mav1hip.ptr(b), 0, 0, 1)
        assert np.array_equal(a, b)
    assert n > 500


@needs_ref
def test_forward_transform_and_quantiser():
    rng = np.random.default_rng(2)
    mats = libs.ref_mats()
    o, r = libs.Impl("oracle"), libs.Impl("ref")
    for tx in range(19):
        for t in range(16):
            if not libs.valid_tx(tx, t):
                continue
            w, h = libs.TXW[tx], libs.TXH[tx]
            res = np.zeros((64, 64), np.int32)
            res[:h, :w] = rng.integers(-255, 256, (h, w))
            a = np.zeros((64, 64), np.int32)
            b = np.zeros((64, 64), np.int32)
            o.fn("fwd_2d")(mav1hip.ptr(res), mav1hip.ptr(a), tx, t, mav1hip.ptr(mats))
            r.fn("fwd_2d")(mav1hip.ptr(res), mav1hip.ptr(b), tx, t, mav1hip.ptr(mats))
            assert np.array_equal(a, b), f"fwd size {tx} type {t}"
            for q in (4, 37, 140, 600, 1336):
                qa = np.zeros(1024, np.int32)
                qb = np.zeros(1024, np.int32)
                o.fn("quantize")(mav1hip.ptr(a), mav1hip.ptr(qa), tx, q, q + 3)
                r.fn("quantize")(mav1hip.ptr(a), mav1hip.ptr(qb), tx, q, q + 3)
                assert np.array_equal(qa, qb), f"quantize size {tx} q {q}"
                da = np.zeros((64, 64), np.int32)
                db = np.zeros((64, 64), np.int32)
                o.fn("dequant")(mav1hip.ptr(qa), mav1hip.ptr(da), tx, q, q + 3)
                r.fn("dequant")(mav1hip.ptr(qa), mav1hip.ptr(db), tx, q, q + 3)
                assert np.array_equal(da, db)


def test_quantiser_reciprocal_equals_division():
    """The reference quantiser's reciprocal path, level = t * (2^44 / q + 1) >> 44
    for t < 2^19, equals the exact division t / q (the formula of the
    reference's tests/test_neon.c:74-86) for every q of the 8-bit tables."""
    import re
    import os
    txt = open(os.path.join(libs.ROOT, "mav1-encoder_amd", "csrc", "av1_tables.h")).read()
    def table(name):
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", txt, re.S).group(1)
        return [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    qs = sorted(set(table("Dc_Qlookup")[:256] + table("Ac_Qlookup")[:256]))
    t = np.arange(1 << 19, dtype=np.uint64)
    for q in qs:
        inv = np.uint64((1 << 44) // q + 1)
        fast = (t * inv) >> np.uint64(44)
        assert np.array_equal(fast, t // np.uint64(q)), f"q {q}"


def test_quantiser_float_division_is_exact():
    """The device quantiser's division (tx_qdiv in csrc/dev_tx.h): the
    truncation of float(t) * (1.0f / q), both products rounded to nearest
    float as the GPU's v_mul_f32 / division do, corrected by the remainder
    t - l q once in either direction, equals t // q for every t < 2^19 and
    every q of the 8-bit tables; without the correction it does not."""
    import re
    import os
    txt = open(os.path.join(libs.ROOT, "mav1-encoder_amd", "csrc", "av1_tables.h")).read()
    def table(name):
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", txt, re.S).group(1)
        return [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    qs = sorted(set(table("Dc_Qlookup")[:256] + table("Ac_Qlookup")[:256]))
    t = np.arange(1 << 19, dtype=np.int64)
    tf = t.astype(np.float32)
    corrected_any = False
    for q in qs:
        rq = np.float32(1.0) / np.float32(q)
        l0 = (tf * rq).astype(np.int64)  # truncation toward zero, t >= 0
        r = t - l0 * q
        l = l0 + (r >= q) - (r < 0)
        assert np.array_equal(l, t // q), f"q {q}"
        corrected_any |= not np.array_equal(l0, t // q)
    assert corrected_any


def test_inverse_dct_is_the_scaled_orthonormal_dct():
    """Closed form: the integer inverse DCT of length N applied to a unit
    coefficient at k approximates sqrt(2/N) c_k cos(pi (2i+1) k / 2N) * 2^-?
    scaled so that a DC coefficient of 4096 gives a flat row; checked to
    within rounding for every length (catches a wrong angle or permutation)."""
    o = libs.Impl("oracle")
    for tx, n in ((0, 4), (1, 8), (2, 16), (3, 32), (4, 64)):
        for k in range(min(n, 32)):
            dq = np.zeros((64, 64), np.int32)
            dq[0, k] = 1024 * 8
            res = np.zeros((64, 64), np.int32)
            o.fn("itx_2d")(mav1hip.ptr(dq), mav1hip.ptr(res), tx, 0, 0)
            row = res[0, :n].astype(np.float64)
            basis = np.cos(np.pi * (2 * np.arange(n) + 1) * k / (2 * n))
            if k == 0:
                assert np.ptp(row) <= 1
            else:
                corr = np.dot(row, basis) / (np.linalg.norm(row) * np.linalg.norm(basis) + 1e-9)
                assert corr > 0.999, f"length {n} k {k}: {corr}"


def _pipeline(w, h, kind, seed, turbo=1, **kw):
    cfg = mav1hip.default_config(w, h, turbo=turbo, qindex=kw.pop("q", 125), **kw)
    fr = mav1hip.turbo_frame(kind, 1) if kind in ("key", "arf", "leaf") else mav1hip.frame_info(order_hint=1, gold_differs=1)
    fp, _ = mav1hip.setup_params(cfg, fr)
    geo, frames = libs.make_sequence(w, h, 3, seed, n_objects=8, max_vel=10)
    r = libs.Impl("ref")
    mvf = mvg = None
    if not fp.frame_is_intra:
        mvf, mvg = r.motion_search(fp, frames[1], frames[0], frames[2])
    if mvg is None:
        mvg = np.zeros(fp.sb_cols * fp.sb_rows * 170, np.int32)
    if mvf is None:
        mvf = np.zeros_like(mvg)
    _, _, blocks, sbs = r.analyze(fp, frames[1], frames[0], frames[2], mvf, mvg)
    if cfg.intra_edge_filter:
        blocks = libs.edge_types(libs.Impl("oracle"), fp, blocks, sbs)
    return fp, frames, blocks, sbs


@needs_ref
@pytest.mark.parametrize("kind", ["key", "arf", "leaf"])
@pytest.mark.parametrize("w,h,seed", [(352, 288, 3), (200, 136, 9), (1920, 1080, 5)])
def test_recon_turbo(kind, w, h, seed):
    fp, frames, blocks, sbs = _pipeline(w, h, kind, seed)
    mats = libs.ref_mats()
    a = libs.Impl("oracle").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    b = libs.Impl("ref").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    assert np.array_equal(a[0], b[0]), "reconstruction differs"
    assert np.array_equal(a[1], b[1]), "coefficients differ"
    assert np.array_equal(a[2], b[2]), "block outputs differ"


@needs_ref
@pytest.mark.parametrize("variant", ["full_rd", "key_full", "lossless", "q_low", "no_edge"])
def test_recon_other_presets(variant):
    w, h = 136, 104
    if variant == "full_rd":
        fp, frames, blocks, sbs = _pipeline(w, h, "inter", 7, turbo=0, multi_ref=1)
    elif variant == "key_full":
        fp, frames, blocks, sbs = _pipeline(w, h, "key", 7, turbo=0)
    elif variant == "lossless":
        fp, frames, blocks, sbs = _pipeline(w, h, "key", 7, turbo=0, q=0)
    elif variant == "q_low":
        fp, frames, blocks, sbs = _pipeline(w, h, "arf", 7, q=30)
    else:
        fp, frames, blocks, sbs = _pipeline(w, h, "key", 7, turbo=0, intra_edge_filter=0)
    mats = libs.ref_mats()
    a = libs.Impl("oracle").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    b = libs.Impl("ref").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    assert np.array_equal(a[0], b[0]), "reconstruction differs"
    assert np.array_equal(a[1], b[1]), "coefficients differ"
    assert np.array_equal(a[2], b[2]), "block outputs differ"


def test_edge_types_known_layout():
    """Hand-built layout: an intra SMOOTH block above and a SMOOTH_V block on
    the left of a block set its luma flag; chroma follows the uv modes; the
    left neighbour does not count across a tile column boundary."""
    w, h = 256, 64
    cfg = mav1hip.default_config(w, h, turbo=1, tile_cols_log2=2)
    fp, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("key", 0))
    assert fp.tile_w_sb == 1
    nsb = fp.sb_cols * fp.sb_rows
    blocks = np.zeros((nsb * 64, 27), np.int32)
    sbs = np.zeros((nsb, 89), np.int32)
    # SB 0: 4 blocks of 32x32; SB 1: one 64x64 block (another tile column)
    B32, B64 = 9, 12
    layout = {0: [(0, 0, B32, 9, 0), (0, 8, B32, 0, 10), (8, 0, B32, 0, 0), (8, 8, B32, 10, 9)],
              1: [(0, 16, B64, 0, 0)], 2: [(0, 32, B64, 0, 0)], 3: [(0, 48, B64, 9, 0)]}
    for s in range(nsb):
        bl = layout.get(s, [])
        sbs[s, :4] = (s * 64, len(bl), 0, s)
        for k, (r, c, bs, ym, uvm) in enumerate(bl):
            blocks[s * 64 + k, :8] = (r, c, bs, 0, ym, uvm, 0, 0)
    out = libs.edge_types(libs.Impl("oracle"), fp, blocks, sbs)
    f = out[:, 22]  # interp_filter
    # block (0,8): left neighbour SMOOTH luma (bit 4); its uv SMOOTH_V is its own mode
    assert f[1] == 0x10
    # block (8,0): above is SMOOTH luma; above's uv DC; chroma above uses (r-1, c+1): block (0,0) uv DC
    assert f[2] == 0x10
    # block (8,8): above (0,8) luma DC, uv SMOOTH_V -> chroma flag; left (8,0) luma DC
    assert f[3] == 0x20
    # SB 1 block: left neighbour (0,15) belongs to the other tile column: not available
    assert f[64] == 0


@needs_ref
def test_forward_matrices_equal_reference():
    """The backend's own FP64 derivation of the Q20 forward matrices equals
    the reference's fwd_build_matrices entry for entry."""
    mine = mav1hip.forward_matrices()
    assert np.array_equal(mine, libs.ref_mats())
    assert np.count_nonzero(mine) > 3500


@needs_ref
@pytest.mark.parametrize("kind,w,h,seed", [("key", 352, 288, 11), ("arf", 352, 288, 13)])
def test_recon_forced_intra_modes(kind, w, h, seed):
    """Every intra mode, angle delta and CfL on the reference's own recon."""
    fp, frames, blocks, sbs = _pipeline(w, h, kind, seed)
    blocks, n_cfl, n_dir = libs.force_intra_modes(fp, blocks, sbs, seed)
    assert n_cfl > 0 and n_dir > 0
    blocks = libs.edge_types(libs.Impl("oracle"), fp, blocks, sbs)
    mats = libs.ref_mats()
    a = libs.Impl("oracle").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    b = libs.Impl("ref").recon(fp, blocks, sbs, frames[1], frames[0], frames[2], mats)
    assert np.array_equal(a[0], b[0]), "reconstruction differs"
    assert np.array_equal(a[1], b[1]), "coefficients differ"
    assert np.array_equal(a[2], b[2]), "block outputs differ"
