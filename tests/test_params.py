"""Pins the driver arithmetic that every parity test shares: the encoder
configuration adjustments, the frame geometry, the per-frame quantizer
cascade, the frame parameters and the filter plan (mav1hip_config_normalize,
mav1hip_geometry, mav1hip_frame_qindex, mav1hip_setup_params,
mav1hip_filter_plan). The GPU and the oracle both consume these values, so
only a check independent of params.c catches a wrong constant in it.

The expected values below are a separate restatement of the reference in
Python: create_enc (src/encoder.c:75-120), frame_qindex (:585-596),
setup_params (:407-510), run_frame_core's filter decisions (:618-640), and
the leaf savings of alt_code (:1007-1008). The quantizer tables come from a
fixture taken from the reference's own table file
(tests/golden/qlookup_8bit.json, src/spec_tables.h:608, 692), which also
cross-checks the product's generated copy. A few literal values anchor the
restatement itself (hand-derived from the formulas and tables)."""
import ctypes as C
import itertools
import json
import os

import numpy as np
import pytest

import mav1hip

HERE = os.path.dirname(os.path.abspath(__file__))
QL = json.load(open(os.path.join(HERE, "golden", "qlookup_8bit.json")))
DCQ, ACQ = QL["dc"], QL["ac"]

MAX_TILE_WIDTH = 4096   # spec constants (Annex A / 7.x)
MAX_TILE_COLS = 64
MAX_LOOP_FILTER = 63
RESTORATION_TILESIZE_MAX = 256
GOLDEN_FRAME, ALTREF_FRAME = 4, 7
CDEF_PRESETS = [(0, 0, 0, 0), (2, 1, 1, 0), (4, 2, 2, 1), (8, 4, 4, 2)]


def normalize(c):
    """create_enc's adjustments (src/encoder.c:75-92) on a dict config."""
    c = dict(c)
    if c["turbo"]:
        c["fast"] = 1
        c["altref"] = max(c["altref"], 2)
        c["leaf"] = 1
    if c["altref"] == 1:
        c["altref"] = 0
    c["altref"] = min(c["altref"], 16)
    if c["altref"]:
        c["multi_ref"] = c["nonref_odd"] = 0
    if not c["altref"]:
        c["leaf"] = 0
    c["qindex"] = min(max(c["qindex"], 0), 255)
    return c


def geometry(w, h, tile_log2):
    """src/encoder.c:93-120."""
    mi_cols, mi_rows = 2 * ((w + 7) >> 3), 2 * ((h + 7) >> 3)
    sb_cols, sb_rows = (mi_cols + 15) >> 4, (mi_rows + 15) >> 4
    lo = 0
    while ((MAX_TILE_WIDTH >> 6) << lo) < sb_cols:
        lo += 1
    hi = 0
    while (1 << hi) < min(sb_cols, MAX_TILE_COLS):
        hi += 1
    l = min(max(tile_log2, lo), hi)
    tile_w_sb = (sb_cols + (1 << l) - 1) >> l
    stride = [sb_cols * 64, sb_cols * 32, sb_cols * 32]
    plane_h = [sb_rows * 64, sb_rows * 32, sb_rows * 32]
    offset = [0, stride[0] * plane_h[0], stride[0] * plane_h[0] + stride[1] * plane_h[1]]
    return dict(mi_cols=mi_cols, mi_rows=mi_rows, sb_cols=sb_cols, sb_rows=sb_rows, tile_w_sb=tile_w_sb,
                stride=stride, plane_h=plane_h, offset=offset,
                frame_bytes=offset[2] + stride[2] * plane_h[2])


def frame_q(c, fr):
    """frame_qindex (src/encoder.c:585-596), default knob offsets."""
    q = c["qindex"]
    if q == 0:
        return 0
    key = fr["is_key"]
    off = -24 if key else (-8 if fr["order_hint"] % 4 == 0 else 4)
    if not key and fr["nonref"]:
        off = 12
    if not key and c["altref"]:
        off = -8 if fr["frame_kind"] == 1 else 4
    if not key and c["leaf"] and fr["frame_kind"] == 3:
        off = 0
    if not key and c["leaf"] and fr["nonref"]:
        off = 8
    return min(max(q + off, 1), 255)


def params(c, fr, q):
    """setup_params (src/encoder.c:407-510) with the leaf savings of
    alt_code (:1007-1008); returns (fields, presets the filter stage
    evaluates)."""
    g = geometry(c["width"], c["height"], c["tile_cols_log2"])
    key = fr["is_key"]
    leafopt_cfg = c["knob_leafopt"] if c["knob_leafopt"] >= 0 else 0x3F
    nonref = fr["nonref"]
    leafopt = leafopt_cfg if (c["altref"] and nonref) else 0
    p = {}
    p["mi_cols"], p["mi_rows"], p["sb_cols"], p["sb_rows"] = g["mi_cols"], g["mi_rows"], g["sb_cols"], g["sb_rows"]
    p["plane_w"], p["plane_h"], p["stride"], p["offset"] = g["stride"], g["plane_h"], g["stride"], g["offset"]
    p["base_q_idx"] = q
    p["dc_q"] = [DCQ[q]] * 3
    p["ac_q"] = [ACQ[q]] * 3
    p["frame_is_intra"] = key
    p["reduced_tx_set"] = 0
    p["enable_intra_edge_filter"] = c["intra_edge_filter"]
    p["tx_search"] = c["tx_search"]
    p["lossless"] = int(q == 0)
    p["tx_mode_select"] = int(q != 0 and bool(c["tx_depth"]))
    p["switchable_interp"] = int(not key and bool(c["switchable_interp"]))
    p["golden"] = int(not key and bool(c["multi_ref"]) and bool(fr["gold_differs"]))
    p["ref2_frame"] = GOLDEN_FRAME
    p["compound"] = 0
    if c["altref"]:
        p["golden"] = int(not key and fr["frame_kind"] >= 2)
        p["ref2_frame"] = ALTREF_FRAME
        p["compound"] = fr["reference_select"]
    p["fast"] = c["fast"]
    p["turbo"] = (c["knob_turbo_mask"] if c["knob_turbo_mask"] >= 0 else 0x47CBF1) if c["turbo"] else 0
    p["leaf"] = (leafopt & 1) if c["leaf"] else nonref
    if leafopt & 16:
        p["tx_search"] = 0
    dt = c["knob_dt"] if c["knob_dt"] >= 0 else (320 if g["sb_rows"] >= 16 else 352)
    bias = c["knob_dt_bias"] if c["knob_dt_bias"] >= 0 else 0
    p["depth_thr"] = (dt * (q + bias)) // (128 + bias) if c["turbo"] else 0
    p["tile_w_sb"] = g["tile_w_sb"]
    acq = ACQ[q]
    p["lambda_"] = acq * acq * 256 // 430
    p["frame_w"], p["frame_h"] = c["width"], c["height"]
    p["me_range"] = c["me_range"]
    if (leafopt & 32) and p["me_range"] > 16:
        p["me_range"] = 16
    lvl = 0
    lf_off = (leafopt & 2) if c["leaf"] else nonref
    if c["loop_filter"] and q > 0 and not lf_off:
        lvl = (acq * 20723 + 1015158 + (1 << 17)) >> 18
        if not key:
            lvl -= 4
        lvl = min(max(lvl, 0), MAX_LOOP_FILTER)
    p["lf_level"] = [lvl] * 4
    p["lf_sharpness"] = 0
    p["cdef_damping"] = 3 + (q >> 6)
    p["cdef_bits"] = 2
    if leafopt & 4:
        p["cdef_damping"], p["cdef_bits"] = 3, 0
    pi = [0, 1, 2, 3]
    if (p["turbo"] & 2048) and p["cdef_bits"] == 2:
        p["cdef_bits"] = 1
        pi[1] = c["knob_cdef_preset"] if c["knob_cdef_preset"] >= 0 else 2
    p["cdef_y_pri"] = [CDEF_PRESETS[pi[i]][0] for i in range(4)] + [0] * 4
    p["cdef_y_sec"] = [CDEF_PRESETS[pi[i]][1] for i in range(4)] + [0] * 4
    p["cdef_uv_pri"] = [CDEF_PRESETS[pi[i]][2] for i in range(4)] + [0] * 4
    p["cdef_uv_sec"] = [CDEF_PRESETS[pi[i]][3] for i in range(4)] + [0] * 4
    size = RESTORATION_TILESIZE_MAX >> 2
    off = 0
    p["lr_type"], p["lr_unit_size"], p["lr_unit_rows"], p["lr_unit_cols"], p["lr_unit_off"] = [], [], [], [], []
    for pl in range(3):
        sub = pl > 0
        pw, ph = (c["width"] + sub) >> sub, (c["height"] + sub) >> sub
        rows, cols = max((ph + (size >> 1)) // size, 1), max((pw + (size >> 1)) // size, 1)
        p["lr_type"].append(0)
        p["lr_unit_size"].append(size)
        p["lr_unit_rows"].append(rows)
        p["lr_unit_cols"].append(cols)
        p["lr_unit_off"].append(off)
        off += rows * cols
    p["wave"], p["num_blocks"] = 0, 0
    npre = (1 << p["cdef_bits"]) if c["cdef"] else 0
    if leafopt & 4:
        npre = 0
    return p, npre


def plan(c, fr, p, npre):
    """run_frame_core's filter decisions (src/encoder.c:618-640)."""
    leafopt_cfg = c["knob_leafopt"] if c["knob_leafopt"] >= 0 else 0x3F
    leafopt = leafopt_cfg if (c["altref"] and fr["nonref"]) else 0
    cdef_off = (leafopt & 4) != 0
    lf_on = any(p["lf_level"])
    filt = lf_on or (npre and not cdef_off) or not c["turbo"]
    return dict(filter_info=int(bool(filt)), deblock=int(bool(filt and lf_on)),
                cdef_sets=npre if (filt and not cdef_off) else 0,
                cdef_select=int(bool(c["cdef"] and p["base_q_idx"] > 0 and c["qindex"] > 0 and not cdef_off
                                     and npre > 0)),
                restoration=int(bool(c["restoration"] and c["qindex"] > 0 and not (leafopt & 8))))


def default_dict(w, h, turbo):
    """mav1_default_config (src/encoder.c:51-66) plus the CLI presets
    (tools/mav1enc.c:57, 98-100): --turbo sets fast, 2 tile columns in log2,
    and fast turns off tx search, tx depth, restoration and switchable
    filters."""
    c = dict(width=w, height=h, qindex=100, keyint=0, tx_search=1, tx_depth=1, loop_filter=1, cdef=1, restoration=1,
             tile_cols_log2=2 if turbo else 0, intra_edge_filter=1, nonref_odd=0, fast=turbo, multi_ref=0, altref=0,
             leaf=0, switchable_interp=1, me_range=32, turbo=turbo, knob_turbo_mask=-1, knob_dt=-1, knob_dt_bias=-1,
             knob_cdef_preset=-1, knob_leafopt=-1)
    if turbo:
        c.update(tx_search=0, tx_depth=0, restoration=0, switchable_interp=0)
    return c


def to_cfg(c):
    cfg = mav1hip.Config()
    for k, v in c.items():
        setattr(cfg, k, v)
    return cfg


FRAMES = {
    "key": dict(is_key=1, frame_kind=0, nonref=0, order_hint=0, reference_select=0, gold_differs=0),
    "arf": dict(is_key=0, frame_kind=1, nonref=0, order_hint=2, reference_select=0, gold_differs=0),
    "leaf": dict(is_key=0, frame_kind=2, nonref=1, order_hint=1, reference_select=1, gold_differs=0),
    "regular": dict(is_key=0, frame_kind=0, nonref=0, order_hint=3, reference_select=0, gold_differs=1),
    "regular4": dict(is_key=0, frame_kind=0, nonref=0, order_hint=4, reference_select=0, gold_differs=1),
    "middle": dict(is_key=0, frame_kind=3, nonref=0, order_hint=8, reference_select=1, gold_differs=0),
    "nonref": dict(is_key=0, frame_kind=0, nonref=1, order_hint=5, reference_select=0, gold_differs=0),
}
SIZES = [(352, 288), (854, 480), (1280, 720), (1920, 1080), (3840, 2160), (200, 136), (204, 116), (17, 33),
         (16, 16), (4100, 64)]


def to_frame(f):
    return mav1hip.frame_info(**f)


def fp_dict(fp):
    out = {}
    for name, _ in mav1hip.FrameParams._fields_:
        v = getattr(fp, name)
        out[name] = list(v) if not isinstance(v, int) else v
    return out


def check(c, fr, q=None):
    cn = normalize(c)
    cfg = to_cfg(c)
    mav1hip.lib().mav1hip_config_normalize(C.byref(cfg))
    got_cfg = {k: getattr(cfg, k) for k in c}
    assert got_cfg == cn, f"normalize: {got_cfg} != {cn}"
    f = to_frame(fr)
    qx = frame_q(cn, fr) if q is None else q
    assert mav1hip.lib().mav1hip_frame_qindex(C.byref(cfg), C.byref(f)) == frame_q(cn, fr)
    fp, npre = mav1hip.setup_params(cfg, f, qx)
    exp, enpre = params(cn, fr, qx)
    got = fp_dict(fp)
    bad = {k: (got[k], exp[k]) for k in exp if got[k] != exp[k]}
    assert not bad, f"{c['width']}x{c['height']} q{qx} {fr}: field (got, expected): {bad}"
    assert npre == enpre
    pl = mav1hip.filter_plan(cfg, f, fp, npre)
    gp = {k: getattr(pl, k) for k, _ in mav1hip.FilterPlan._fields_}
    assert gp == plan(cn, fr, exp, enpre), (gp, plan(cn, fr, exp, enpre))
    return exp


def test_fixture_matches_product_tables():
    """The fixture (the reference's table file) and the product's generated
    tables agree at every q (through ac_q / dc_q of the parameters)."""
    cfg = mav1hip.default_config(352, 288, turbo=1)
    for q in range(256):
        fp, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("key", 0), q)
        assert fp.dc_q[0] == DCQ[q] and fp.ac_q[0] == ACQ[q], q


def test_literal_anchors():
    """Hand-derived values of the formulas at known points: AV1's 8-bit
    tables end at dc 1336 / ac 1828 and start at 4; q 117 (the 1080p turbo
    ALTREF at -q 125) gives ac 146 (hence lambda = 146^2 * 256 / 430 =
    12690, LF level (146*20723 + 1015158 + 131072) >> 18 = 15, minus 4 on
    an inter frame = 11, depth_thr 320 * 117 / 128 = 292, CDEF damping
    3 + (117 >> 6) = 4)."""
    assert DCQ[0] == 4 and ACQ[0] == 4 and DCQ[255] == 1336 and ACQ[255] == 1828
    assert ACQ[117] == 146
    cfg = mav1hip.default_config(1920, 1080, turbo=1, qindex=125)
    fp, npre = mav1hip.setup_params(cfg, mav1hip.turbo_frame("arf", 2))
    assert fp.base_q_idx == 117
    assert fp.lambda_ == 12690
    assert list(fp.lf_level) == [11] * 4
    assert fp.depth_thr == 292
    assert fp.cdef_damping == 4 and fp.cdef_bits == 1 and npre == 2
    assert list(fp.cdef_y_pri[:2]) == [0, 4] and list(fp.cdef_uv_sec[:2]) == [0, 1]
    assert fp.tile_w_sb == 8 and fp.sb_cols == 30 and fp.sb_rows == 17
    assert list(fp.lr_unit_rows) == [17, 8, 8] and list(fp.lr_unit_cols) == [30, 15, 15]
    fpl, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("leaf", 1))
    assert fpl.base_q_idx == 133 and list(fpl.lf_level) == [0] * 4 and fpl.me_range == 16 and fpl.leaf == 1
    fpk, _ = mav1hip.setup_params(cfg, mav1hip.turbo_frame("key", 0))
    assert fpk.base_q_idx == 101 and fpk.frame_is_intra == 1


@pytest.mark.parametrize("w,h", SIZES)
def test_geometry(w, h):
    for turbo, tiles in itertools.product((0, 1), (0, 1, 2, 3, 6)):
        c = default_dict(w, h, turbo)
        c["tile_cols_log2"] = tiles
        cfg = to_cfg(normalize(c))
        out = [C.c_int32() for _ in range(5)]
        st, of, ph = (C.c_int32 * 3)(), (C.c_int32 * 3)(), (C.c_int32 * 3)()
        fb = mav1hip.lib().mav1hip_geometry(C.byref(cfg), *[C.byref(o) for o in out], st, of, ph)
        g = geometry(w, h, tiles)
        assert fb == g["frame_bytes"]
        assert [o.value for o in out] == [g["mi_cols"], g["mi_rows"], g["sb_cols"], g["sb_rows"], g["tile_w_sb"]]
        assert list(st) == g["stride"] and list(of) == g["offset"] and list(ph) == g["plane_h"]


def test_every_q_turbo_kinds():
    """q 0..255 for every frame kind of the turbo GOP at 1080p."""
    for kind in ("key", "arf", "leaf", "regular"):
        for q in range(256):
            c = default_dict(1920, 1080, 1)
            c["qindex"] = q
            check(c, FRAMES[kind])
            c["restoration"] = 1
            check(c, FRAMES[kind])


@pytest.mark.parametrize("w,h", SIZES)
def test_sizes_and_configs(w, h):
    rng = np.random.default_rng(w * 7 + h)
    for q in [0, 1, 20, 63, 64, 80, 125, 127, 128, 170, 191, 192, 215, 255] + list(rng.integers(0, 256, 6)):
        for kind in FRAMES:
            for turbo in (0, 1):
                c = default_dict(w, h, turbo)
                c["qindex"] = int(q)
                check(c, FRAMES[kind])
                # full RD: altref 16 pyramid, the CLI's --golden (a no-op with altref)
                c2 = dict(c, altref=16, multi_ref=1, leaf=1, restoration=1, tx_search=1, tx_depth=1,
                          switchable_interp=1)
                check(c2, FRAMES[kind])
                # low delay with GOLDEN: no altref, multi_ref, the reference's odd non-reference frames
                c3 = dict(c, turbo=0, fast=0, altref=0, multi_ref=1, nonref_odd=1)
                check(c3, FRAMES[kind])


def test_knobs_and_switches():
    """The reference's testing knobs and the per-tool switches."""
    rng = np.random.default_rng(3)
    for _ in range(400):
        w, h = SIZES[int(rng.integers(0, len(SIZES)))]
        c = default_dict(w, h, int(rng.integers(0, 2)))
        c.update(qindex=int(rng.integers(-5, 260)), loop_filter=int(rng.integers(0, 2)), cdef=int(rng.integers(0, 2)),
                 restoration=int(rng.integers(0, 2)), tx_depth=int(rng.integers(0, 2)),
                 tx_search=int(rng.integers(0, 2)), intra_edge_filter=int(rng.integers(0, 2)),
                 switchable_interp=int(rng.integers(0, 2)), altref=int(rng.integers(0, 20)),
                 leaf=int(rng.integers(0, 2)), multi_ref=int(rng.integers(0, 2)), me_range=int(rng.integers(4, 64)),
                 knob_turbo_mask=int(rng.choice([-1, 0x47CBF1, 0x47C3F1, 0x7FFFFF, 0])),
                 knob_dt=int(rng.choice([-1, 0, 200, 320, 352, 500])), knob_dt_bias=int(rng.choice([-1, 0, 16, 64])),
                 knob_cdef_preset=int(rng.choice([-1, 0, 1, 2, 3])),
                 knob_leafopt=int(rng.choice([-1, 0, 1, 2, 4, 8, 16, 32, 0x3F, 0x15])))
        kind = list(FRAMES)[int(rng.integers(0, len(FRAMES)))]
        check(c, FRAMES[kind])


def test_a_wrong_constant_is_caught():
    """The restatement is sensitive to the constants it pins: perturbing one
    formula in the expected values makes the comparison fail."""
    c = default_dict(1920, 1080, 1)
    c["qindex"] = 125
    fr = FRAMES["arf"]
    exp, _ = params(normalize(c), fr, 117)
    fp, _ = mav1hip.setup_params(to_cfg(normalize(c)), to_frame(fr), 117)
    got = fp_dict(fp)
    assert got["lambda_"] == exp["lambda_"]
    assert got["lambda_"] != ACQ[117] * ACQ[117] * 256 // 431
    assert got["lf_level"][0] != ((ACQ[117] * 20723 + 1015158 + (1 << 17)) >> 18)  # the inter -4 applies
    assert got["cdef_damping"] != 3 + (117 >> 5)
