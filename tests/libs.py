"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's
own kernels (oracle/_ref/libmav1ref.so), for the tests only."""
import ctypes as C
import os

import numpy as np

import mav1hip
import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libmav1ref.so")

P = C.c_void_p
FP = C.POINTER(mav1hip.FrameParams)


def _bind(lib, prefix):
    sigs = {
        "qp_entry_bytes": (C.c_long, [FP]),
        "qpel_build": (None, [FP, P, P]),
        "dec_build": (None, [FP, P, P]),
        "motion_search": (None, [FP, P, P, P, P, P]),
        "analyze": (None, [FP, P, P, P, P, P, P, P, P, P]),
        "predict_intra": (None, [P, C.c_int, C.c_int, C.c_int, P, P, C.c_int]),
        "recon": (None, [FP, P, P, P, P, P, P, P, P]),
        "edge_types": (None, [FP, P, P]),
        "itx_2d": (None, [P, P, C.c_int, C.c_int, C.c_int]),
        "fwd_2d": (None, [P, P, C.c_int, C.c_int, P]),
        "quantize": (None, [P, P, C.c_int, C.c_int, C.c_int]),
        "dequant": (None, [P, P, C.c_int, C.c_int, C.c_int]),
        "forward_matrices": (None, [P]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, prefix + name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    return lib


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so is missing (make -C oracle)")
        _oracle = _bind(C.CDLL(ORACLE_PATH), "oracle_")
    return _oracle


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = _bind(C.CDLL(REF_PATH), "ref_")
    return _ref


class Impl:
    """Uniform calls into the oracle ('oracle') or the reference ('ref')."""

    def __init__(self, which):
        self.which = which
        self.lib = oracle() if which == "oracle" else ref()
        self.pre = "oracle_" if which == "oracle" else "ref_"

    def fn(self, name):
        return getattr(self.lib, self.pre + name)

    def motion_search(self, fp, src, ref_, gold=None):
        n = fp.sb_cols * fp.sb_rows * mav1hip.NUM_NODES * 2
        mv = np.zeros(n, np.int32)
        mvg = np.zeros(n, np.int32)
        self.fn("motion_search")(C.byref(fp), mav1hip.ptr(src), mav1hip.ptr(ref_), mav1hip.ptr(gold), mav1hip.ptr(mv),
                                 mav1hip.ptr(mvg))
        return mv, (mvg if fp.golden else None)

    def analyze(self, fp, src, ref_, gold, mvf, mvg):
        nsb = fp.sb_cols * fp.sb_rows
        costs = np.zeros(nsb * mav1hip.NUM_NODES * mav1hip.A_SLOTS, np.int64)
        infos = np.zeros((nsb * mav1hip.NUM_NODES * mav1hip.A_SLOTS, 27), np.int32)
        blocks = np.zeros((nsb * 64, 27), np.int32)
        sbs = np.zeros((nsb, 89), np.int32)
        if mvf is None:
            mvf = np.zeros(nsb * mav1hip.NUM_NODES * 2, np.int32)
        self.fn("analyze")(C.byref(fp), mav1hip.ptr(src), mav1hip.ptr(ref_), mav1hip.ptr(gold), mav1hip.ptr(mvf),
                           mav1hip.ptr(mvg), mav1hip.ptr(costs), mav1hip.ptr(infos), mav1hip.ptr(blocks),
                           mav1hip.ptr(sbs))
        return costs, infos, blocks, sbs

    def predict_intra(self, px, stride, x, y, args, w, h):
        a = np.asarray(args, np.int32)
        dst = np.zeros((h, w), np.uint8)
        self.fn("predict_intra")(mav1hip.ptr(px), stride, x, y, mav1hip.ptr(a), mav1hip.ptr(dst), w)
        return dst

    def recon(self, fp, blocks, sbs, src, ref_, gold, mats, rec=None):
        """Returns (rec, coefs, blocks) after reconstructing the frame."""
        blocks = np.ascontiguousarray(blocks, np.int32).copy()
        sbs = np.ascontiguousarray(sbs, np.int32)
        rec = np.zeros_like(src) if rec is None else rec.copy()
        coefs = np.zeros(fp.sb_cols * fp.sb_rows * 6144, np.int32)
        self.fn("recon")(C.byref(fp), mav1hip.ptr(blocks), mav1hip.ptr(sbs), mav1hip.ptr(src), mav1hip.ptr(rec),
                         mav1hip.ptr(ref_), mav1hip.ptr(gold), mav1hip.ptr(coefs), mav1hip.ptr(mats))
        return rec, coefs, blocks

    def qpel_build(self, fp, ref_):
        out = np.zeros(16 * fp.frame_w * fp.frame_h, np.uint8)
        self.fn("qpel_build")(C.byref(fp), mav1hip.ptr(ref_), mav1hip.ptr(out))
        return out

    def dec_build(self, fp, frame):
        out = np.zeros(((fp.frame_w + 3) >> 2) * ((fp.frame_h + 3) >> 2), np.uint8)
        self.fn("dec_build")(C.byref(fp), mav1hip.ptr(frame), mav1hip.ptr(out))
        return out


def make_sequence(width, height, nframes, seed, **kw):
    """Padded frames of a synthetic sequence."""
    geo = synth.Geometry(width, height)
    s = synth.Synth(width, height, seed, **kw)
    return geo, [synth.pad_frame(geo, s.frame(t)) for t in range(nframes)]


def turbo_params(width, height, kind, q=125, order_hint=1, **cfgkw):
    cfg = mav1hip.default_config(width, height, turbo=1, qindex=q, **cfgkw)
    fr = mav1hip.turbo_frame(kind, order_hint)
    fp, npresets = mav1hip.setup_params(cfg, fr)
    return cfg, fp


# BlockInfo columns (int32) that eval_block defines; eob / coef_off (11..16)
# are filled by pass 2 only.
INFO_COLS = [i for i in range(27) if not 11 <= i <= 16]


def compare_analysis(a, b, nsb):
    """Equal pass-1 costs everywhere, equal decisions where evaluated, equal
    blocks and superblock records."""
    ca, ia, ba, sa = a
    cb, ib, bb, sb = b
    assert np.array_equal(ca, cb), f"pass-1 costs differ at {np.nonzero(ca != cb)[0][:8]}"
    ev = ca >= 0
    assert np.array_equal(ia[ev][:, INFO_COLS], ib[ev][:, INFO_COLS]), "pass-1 decisions differ"
    assert np.array_equal(sa, sb), "superblock records differ"
    for s in range(nsb):
        n = sa[s, 1]
        assert np.array_equal(ba[s * 64:s * 64 + n], bb[s * 64:s * 64 + n]), f"blocks of SB {s} differ"


TXW = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64]
TXH = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16]
ROW_KIND = [0, 0, 1, 1, 0, 1, 1, 1, 1, 2, 0, 2, 1, 2, 1, 2]  # per tx type: 0 DCT, 1 ADST, 2 identity
COL_KIND = [0, 1, 0, 1, 1, 0, 1, 1, 1, 2, 2, 0, 2, 1, 2, 1]


def valid_tx(txsz, t):
    """Transform types a size can carry: ADST up to 16, identity up to 32."""
    def ok(kind, n):
        return kind == 0 or (kind == 1 and n <= 16) or (kind == 2 and n <= 32)
    return ok(ROW_KIND[t], TXW[txsz]) and ok(COL_KIND[t], TXH[txsz])


def ref_mats():
    m = np.zeros(15 * 4096, np.int32)
    ref().ref_forward_matrices(mav1hip.ptr(m))
    return m


def edge_types(impl, fp, blocks, sbs):
    b = np.ascontiguousarray(blocks, np.int32).copy()
    impl.fn("edge_types")(C.byref(fp), mav1hip.ptr(b), mav1hip.ptr(np.ascontiguousarray(sbs, np.int32)))
    return b


def _bind_filters(lib, pre):
    sigs = {
        "loop_filter": (None, [FP, P, P]),
        "cdef": (None, [FP, P, P, C.c_long, P, C.c_int]),
        "lr_apply": (None, [FP, P, P, P, P]),
        "lr_select": (None, [FP, P, P, P, P, P, C.c_int]),
        "filter_info": (None, [FP, P, P, P]),
        "cdef_select": (None, [FP, P, P, C.c_long, C.c_int, P, P]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, pre + name, None)
        if f is not None:
            f.restype = res
            f.argtypes = args


def filter_info(fp, blocks, sbs):
    """The oracle's per-mi filter info bytes."""
    lib = oracle()
    _bind_filters(lib, "oracle_")
    info = np.zeros(fp.mi_rows * fp.mi_cols, np.uint8)
    lib.oracle_filter_info(C.byref(fp), mav1hip.ptr(np.ascontiguousarray(blocks, np.int32)),
                           mav1hip.ptr(np.ascontiguousarray(sbs, np.int32)), mav1hip.ptr(info))
    return info


class Filters:
    """In-loop filter stages of the oracle ('oracle') or the reference ('ref')."""

    def __init__(self, which):
        self.lib = oracle() if which == "oracle" else ref()
        self.pre = "oracle_" if which == "oracle" else "ref_"
        _bind_filters(self.lib, self.pre)

    def fn(self, name):
        return getattr(self.lib, self.pre + name)

    def loop_filter(self, fp, rec, info):
        out = np.ascontiguousarray(rec, np.uint8).copy()
        self.fn("loop_filter")(C.byref(fp), mav1hip.ptr(out), mav1hip.ptr(info))
        return out

    def cdef(self, fp, rec, info, n):
        fb = rec.size
        out = np.zeros(n * fb, np.uint8)
        self.fn("cdef")(C.byref(fp), mav1hip.ptr(rec), mav1hip.ptr(out), fb, mav1hip.ptr(info), n)
        return out.reshape(n, fb)

    def lr_apply(self, fp, cdef, dbk, units):
        out = np.zeros_like(cdef)
        self.fn("lr_apply")(C.byref(fp), mav1hip.ptr(cdef), mav1hip.ptr(dbk), mav1hip.ptr(out),
                            mav1hip.ptr(np.ascontiguousarray(units, np.int32)))
        return out

    def lr_select(self, fp, src, cdef, dbk, lr_test=0):
        """Returns (units int32[n, 12], lr_type[3])."""
        n = sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
        units = np.zeros((n, 12), np.int32)
        tmp = np.zeros_like(cdef)
        f = mav1hip.FrameParams.from_buffer_copy(fp)
        self.fn("lr_select")(C.byref(f), mav1hip.ptr(src), mav1hip.ptr(cdef), mav1hip.ptr(dbk), mav1hip.ptr(tmp),
                             mav1hip.ptr(units), lr_test)
        return units, [f.lr_type[p] for p in range(3)]


def cdef_select(fp, src, cdefbuf, rec):
    """The oracle's per-superblock CDEF choice: (rec, choice)."""
    lib = oracle()
    _bind_filters(lib, "oracle_")
    out = np.ascontiguousarray(rec, np.uint8).copy()
    choice = np.zeros(fp.sb_cols * fp.sb_rows, np.int8)
    lib.oracle_cdef_select(C.byref(fp), mav1hip.ptr(src), mav1hip.ptr(cdefbuf), cdefbuf.shape[1], cdefbuf.shape[0],
                           mav1hip.ptr(out), mav1hip.ptr(choice))
    return out, choice


def force_intra_modes(fp, blocks, sbs, seed):
    """Random luma / chroma intra modes and angles on every intra block (CfL
    on blocks up to 32x32); returns (blocks, CfL count, directional count)."""
    rng = np.random.default_rng(seed)
    blocks = blocks.copy()
    n_cfl = n_dir = 0
    for s in range(fp.sb_cols * fp.sb_rows):
        for k in range(sbs[s, 0], sbs[s, 0] + sbs[s, 1]):
            b = blocks[k]
            if b[3]:
                continue
            ym = int(rng.integers(0, 13))
            b[4], b[6] = ym, int(rng.integers(-3, 4)) if 1 <= ym <= 8 else 0
            um = int(rng.integers(0, 14 if b[2] <= 9 else 13))
            b[5], b[7] = um, int(rng.integers(-3, 4)) if 1 <= um <= 8 else 0
            n_cfl += um == 13
            n_dir += 1 <= ym <= 8
    return blocks, n_cfl, n_dir
