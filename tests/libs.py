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
