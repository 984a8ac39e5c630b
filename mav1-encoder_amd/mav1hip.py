"""Python binding of libmav1hip.so (include/mav1_hip.h) over ctypes.

The records are the ABI structs of include/mav1_types.h; frames and
motion fields are numpy arrays. Loading fails loudly when the library has
not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmav1hip.so")

NUM_NODES = 85
A_SLOTS = 13
TURBO_DEFAULT = 0x47CBF1
(BUF_SRC, BUF_REF, BUF_GOLD, BUF_MVFIELD, BUF_MVGOLD, BUF_PLANES, BUF_ACOSTS, BUF_AINFOS, BUF_BLOCKS,
 BUF_SBS, BUF_REC, BUF_COEFS, BUF_LFMASK, BUF_CDEFBUF, BUF_DBK, BUF_LROUT, BUF_LRUNITS, BUF_CDEF_CHOICE,
 BUF_DECGRID) = range(19)
MAX_CDEF_SETS = 8
STATIC_LR_UNITS = 3072  # MAV1HIP_STATIC_LR_UNITS


class FrameParams(C.Structure):
    _fields_ = [
        ("mi_cols", C.c_int32), ("mi_rows", C.c_int32), ("sb_cols", C.c_int32), ("sb_rows", C.c_int32),
        ("plane_w", C.c_int32 * 3), ("plane_h", C.c_int32 * 3), ("stride", C.c_int32 * 3),
        ("offset", C.c_int32 * 3), ("base_q_idx", C.c_int32), ("dc_q", C.c_int32 * 3), ("ac_q", C.c_int32 * 3),
        ("frame_is_intra", C.c_int32), ("reduced_tx_set", C.c_int32), ("enable_intra_edge_filter", C.c_int32),
        ("tx_search", C.c_int32), ("lambda_", C.c_int32), ("wave", C.c_int32), ("num_blocks", C.c_int32),
        ("frame_w", C.c_int32), ("frame_h", C.c_int32), ("me_range", C.c_int32), ("lossless", C.c_int32),
        ("tx_mode_select", C.c_int32), ("tile_w_sb", C.c_int32), ("fast", C.c_int32), ("leaf", C.c_int32),
        ("depth_thr", C.c_int32), ("switchable_interp", C.c_int32), ("golden", C.c_int32),
        ("ref2_frame", C.c_int32), ("compound", C.c_int32), ("lf_level", C.c_int32 * 4),
        ("lf_sharpness", C.c_int32), ("cdef_damping", C.c_int32), ("cdef_bits", C.c_int32),
        ("cdef_y_pri", C.c_int32 * 8), ("cdef_y_sec", C.c_int32 * 8), ("cdef_uv_pri", C.c_int32 * 8),
        ("cdef_uv_sec", C.c_int32 * 8), ("lr_type", C.c_int32 * 3), ("lr_unit_size", C.c_int32 * 3),
        ("lr_unit_rows", C.c_int32 * 3), ("lr_unit_cols", C.c_int32 * 3), ("lr_unit_off", C.c_int32 * 3),
        ("turbo", C.c_int32),
    ]


assert C.sizeof(FrameParams) == 392


class BlockInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "mi_row", "mi_col", "bsize", "is_inter", "y_mode", "uv_mode", "angle_y", "angle_uv", "mv_row", "mv_col",
        "tx_type")] + [("eob", C.c_int32 * 3), ("coef_off", C.c_int32 * 3)] + [(n, C.c_int32) for n in (
            "cfl_alpha_u", "cfl_alpha_v", "tx_depth", "tx_types_lo", "tx_types_hi", "interp_filter", "ref_frame",
            "ref_frame1", "mv1_row", "mv1_col")]


assert C.sizeof(BlockInfo) == 108


class SbInfo(C.Structure):
    _fields_ = [("first_block", C.c_int32), ("num_blocks", C.c_int32), ("sb_row", C.c_int32),
                ("sb_col", C.c_int32), ("part", C.c_int32 * 85)]


assert C.sizeof(SbInfo) == 356


class Config(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "width", "height", "qindex", "keyint", "tx_search", "tx_depth", "loop_filter", "cdef", "restoration",
        "tile_cols_log2", "intra_edge_filter", "nonref_odd", "fast", "multi_ref", "altref", "leaf",
        "switchable_interp", "me_range", "turbo", "knob_turbo_mask", "knob_dt", "knob_dt_bias", "knob_cdef_preset",
        "knob_leafopt")]


class Frame(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "is_key", "frame_kind", "nonref", "order_hint", "reference_select", "gold_differs")]


class FilterPlan(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("filter_info", "deblock", "cdef_sets", "cdef_select", "restoration")]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32 = C.c_int32
    sig = {
        "mav1hip_config_default": (None, [C.POINTER(Config), i32, i32, i32]),
        "mav1hip_config_normalize": (None, [C.POINTER(Config)]),
        "mav1hip_geometry": (C.c_size_t, [C.POINTER(Config)] + [C.POINTER(i32)] * 5 + [C.POINTER(i32)] * 3),
        "mav1hip_frame_qindex": (i32, [C.POINTER(Config), C.POINTER(Frame)]),
        "mav1hip_setup_params": (i32, [C.POINTER(Config), C.POINTER(Frame), i32, C.POINTER(FrameParams)]),
        "mav1hip_forward_matrices": (None, [P]),
        "mav1hip_create": (P, [C.POINTER(Config), C.c_char_p, C.c_size_t]),
        "mav1hip_destroy": (None, [P]),
        "mav1hip_error": (C.c_char_p, [P]),
        "mav1hip_device_name": (C.c_char_p, [P]),
        "mav1hip_motion_search": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, P, P]),
        "mav1hip_buffer_bytes": (C.c_size_t, [P, i32]),
        "mav1hip_planes_entry_bytes": (C.c_size_t, [P]),
        "mav1hip_upload": (C.c_int, [P, i32, P, C.c_size_t]),
        "mav1hip_download": (C.c_int, [P, i32, P, C.c_size_t]),
        "mav1hip_set_params": (C.c_int, [P, C.POINTER(FrameParams)]),
        "mav1hip_launch_motion_search": (C.c_int, [P]),
        "mav1hip_analyze": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, P, P, P, P, P, P]),
        "mav1hip_launch_analyze": (C.c_int, [P]),
        "mav1hip_edge_types": (C.c_int, [P, C.POINTER(FrameParams), P, P]),
        "mav1hip_recon": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, P, P, P, P]),
        "mav1hip_launch_edge_types": (C.c_int, [P]),
        "mav1hip_launch_recon": (C.c_int, [P]),
        "mav1hip_sync": (C.c_int, [P]),
        "mav1hip_pool_alloc": (C.c_int, [P, i32]),
        "mav1hip_pool_upload": (C.c_int, [P, i32, P]),
        "mav1hip_bind_frames": (C.c_int, [P, i32, i32, i32]),
        "mav1hip_pool_store": (C.c_int, [P, i32, i32]),
        "mav1hip_pool_import": (C.c_int, [P, i32, P, i32]),
        "mav1hip_timing_enable": (C.c_int, [P, i32]),
        "mav1hip_timing_select": (C.c_int, [P, C.c_char_p]),
        "mav1hip_timing_reset": (C.c_int, [P]),
        "mav1hip_timing_get": (C.c_int, [P, C.c_char_p, C.POINTER(i32), C.POINTER(C.c_double)]),
        "mav1hip_filter_plan": (i32, [C.POINTER(Config), C.POINTER(Frame), C.POINTER(FrameParams), i32,
                                      C.POINTER(FilterPlan)]),
        "mav1hip_filter_info": (C.c_int, [P, C.POINTER(FrameParams), P, P, P]),
        "mav1hip_filter": (C.c_int, [P, C.POINTER(FrameParams), P, P, i32, P]),
        "mav1hip_cdef_select": (C.c_int, [P, C.POINTER(FrameParams), P, P, i32, P, P]),
        "mav1hip_restore": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, P]),
        "mav1hip_launch_filter_info": (C.c_int, [P]),
        "mav1hip_launch_deblock": (C.c_int, [P]),
        "mav1hip_launch_cdef": (C.c_int, [P, i32]),
        "mav1hip_launch_save_deblocked": (C.c_int, [P]),
        "mav1hip_launch_cdef_select": (C.c_int, [P, i32]),
        "mav1hip_launch_restore": (C.c_int, [P]),
        "mav1hip_launch_lr_select": (C.c_int, [P, i32]),
        "mav1hip_decision_record_bytes": (C.c_size_t, [C.POINTER(FrameParams)]),
        "mav1hip_decision_record": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, P]),
        "mav1hip_launch_decision_grid": (C.c_int, [P]),
        "mav1hip_lr_select": (C.c_int, [P, C.POINTER(FrameParams), P, P, P, i32, P, P, P]),
        "mav1hip_seq_create": (P, [C.POINTER(Config), i32, i32, C.c_char_p, C.c_size_t]),
        "mav1hip_seq_destroy": (None, [P]),
        "mav1hip_seq_error": (C.c_char_p, [P]),
        "mav1hip_seq_send_frame": (C.c_int, [P, P, P]),
        "mav1hip_seq_send_device_frame": (C.c_int, [P, P]),
        "mav1hip_seq_receive": (C.c_int, [P, P, i32]),
        "mav1hip_seq_engine": (P, [P, i32]),
        "mav1hip_batch_create": (P, [P, i32, i32, i32, C.c_char_p, C.c_size_t]),
        "mav1hip_batch_destroy": (None, [P]),
        "mav1hip_batch_error": (C.c_char_p, [P]),
        "mav1hip_batch_send_frame": (C.c_int, [P, i32, P, P]),
        "mav1hip_batch_send_device_frame": (C.c_int, [P, i32, P]),
        "mav1hip_batch_receive": (C.c_int, [P, i32, P, i32]),
        "mav1hip_batch_stream": (P, [P, i32]),
        "mav1hip_batch_set_knob": (C.c_int, [P, C.c_char_p, i32]),
        "mav1hip_batch_set_pruner": (C.c_int, [P, P, C.c_float]),
        "mav1hip_pruner_num_params": (i32, []),
        "mav1hip_pruner_create": (P, [i32, C.c_char_p, C.c_size_t]),
        "mav1hip_pruner_destroy": (None, [P]),
        "mav1hip_pruner_error": (C.c_char_p, [P]),
        "mav1hip_pruner_set_params": (C.c_int, [P, P]),
        "mav1hip_pruner_get_params": (C.c_int, [P, P]),
        "mav1hip_pruner_load": (C.c_int, [P, P, P, P, i32]),
        "mav1hip_pruner_forward": (C.c_int, [P, P, P]),
        "mav1hip_pruner_backward": (C.c_int, [P, P]),
        "mav1hip_pruner_adam": (C.c_int, [P, C.c_float, C.c_float, C.c_float, C.c_float, i32]),
        "mav1hip_pruner_sync": (C.c_int, [P]),
        "mav1hip_y4m_open": (P, [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                 C.c_char_p, C.c_size_t]),
        "mav1hip_y4m_read": (C.c_int, [P, P, P]),
        "mav1hip_y4m_close": (None, [P]),
        "mav1hip_psnr": (C.c_double, [P, P, C.c_size_t]),
        "mav1hip_lr_select_static": (C.c_int, [C.POINTER(FrameParams), P, P, P, i32, P, P, P, C.c_char_p,
                                               C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lr_select_static(fp, d_src, d_cdef, d_dbk, lr_test, d_units, d_lr_type, stream=None):
    """mav1hip_lr_select_static: the restoration search on device buffers
    (integer device addresses) with the library's static scratch, enqueued
    on `stream` (an integer hipStream_t, None: the null stream)."""
    err = C.create_string_buffer(512)
    rc = lib().mav1hip_lr_select_static(C.byref(fp), C.c_void_p(d_src), C.c_void_p(d_cdef), C.c_void_p(d_dbk),
                                        lr_test, C.c_void_p(d_units), C.c_void_p(d_lr_type),
                                        C.c_void_p(stream) if stream else None, err, len(err))
    if rc != 0:
        raise RuntimeError(f"lr_select_static failed ({rc}): {err.value.decode(errors='replace')}")


def default_config(width, height, turbo=1, **kw):
    cfg = Config()
    lib().mav1hip_config_default(C.byref(cfg), width, height, turbo)
    for k, v in kw.items():
        setattr(cfg, k, v)
    lib().mav1hip_config_normalize(C.byref(cfg))
    return cfg


def frame_info(is_key=0, frame_kind=0, nonref=0, order_hint=0, reference_select=0, gold_differs=0):
    return Frame(is_key, frame_kind, nonref, order_hint, reference_select, gold_differs)


def turbo_frame(kind, order_hint):
    """Frame state of the turbo mini-GOP of 2 (altref 2, leaf): kind 'key',
    'arf' (hidden ALTREF, LAST only) or 'leaf' (non-reference, LAST + ALTREF
    compound) (src/encoder.c:1004-1030, 1252-1275)."""
    if kind == "key":
        return frame_info(is_key=1, order_hint=order_hint)
    if kind == "arf":
        return frame_info(frame_kind=1, order_hint=order_hint)
    if kind == "leaf":
        return frame_info(frame_kind=2, nonref=1, order_hint=order_hint, reference_select=1)
    raise ValueError(kind)


def setup_params(cfg, fr, q=None):
    if q is None:
        q = lib().mav1hip_frame_qindex(C.byref(cfg), C.byref(fr))
    fp = FrameParams()
    n = lib().mav1hip_setup_params(C.byref(cfg), C.byref(fr), q, C.byref(fp))
    return fp, n


def filter_plan(cfg, fr, fp, npresets):
    plan = FilterPlan()
    lib().mav1hip_filter_plan(C.byref(cfg), C.byref(fr), C.byref(fp), npresets, C.byref(plan))
    return plan


def forward_matrices():
    m = np.zeros(15 * 4096, np.int32)
    lib().mav1hip_forward_matrices(ptr(m))
    return m


class HipError(RuntimeError):
    pass


def psnr(a, b):
    """mav1hip_psnr over two equally sized uint8 arrays."""
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    if a.size != b.size:
        raise ValueError("psnr: sizes differ")
    return lib().mav1hip_psnr(ptr(a), ptr(b), a.size)


class Y4mReader:
    """mav1hip_y4m_*: I420 frames of a y4m file as (Y, U, V) arrays."""

    def __init__(self, path):
        w, h, fn, fd = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        err = C.create_string_buffer(256)
        self.h_ = lib().mav1hip_y4m_open(path.encode(), C.byref(w), C.byref(h), C.byref(fn), C.byref(fd), err, 256)
        if not self.h_:
            raise HipError(err.value.decode())
        self.width, self.height, self.fps = w.value, h.value, (fn.value, fd.value)

    def read(self):
        cw, ch = (self.width + 1) >> 1, (self.height + 1) >> 1
        pl = [np.empty((self.height, self.width), np.uint8), np.empty((ch, cw), np.uint8), np.empty((ch, cw), np.uint8)]
        arr = (C.c_void_p * 3)(*[p.ctypes.data for p in pl])
        st = (C.c_int32 * 3)(self.width, cw, cw)
        rc = lib().mav1hip_y4m_read(self.h_, arr, st)
        if rc == 1:
            return None
        if rc:
            raise HipError("y4m: truncated or malformed frame")
        return tuple(pl)

    def close(self):
        if self.h_:
            lib().mav1hip_y4m_close(self.h_)
            self.h_ = None


EBUSY = -5


class CodedFrame(C.Structure):
    """Mav1HipCodedFrame: one coded frame of the sequence driver."""
    _fields_ = [(n, C.c_int32) for n in (
        "coded_index", "disp", "is_key", "frame_kind", "nonref", "hidden", "order_hint", "refresh_frame_flags",
        "last_slot", "arf_slot", "last_hint", "arf_hint", "engine")] + [
        ("fp", FrameParams), ("plan", FilterPlan), ("blocks", C.c_void_p), ("sbs", C.c_void_p),
        ("coefs", C.c_void_p), ("cdef_choice", C.c_void_p), ("lr_units", C.c_void_p), ("recon", C.c_void_p),
        ("decision", C.c_void_p), ("decision_bytes", C.c_size_t)]


def _host_array(p, dtype, count):
    if not p:
        return None
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), (count,)).copy()


class Sequence:
    """The sequence driver (mav1hip_seq_*): frames in, in display order;
    coded frames out, in coding order, with the entropy coder's inputs."""

    def __init__(self, cfg, lr_test=0, dump_decisions=False):
        err = C.create_string_buffer(512)
        h = lib().mav1hip_seq_create(C.byref(cfg), lr_test, int(bool(dump_decisions)), err, 512)
        if not h:
            raise HipError(err.value.decode())
        self.h = h
        self.cfg = cfg

    def close(self):
        if self.h:
            lib().mav1hip_seq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, what, rc):
        return HipError(f"{what}: {lib().mav1hip_seq_error(self.h).decode()} (rc {rc})")

    def engine(self, i):
        """Engine i's context (owned by the driver), for the timing calls."""
        h = lib().mav1hip_seq_engine(self.h, i)
        if not h:
            raise HipError(f"no engine {i}")
        ctx = Context.__new__(Context)
        ctx.h, ctx.cfg, ctx._borrowed = h, self.cfg, True
        ctx.frame_bytes = lib().mav1hip_buffer_bytes(h, BUF_SRC)
        ctx.mv_ints = lib().mav1hip_buffer_bytes(h, BUF_MVFIELD) // 4
        return ctx

    def send(self, planes):
        """I420 planes (Y, U, V 2-D uint8 arrays) or None to end the stream.
        Returns False when the driver needs a receive first (EBUSY)."""
        if planes is None:
            rc = lib().mav1hip_seq_send_frame(self.h, None, None)
        else:
            pl = [np.ascontiguousarray(p, np.uint8) for p in planes]
            arr = (C.c_void_p * 3)(*[p.ctypes.data for p in pl])
            st = (C.c_int32 * 3)(*[p.shape[1] for p in pl])
            rc = lib().mav1hip_seq_send_frame(self.h, arr, st)
        if rc == EBUSY:
            return False
        if rc:
            raise self._err("send_frame", rc)
        return True

    def send_device(self, dev_ptr):
        """A padded frame resident in device memory (integer address)."""
        rc = lib().mav1hip_seq_send_device_frame(self.h, C.c_void_p(dev_ptr))
        if rc == EBUSY:
            return False
        if rc:
            raise self._err("send_device_frame", rc)
        return True

    def receive_raw(self, wait=True):
        """The next CodedFrame (its pointers valid until the next call) or
        None."""
        out = CodedFrame()
        rc = lib().mav1hip_seq_receive(self.h, C.byref(out), int(bool(wait)))
        if rc == 1:
            return None
        if rc:
            raise self._err("receive", rc)
        return out

    def receive(self, wait=True):
        """The next coded frame as a dict of numpy copies, or None."""
        return coded_dict(self.receive_raw(wait))


def coded_dict(out):
    """A CodedFrame's fields and buffers as a dict of numpy copies (None for
    None)."""
    if out is None:
        return None
    fp = out.fp
    nsb = fp.sb_cols * fp.sb_rows
    nunits = fp.lr_unit_off[2] + fp.lr_unit_rows[2] * fp.lr_unit_cols[2]
    fb = fp.offset[2] + fp.stride[2] * fp.plane_h[2]
    d = {n: getattr(out, n) for n, _ in CodedFrame._fields_[:13]}
    d["fp"] = FrameParams.from_buffer_copy(fp)
    d["plan"] = FilterPlan.from_buffer_copy(out.plan)
    d["blocks"] = _host_array(out.blocks, np.int32, nsb * 64 * 27).reshape(-1, 27)
    d["sbs"] = _host_array(out.sbs, np.int32, nsb * 89).reshape(-1, 89)
    d["coefs"] = _host_array(out.coefs, np.int32, nsb * 6144)
    d["cdef_choice"] = _host_array(out.cdef_choice, np.int8, nsb)
    d["lr_units"] = None if not out.lr_units else _host_array(out.lr_units, np.int32, nunits * 12).reshape(-1, 12)
    d["recon"] = _host_array(out.recon, np.uint8, fb)
    d["decision"] = None if not out.decision else _host_array(out.decision, np.uint8, out.decision_bytes)
    return d


class Batch:
    """Several streams coded together (mav1hip_batch_*): every stage one
    launch over the frames of all streams at the same coding position. Feed
    the streams in step; each stream's coded frames equal its own
    Sequence's."""

    def __init__(self, cfgs, lr_test=0, dump_decisions=False):
        arr = (Config * len(cfgs))(*cfgs)
        err = C.create_string_buffer(512)
        h = lib().mav1hip_batch_create(arr, len(cfgs), lr_test, int(bool(dump_decisions)), err, 512)
        if not h:
            raise HipError(err.value.decode())
        self.h = h
        self.cfgs = list(cfgs)

    def close(self):
        if self.h:
            lib().mav1hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, what, rc):
        return HipError(f"{what}: {lib().mav1hip_batch_error(self.h).decode()} (rc {rc})")

    def send(self, i, planes):
        """I420 planes for stream i, or None to end it. False on EBUSY."""
        if planes is None:
            rc = lib().mav1hip_batch_send_frame(self.h, i, None, None)
        else:
            pl = [np.ascontiguousarray(p, np.uint8) for p in planes]
            arr = (C.c_void_p * 3)(*[p.ctypes.data for p in pl])
            st = (C.c_int32 * 3)(*[p.shape[1] for p in pl])
            rc = lib().mav1hip_batch_send_frame(self.h, i, arr, st)
        if rc == EBUSY:
            return False
        if rc:
            raise self._err("send_frame", rc)
        return True

    def send_device(self, i, dev_ptr):
        rc = lib().mav1hip_batch_send_device_frame(self.h, i, C.c_void_p(dev_ptr))
        if rc == EBUSY:
            return False
        if rc:
            raise self._err("send_device_frame", rc)
        return True

    def receive_raw(self, i, wait=True):
        out = CodedFrame()
        rc = lib().mav1hip_batch_receive(self.h, i, C.byref(out), int(bool(wait)))
        if rc == 1:
            return None
        if rc:
            raise self._err("receive", rc)
        return out

    def receive(self, i, wait=True):
        return coded_dict(self.receive_raw(i, wait))

    def set_knob(self, name, value):
        """A launch shape of the batched form (mav1hip_batch_set_knob)."""
        rc = lib().mav1hip_batch_set_knob(self.h, name.encode(), int(value))
        if rc:
            raise self._err("set_knob", rc)

    def set_pruner(self, params, tau):
        """Prune the analysis with the learned pruner (params: FP32 vector)
        at confidence 1 - tau; None or tau <= 0 turns it off."""
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        self._pruner_params = p
        rc = lib().mav1hip_batch_set_pruner(self.h, None if p is None else p.ctypes.data, float(tau))
        if rc:
            raise self._err("set_pruner", rc)

    def set_knobs(self, knobs):
        for k, v in (knobs or {}).items():
            self.set_knob(k, v)


class Pruner:
    """The learned pruner's network on the GPU (mav1hip_pruner_*): FP32
    parameters, BF16 MFMA convolutions, forward, backward and Adam."""

    def __init__(self, max_batch=64):
        err = C.create_string_buffer(512)
        h = lib().mav1hip_pruner_create(max_batch, err, 512)
        if not h:
            raise HipError(err.value.decode())
        self.h = h
        self.n = lib().mav1hip_pruner_num_params()
        self.batch = 0

    def close(self):
        if self.h:
            lib().mav1hip_pruner_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, what, rc):
        if rc:
            raise HipError(f"{what}: {lib().mav1hip_pruner_error(self.h).decode()} (rc {rc})")

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n
        self._chk("set_params", lib().mav1hip_pruner_set_params(self.h, p.ctypes.data))

    def get_params(self):
        p = np.empty(self.n, np.float32)
        self._chk("get_params", lib().mav1hip_pruner_get_params(self.h, p.ctypes.data))
        return p

    def load(self, x, q, labels=None):
        x = np.ascontiguousarray(x, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        self.batch = x.shape[0]
        assert x.shape == (self.batch, 64, 64, 2) and q.shape == (self.batch,)
        lab = None if labels is None else np.ascontiguousarray(labels, np.int32)
        self._keep = (x, q, lab)
        self._chk("load", lib().mav1hip_pruner_load(self.h, x.ctypes.data, q.ctypes.data,
                                                    None if lab is None else lab.ctypes.data, self.batch))

    def forward(self):
        logits = np.empty((self.batch, 12), np.float32)
        loss = np.empty(1, np.float32)
        self._chk("forward", lib().mav1hip_pruner_forward(self.h, logits.ctypes.data, loss.ctypes.data))
        return logits, float(loss[0])

    def backward(self):
        g = np.empty(self.n, np.float32)
        self._chk("backward", lib().mav1hip_pruner_backward(self.h, g.ctypes.data))
        return g

    def train_step(self, step, lr=1e-3):
        """forward, backward and one Adam step on the loaded batch, all on
        the device; returns nothing (no host transfers)."""
        self._chk("forward", lib().mav1hip_pruner_forward(self.h, None, None))
        self._chk("backward", lib().mav1hip_pruner_backward(self.h, None))
        self._chk("adam", lib().mav1hip_pruner_adam(self.h, lr, 0.9, 0.999, 1e-8, step))


class Context:
    """A device context for frames of one size (mav1hip_create)."""

    def __init__(self, cfg):
        err = C.create_string_buffer(512)
        h = lib().mav1hip_create(C.byref(cfg), err, 512)
        if not h:
            raise HipError(err.value.decode())
        self.h = h
        self.cfg = cfg
        self.frame_bytes = lib().mav1hip_buffer_bytes(h, BUF_SRC)
        self.mv_ints = lib().mav1hip_buffer_bytes(h, BUF_MVFIELD) // 4

    def close(self):
        if self.h and not getattr(self, "_borrowed", False):
            lib().mav1hip_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise HipError(f"{what}: {lib().mav1hip_error(self.h).decode()} (rc {rc})")

    @property
    def device_name(self):
        return lib().mav1hip_device_name(self.h).decode()

    def motion_search(self, fp, src, ref, gold=None):
        mv = np.zeros(self.mv_ints, np.int32)
        mvg = np.zeros(self.mv_ints, np.int32) if fp.golden else None
        rc = lib().mav1hip_motion_search(self.h, C.byref(fp), ptr(src), ptr(ref), ptr(gold), ptr(mv), ptr(mvg))
        self._chk(rc, "motion_search")
        return mv, mvg

    def analyze(self, fp, src, ref, gold, mvf, mvg, tables=True):
        """Blocks (num_sbs*64 x 27 int32) and superblock records (num_sbs x 89);
        with tables, also the pass-1 costs and decisions."""
        nsb = fp.sb_cols * fp.sb_rows
        blocks = np.zeros((nsb * 64, 27), np.int32)
        sbs = np.zeros((nsb, 89), np.int32)
        costs = np.zeros(nsb * NUM_NODES * A_SLOTS, np.int64) if tables else None
        infos = np.zeros((nsb * NUM_NODES * A_SLOTS, 27), np.int32) if tables else None
        rc = lib().mav1hip_analyze(self.h, C.byref(fp), ptr(src), ptr(ref), ptr(gold), ptr(mvf), ptr(mvg),
                                   ptr(blocks), ptr(sbs), ptr(costs), ptr(infos))
        self._chk(rc, "analyze")
        return costs, infos, blocks, sbs

    def edge_types(self, fp, blocks, sbs):
        b = np.ascontiguousarray(blocks, np.int32).copy()
        self._chk(lib().mav1hip_edge_types(self.h, C.byref(fp), ptr(b), ptr(np.ascontiguousarray(sbs, np.int32))),
                  "edge_types")
        return b

    def recon(self, fp, blocks, sbs, src, ref, gold, rec=None):
        """Returns (rec, coefs, blocks)."""
        b = np.ascontiguousarray(blocks, np.int32).copy()
        rec = np.zeros_like(src) if rec is None else rec.copy()
        coefs = np.zeros(fp.sb_cols * fp.sb_rows * 6144, np.int32)
        rc = lib().mav1hip_recon(self.h, C.byref(fp), ptr(b), ptr(np.ascontiguousarray(sbs, np.int32)), ptr(src),
                                 ptr(rec), ptr(ref), ptr(gold), ptr(coefs))
        self._chk(rc, "recon")
        return rec, coefs, b

    def filter_info(self, fp, blocks, sbs):
        info = np.zeros(fp.mi_rows * fp.mi_cols, np.uint8)
        rc = lib().mav1hip_filter_info(self.h, C.byref(fp), ptr(np.ascontiguousarray(blocks, np.int32)),
                                       ptr(np.ascontiguousarray(sbs, np.int32)), ptr(info))
        self._chk(rc, "filter_info")
        return info

    def filter(self, fp, info, rec, nsets):
        """Deblocked copy of rec and the nsets CDEF outputs (nsets x frame_bytes)."""
        out = np.ascontiguousarray(rec, np.uint8).copy()
        cdef = np.zeros((max(nsets, 1), out.size), np.uint8)
        rc = lib().mav1hip_filter(self.h, C.byref(fp), ptr(np.ascontiguousarray(info, np.uint8)), ptr(out), nsets,
                                  ptr(cdef))
        self._chk(rc, "filter")
        return out, cdef[:nsets]

    def cdef_select(self, fp, src, cdefbuf, rec):
        out = np.ascontiguousarray(rec, np.uint8).copy()
        choice = np.zeros(fp.sb_cols * fp.sb_rows, np.int8)
        cdefbuf = np.ascontiguousarray(cdefbuf, np.uint8)
        rc = lib().mav1hip_cdef_select(self.h, C.byref(fp), ptr(src), ptr(cdefbuf), cdefbuf.shape[0], ptr(out),
                                       ptr(choice))
        self._chk(rc, "cdef_select")
        return out, choice

    def restore(self, fp, cdef, dbk, units, out=None):
        out = np.zeros_like(cdef) if out is None else np.ascontiguousarray(out, np.uint8).copy()
        rc = lib().mav1hip_restore(self.h, C.byref(fp), ptr(cdef), ptr(dbk), ptr(np.ascontiguousarray(units, np.int32)),
                                   ptr(out))
        self._chk(rc, "restore")
        return out

    def launch_filter_info(self):
        self._chk(lib().mav1hip_launch_filter_info(self.h), "launch_filter_info")

    def launch_deblock(self):
        self._chk(lib().mav1hip_launch_deblock(self.h), "launch_deblock")

    def launch_cdef(self, nsets):
        self._chk(lib().mav1hip_launch_cdef(self.h, nsets), "launch_cdef")

    def launch_save_deblocked(self):
        self._chk(lib().mav1hip_launch_save_deblocked(self.h), "launch_save_deblocked")

    def launch_cdef_select(self, nsets):
        self._chk(lib().mav1hip_launch_cdef_select(self.h, nsets), "launch_cdef_select")

    def lr_select(self, fp, src, cdef, dbk, lr_test=0):
        """Returns (units int32[n, 12], lr_type[3], reruns[3])."""
        n = sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
        units = np.zeros((n, 12), np.int32)
        lt = np.zeros(3, np.int32)
        rr = np.zeros(3, np.int32)
        rc = lib().mav1hip_lr_select(self.h, C.byref(fp), ptr(src), ptr(cdef), ptr(dbk), lr_test, ptr(units), ptr(lt),
                                     ptr(rr))
        self._chk(rc, "lr_select")
        return units, [int(v) for v in lt], [int(v) for v in rr]

    def launch_lr_select(self, lr_test=0):
        self._chk(lib().mav1hip_launch_lr_select(self.h, lr_test), "launch_lr_select")

    def decision_record(self, fp, blocks, sbs, hdr7):
        """The dump_decisions record of the coded blocks (bytes)."""
        out = np.zeros(lib().mav1hip_decision_record_bytes(C.byref(fp)), np.uint8)
        h = np.asarray(hdr7, np.int32)
        rc = lib().mav1hip_decision_record(self.h, C.byref(fp), ptr(np.ascontiguousarray(blocks, np.int32)),
                                           ptr(np.ascontiguousarray(sbs, np.int32)), ptr(h), ptr(out))
        self._chk(rc, "decision_record")
        return out

    def launch_decision_grid(self):
        self._chk(lib().mav1hip_launch_decision_grid(self.h), "launch_decision_grid")

    def launch_restore(self):
        self._chk(lib().mav1hip_launch_restore(self.h), "launch_restore")

    def launch_edge_types(self):
        self._chk(lib().mav1hip_launch_edge_types(self.h), "launch_edge_types")

    def launch_recon(self):
        self._chk(lib().mav1hip_launch_recon(self.h), "launch_recon")

    def launch_analyze(self):
        self._chk(lib().mav1hip_launch_analyze(self.h), "launch_analyze")

    def upload(self, buf, arr):
        arr = np.ascontiguousarray(arr)
        self._chk(lib().mav1hip_upload(self.h, buf, ptr(arr), arr.nbytes), "upload")

    def download(self, buf, dtype=np.uint8, count=None):
        nbytes = lib().mav1hip_buffer_bytes(self.h, buf)
        itemsize = np.dtype(dtype).itemsize
        out = np.empty(count if count is not None else nbytes // itemsize, dtype)
        self._chk(lib().mav1hip_download(self.h, buf, ptr(out), out.nbytes), "download")
        return out

    def set_params(self, fp):
        self._chk(lib().mav1hip_set_params(self.h, C.byref(fp)), "set_params")

    def launch_motion_search(self):
        self._chk(lib().mav1hip_launch_motion_search(self.h), "launch_motion_search")

    def sync(self):
        self._chk(lib().mav1hip_sync(self.h), "sync")

    def pool_alloc(self, n):
        self._chk(lib().mav1hip_pool_alloc(self.h, n), "pool_alloc")

    def pool_upload(self, idx, frame):
        frame = np.ascontiguousarray(frame, np.uint8)
        self._chk(lib().mav1hip_pool_upload(self.h, idx, ptr(frame)), "pool_upload")

    def bind_frames(self, s, r, g):
        self._chk(lib().mav1hip_bind_frames(self.h, s, r, g), "bind_frames")

    def pool_store(self, idx, buf):
        self._chk(lib().mav1hip_pool_store(self.h, idx, buf), "pool_store")

    def pool_import(self, idx, src, src_idx):
        """Copy pool frame src_idx of Context src into this context's pool
        frame idx, ordered on both contexts' streams."""
        self._chk(lib().mav1hip_pool_import(self.h, idx, src.h, src_idx), "pool_import")

    def timing(self, on=True):
        self._chk(lib().mav1hip_timing_enable(self.h, 1 if on else 0), "timing_enable")

    def timing_select(self, kernel=None):
        """Time only the launches of `kernel` (None: every kernel)."""
        self._chk(lib().mav1hip_timing_select(self.h, kernel.encode() if kernel else None), "timing_select")

    def timing_reset(self):
        self._chk(lib().mav1hip_timing_reset(self.h), "timing_reset")

    def timing_get(self, kernel):
        n = C.c_int32()
        ms = C.c_double()
        self._chk(lib().mav1hip_timing_get(self.h, kernel.encode(), C.byref(n), C.byref(ms)), "timing_get")
        return n.value, ms.value
