"""ORACLE — test infrastructure only.

The encoder's frame driver for the analysis path, restated serially in
Python over the oracle's C stages (oracle/liboracle.so), to check the HIP
sequence driver (mav1hip_seq_*) against. Only tests/ may use it.

What it restates:
  - the display-order queue and the coding order of ALTREF mini-GOPs:
    alt_drain (src/encoder.c:1252-1293) and pyr_code (:1229-1250);
  - each frame's state: alt_code (:1004-1030) — frame kind, order hint,
    the LAST / ALTREF slots, refresh flags, compound, non-reference leaf;
  - run_frame_core (:598-647): motion search (not on key frames), analysis,
    edge filter types, reconstruction, filter info, deblocking, CDEF, the
    deblocked copy, the CDEF choice, the restoration search and apply (the
    output starting as a copy of the reconstruction);
  - the slot writes after the core (:1128-1140): every refreshed slot holds
    the final padded picture;
  - each engine keeps its own coefficient and reconstruction buffers across
    the frames it codes (src/encoder.c:196-208: engines are separate
    encoders), the reference chain on engine 0 and the other frames on
    engine 1 (:1046-1052). Bytes of the padded picture that no block of a
    frame covers (rows and columns past the mi area) therefore keep what the
    engine's earlier frames left there, as in the reference; the swap of the
    reconstruction and restoration buffers (:638-641) leaves those bytes
    alike in both.
The per-frame parameters come from the library's setup_params /
frame_qindex / filter_plan, which tests/test_params.py pins independently.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


def _lib():
    lib = C.CDLL(ORACLE_PATH)
    for n in ("motion_search", "analyze", "edge_types", "recon", "filter_info", "loop_filter", "cdef", "cdef_select",
              "lr_apply", "lr_select", "decision_grid"):
        getattr(lib, "oracle_" + n).restype = None
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleSequence:
    """Serial restatement of the driver: send display-order padded frames,
    then end(); `coded` lists the coded frames in coding order."""

    def __init__(self, mav1hip, cfg, lr_test=0):
        self.m = mav1hip
        self.lib = _lib()
        self.cfg = self.m.Config.from_buffer_copy(cfg)
        self.m.lib().mav1hip_config_normalize(C.byref(self.cfg))
        self.lr_test = lr_test
        self.inq = []
        self.in_disp = 0
        self.key_disp = 0
        self.anchor = 0
        self.slot = [None] * 8
        self.slot_hint = [0] * 8
        self.coded = []
        self.mats = self.m.forward_matrices()
        self.coefs = [None, None]
        self.rec = [None, None]

    # --- alt_drain / pyr_code (src/encoder.c:1229-1293) ---
    def send(self, frame):
        self.inq.append(frame)
        self.in_disp += 1
        self._drain(False)

    def end(self):
        self._drain(True)

    def _drain(self, flush):
        N, keyint = self.cfg.altref, self.cfg.keyint
        while self.inq:
            d0 = self.in_disp - len(self.inq)
            is_key = d0 == 0 or (keyint > 0 and d0 % keyint == 0)
            L = 1 if is_key else N
            if not is_key and keyint > 0:
                L = min(L, (d0 // keyint + 1) * keyint - d0)
            if len(self.inq) < L:
                if not flush:
                    return
                L = len(self.inq)
            if is_key or L == 1:
                self._code(d0, is_key, 0, self.anchor, -1, 0x01, 0)
                self.anchor = 0
            elif self.cfg.leaf:
                arf = 3 if self.anchor == 2 else 2
                self._code(d0 + L - 1, 0, 1, self.anchor, -1, 1 << arf, 1)
                self._pyr(d0 - 1, self.anchor, d0 + L - 1, arf, 0)
                self.anchor = arf
            else:
                arf = 3 if self.anchor == 2 else 2
                self._code(d0 + L - 1, 0, 1, self.anchor, -1, 1 << arf, 1)
                for i in range(L - 1):
                    self._code(d0 + i, 0, 2, self.anchor if i == 0 else 0, arf, 0x01, 0)
                self.anchor = arf
            del self.inq[:L]

    def _pyr(self, a, as_, b, bs, depth):
        if b - a - 1 >= 3 and depth < 4:
            m, ms = (a + b) // 2, 4 + depth
            self._code(m, 0, 3, as_, bs, 1 << ms, 1)
            self._pyr(a, as_, m, ms, depth + 1)
            self._pyr(m, ms, b, bs, depth + 1)
            return
        for d in range(a + 1, b):
            self._code(d, 0, 2, as_, bs, 0, 0)

    # --- alt_code (src/encoder.c:1004-1030) + run_frame_core (:598-647) ---
    def _code(self, disp, is_key, kind, last, arf, refresh, hidden):
        m, cfg = self.m, self.cfg
        d0 = self.in_disp - len(self.inq)
        src = self.inq[disp - d0]
        if is_key:
            self.key_disp = disp
        refresh = 0xFF if is_key else refresh
        nonref = int(bool(cfg.leaf) and not is_key and refresh == 0)
        kind = 0 if is_key else kind
        order_hint = disp - self.key_disp
        fr = m.frame_info(is_key=is_key, frame_kind=kind, nonref=nonref, order_hint=order_hint,
                          reference_select=int(not is_key and kind >= 2 and arf >= 0))
        fp, npre = m.setup_params(cfg, fr)
        plan = m.filter_plan(cfg, fr, fp, npre)
        last_hint = -1 if is_key else self.slot_hint[last]
        arf_hint = self.slot_hint[arf] if (not is_key and arf >= 0) else -1
        engine = 0 if (refresh and kind != 3) else 1
        ref = None if is_key else self.slot[last]
        gold = self.slot[arf] if (not is_key and arf >= 0 and fp.golden) else None
        nsb = fp.sb_cols * fp.sb_rows
        lib, P = self.lib, _p
        mvf = np.zeros(nsb * 170, np.int32)
        mvg = np.zeros(nsb * 170, np.int32)
        if not is_key:
            lib.oracle_motion_search(C.byref(fp), P(src), P(ref), P(gold), P(mvf), P(mvg))
        costs = np.zeros(nsb * 1105, np.int64)
        infos = np.zeros(nsb * 1105 * 27, np.int32)
        blocks = np.zeros((nsb * 64, 27), np.int32)
        sbs = np.zeros((nsb, 89), np.int32)
        refp = src if ref is None else ref
        lib.oracle_analyze(C.byref(fp), P(src), P(refp), P(gold), P(mvf), P(mvg), P(costs), P(infos), P(blocks),
                           P(sbs))
        if cfg.intra_edge_filter:
            lib.oracle_edge_types(C.byref(fp), P(blocks), P(sbs))
        if self.coefs[engine] is None:
            self.coefs[engine] = np.zeros(nsb * 6144, np.int32)
        coefs = self.coefs[engine]
        if self.rec[engine] is None:
            self.rec[engine] = np.zeros_like(src)
        rec = self.rec[engine]
        lib.oracle_recon(C.byref(fp), P(blocks), P(sbs), P(src), P(rec), P(refp), P(gold), P(coefs), P(self.mats))
        leafopt = (cfg.knob_leafopt if cfg.knob_leafopt >= 0 else 0x3F) if (cfg.altref and nonref) else 0
        cdef_off = (leafopt & 4) != 0
        fb = rec.size
        choice = np.zeros(nsb, np.int8)
        info = np.zeros(fp.mi_rows * fp.mi_cols, np.uint8)
        nsets = plan.cdef_sets
        cbuf = np.zeros(max(nsets, 1) * fb, np.uint8)
        if plan.filter_info:
            lib.oracle_filter_info(C.byref(fp), P(blocks), P(sbs), P(info))
            if plan.deblock:
                lib.oracle_loop_filter(C.byref(fp), P(rec), P(info))
            if nsets:
                lib.oracle_cdef(C.byref(fp), P(rec), P(cbuf), C.c_int64(fb), P(info), nsets)
        dbk = rec.copy() if plan.restoration else None
        if plan.cdef_select:
            lib.oracle_cdef_select(C.byref(fp), P(src), P(cbuf), C.c_int64(fb), nsets, P(rec), P(choice))
        units = None
        final = rec
        if plan.restoration:
            nunits = fp.lr_unit_off[2] + fp.lr_unit_rows[2] * fp.lr_unit_cols[2]
            units = np.zeros((nunits, 12), np.int32)
            tmp = np.zeros_like(rec)
            lib.oracle_lr_select(C.byref(fp), P(src), P(rec), P(dbk), P(tmp), P(units), self.lr_test)
            out = rec.copy()
            lib.oracle_lr_apply(C.byref(fp), P(rec), P(dbk), P(out), P(units))
            final = out
        grid = np.zeros(((fp.mi_rows + 1) >> 1) * ((fp.mi_cols + 1) >> 1) * 8, np.uint8)
        lib.oracle_decision_grid(C.byref(fp), P(blocks), P(sbs), P(grid))
        hdr = np.array([0x4345444D, order_hint, last_hint, arf_hint, 0 if is_key else 1, kind, nonref,
                        fp.base_q_idx, (fp.mi_rows + 1) >> 1, (fp.mi_cols + 1) >> 1, 0], np.int32)
        final = final.copy()  # the engine's buffer changes with its next frame
        for k in range(8):
            if (refresh >> k) & 1:
                self.slot[k] = final
                self.slot_hint[k] = order_hint
        self.coded.append({
            "coded_index": len(self.coded), "disp": disp, "is_key": int(is_key), "frame_kind": kind,
            "nonref": nonref, "hidden": hidden, "order_hint": order_hint, "refresh_frame_flags": refresh,
            "last_slot": -1 if is_key else last, "arf_slot": -1 if is_key else arf, "last_hint": last_hint,
            "arf_hint": arf_hint, "engine": engine, "fp": fp, "plan": plan, "blocks": blocks, "sbs": sbs,
            "coefs": coefs.copy(), "cdef_choice": choice, "lr_units": units, "recon": final.copy(),
            "decision": np.concatenate([hdr.view(np.uint8), grid]),
        })
