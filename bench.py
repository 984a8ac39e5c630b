#!/usr/bin/env python3
"""Benchmark of the HIP analysis backend on one MI355X.

Workload: the 1080p configuration (1920x1080 8-bit 4:2:0, synthetic content
with 12 objects moving up to +-48 px/frame, luma noise sigma 3), encoded with
--turbo -q 125: frame 0 is a key frame, then mini-GOPs of 2 in coding order
(hidden ALTREF frame predicting from LAST, then a non-reference leaf frame
predicting from LAST and ALTREF). A step is one frame of that sequence
through the backend stages built on the GPU so far (listed in
config["stages"]), with every source frame already resident in HBM (a device
frame pool filled before the timed region). Reference pictures are the
reconstructed, filtered pictures of the referenced frames, kept in two
device slots (closed loop, starting from a coded key frame).

Prints one JSON line (rank 0) with the throughput in frames/s, the roofline
of the dominant kernel (per-launch durations from HIP events recorded on the
context's stream during the timed region) and the reference CPU baseline
(the reference's own kernels, single-threaded, on a bounded sample).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mav1-encoder_amd"))

import numpy as np  # noqa: E402

import mav1hip  # noqa: E402
import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
W, H = 1920, 1080
Q = 125


def coding_order(n_steps):
    """(kind, frame, last, altref) for the first n_steps coded frames after
    the key frame."""
    out = []
    j = 0
    while len(out) < n_steps:
        a, leaf = 2 * j + 2, 2 * j + 1
        out.append(("arf", a, 2 * j, -1))
        out.append(("leaf", leaf, 2 * j, a))
        j += 1
    return out[:n_steps]


def algorithmic_bytes(kernel, fp, nref):
    """Compulsory HBM bytes of one launch: each input read once, each output
    written once."""
    Y = fp.frame_w * fp.frame_h
    dec = ((fp.frame_w + 3) >> 2) * ((fp.frame_h + 3) >> 2)
    nsb = fp.sb_cols * fp.sb_rows
    if kernel == "qpel_build":
        return Y + 16 * Y
    if kernel == "dec_build":
        return Y + dec
    if kernel.startswith("me_level"):
        level = int(kernel[-1])
        field = nref * nsb * (4 ** level) * 8 + (nref * nsb * (4 ** (level - 1)) * 8 if level else 0)
        return nref * (Y + Y) + field
    F = Y + 2 * (((fp.frame_w + 1) >> 1) * ((fp.frame_h + 1) >> 1))
    if kernel.startswith("analyze_level"):
        level = int(kernel[-1])
        items = nsb * (4 ** level)  # turbo: one slot per node
        return F * (1 + nref) + items * (nref * 8 + 8 + 108)
    if kernel == "analyze_table":
        return nsb * 85 * (8 + 108) + nsb * (64 * 108 + 356)
    if kernel == "recon_inter":
        # prediction reads of the references, source read, reconstruction and levels written, blocks in/out
        return F * (nref + 1) + F + 4 * F + nsb * 64 * 108 * 2
    if kernel == "edge_types":
        return nsb * 64 * 108 * 2 + 2 * fp.mi_rows * fp.mi_cols
    mi = fp.mi_rows * fp.mi_cols
    if kernel == "filter_info":
        return nsb * 64 * 108 + mi
    if kernel == "deblock":
        return 2 * F + mi
    if kernel == "cdef":
        return F + 2 * F + mi  # turbo: two parameter sets written
    if kernel == "cdef_select":
        return 2 * F + F + F
    if kernel in ("lr_fit", "lr_path"):
        return 3 * F
    if kernel == "lr_chain":
        return 48 * sum(fp.lr_unit_rows[p] * fp.lr_unit_cols[p] for p in range(3))
    if kernel == "lr_apply":
        return 3 * F
    return 0


# Block widths / heights by bsize (BLOCK_4X4 .. BLOCK_64X16).
BW = [4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 4, 16, 8, 32, 16, 64]
BH = [4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 64, 32, 64, 128, 64, 128, 16, 4, 32, 8, 64, 16]

# HBM traffic per recon_intra launch over the timed frames of the default run
# (40 steps after 6 warm-up ones, ALTREF / leaf alternating), from the
# FETCH_SIZE and WRITE_SIZE passes of profiles/recon_intra_1080p_turbo_pmc.md,
# scaled by the frame copy's factors there (fetch x1.985, write x1.0).
RECON_INTRA_TRAFFIC = 6617313


def intra_phase_bytes(fp, blocks, sbs):
    """Bytes the intra phase of one frame must move: every superblock's
    decode-order scan of its block records (is_inter, 4 B each); for each
    superblock holding intra blocks its reconstruction tile with the row
    above and the columns left (luma 132 + 64 x 68, chroma 68 + 32 x 36
    each), its source and the tile written back (6144 B each), the short
    transform matrices (4032 B), its block records (108 B) and the intra
    ones written back, and the intra blocks' levels (4 B per coefficient:
    luma Min(32, w) x Min(32, h), chroma likewise at half size)."""
    total = 0
    for s in range(fp.sb_cols * fp.sb_rows):
        first, n = int(sbs[s, 0]), int(sbs[s, 1])
        b = blocks[first:first + n]
        intra = b[b[:, 3] == 0]
        total += 4 * n
        if len(intra):
            total += (132 + 64 * 68) + 2 * (68 + 32 * 36) + 6144 + 6144 + 4032 + 108 * (n + len(intra))
            for bs in intra[:, 2]:
                w, h = BW[bs], BH[bs]
                total += 4 * (min(32, w) * min(32, h) + 2 * min(32, w // 2) * min(32, h // 2))
    return total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    # deblocking + CDEF + loop restoration on (turbo turns restoration off unless asked for)
    cfg = mav1hip.default_config(W, H, turbo=1, qindex=Q, restoration=1)
    geo = synth.Geometry(W, H, cfg.tile_cols_log2)
    order = coding_order(args.warmup + args.steps)
    nframes = max(max(f, l, a) for _, f, l, a in order) + 1
    gen = synth.config_synth(2, seed=2024)
    uniq = min(nframes, 50)
    frames = [synth.pad_frame(geo, gen.frame(t)) for t in range(uniq)]

    # Two frame engines, as the reference's driver runs a stream's mini-GOP
    # (encoder.c:1150-1224: engine encoders with a backend instance each, on
    # their own queues): engine A codes the key frame and the ALTREF chain,
    # engine B the leaf frames. Each engine is a context with its own stream
    # and buffers; an ALTREF picture passes from A's pool to B's with
    # mav1hip_pool_import, ordered on both streams.
    engA, engB = mav1hip.Context(cfg), mav1hip.Context(cfg)
    # pool: the source frames, then two reference slots
    for eng in (engA, engB):
        eng.pool_alloc(uniq + 2)
        for i, f in enumerate(frames):
            eng.pool_upload(i, f)
    params, plans = {}, {}
    for kind in ("key", "arf", "leaf"):
        fr = mav1hip.turbo_frame(kind, 0 if kind == "key" else 1)
        fp, npre = mav1hip.setup_params(cfg, fr)
        params[kind], plans[kind] = fp, mav1hip.filter_plan(cfg, fr, fp, npre)
    slot_of = {}    # frame -> engine A's slot holding its final picture
    b_slot_of = {}  # frame -> engine B's slot holding its imported copy

    def early_stages(ctx, kind):
        if kind != "key":
            ctx.launch_motion_search()
        ctx.launch_analyze()
        ctx.launch_edge_types()

    def late_stages(ctx, kind):
        plan = plans[kind]
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

    def run_stages(ctx, kind):
        early_stages(ctx, kind)
        return late_stages(ctx, kind)

    last_run = [None, None]  # engine and final buffer of the last coded frame
    # A leaf frame waits here for the next ALTREF: it is enqueued after that
    # ALTREF's motion search, analysis and edge types, so its imports (which
    # order it after everything engine A has enqueued) hold it until then and
    # it runs beside the ALTREF's reconstruction, whose intra phase leaves
    # most CUs idle. Only the enqueue order changes, not the results.
    pending = []

    def step(kind, f, last, alt, overlap=True):
        if kind == "arf":  # engine A: LAST is the previous ALTREF (or the key frame)
            ref = slot_of[last]
            engA.bind_frames(f % uniq, ref, -1)
            engA.set_params(params[kind])
            early_stages(engA, kind)
            if pending:
                run_leaf(*pending.pop())
            final = late_stages(engA, kind)
            target = uniq if ref != uniq else uniq + 1  # the slot LAST does not hold
            engA.pool_store(target, final)
            slot_of[f] = target
            last_run[:] = [engA, final]
            return
        if overlap:
            pending.append((f, last, alt))
        else:
            run_leaf(f, last, alt)

    def flush():
        if pending:
            run_leaf(*pending.pop())

    def run_leaf(f, last, alt):
        # engine B: LAST and ALTREF come from engine A, each imported once
        # into the slot the other one does not hold
        kind = "leaf"
        engB.bind_frames(f % uniq, -1, -1)
        for need, other in ((last, alt), (alt, last)):
            if need not in b_slot_of:
                keep = b_slot_of.get(other)
                slot = uniq if keep != uniq else uniq + 1
                for g in [g for g, sl in b_slot_of.items() if sl == slot]:
                    del b_slot_of[g]
                engB.pool_import(slot, engA, slot_of[need])
                b_slot_of[need] = slot
        engB.bind_frames(f % uniq, b_slot_of[last], b_slot_of[alt])
        engB.set_params(params[kind])
        last_run[:] = [engB, run_stages(engB, kind)]

    def code_key():
        """Frame 0: the key frame (untimed) on engine A, its reconstruction in A's first slot."""
        slot_of.clear()
        b_slot_of.clear()
        engA.bind_frames(0, -1, -1)
        engA.set_params(params["key"])
        engA.pool_store(uniq, run_stages(engA, "key"))
        slot_of[0] = uniq

    def sync():
        engA.sync()
        engB.sync()

    kernels = ["qpel_build", "dec_build", "me_level0", "me_level1", "me_level2", "me_level3", "analyze_level0",
               "analyze_level1", "analyze_level2", "analyze_level3", "analyze_table", "edge_types", "recon_inter",
               "intra_sched", "recon_intra", "filter_info", "deblock", "cdef", "cdef_select", "lr_fit", "lr_path", "lr_chain", "lr_apply"]

    def kernel_stats():
        out = {}
        for k in kernels:
            na, ta = engA.timing_get(k)
            nb, tb = engB.timing_get(k)
            out[k] = (na + nb, ta + tb)
        return out

    def timing(on, only=None):
        for eng in (engA, engB):
            eng.timing_select(only)
            eng.timing(on)
            eng.timing_reset()

    code_key()

    # warm-up, every kernel timed: it names the dominant kernel (largest total)
    timing(True)
    for s in order[:args.warmup]:
        step(*s)
    flush()
    sync()
    warm = kernel_stats()
    dom = max(warm, key=lambda k: warm[k][1]) if any(v[0] for v in warm.values()) else None
    # the timed region: events around the dominant kernel's launches only
    # (each timed launch adds two event records to its stream); without a
    # warm-up, around every launch, and the dominant kernel named after
    timing(True, dom)
    t0 = time.perf_counter()
    for s in order[args.warmup:]:
        step(*s)
    flush()
    sync()
    elapsed = time.perf_counter() - t0
    timed = kernel_stats()
    if dom is None:
        dom = max(timed, key=lambda k: timed[k][1])
    n_launch, tot_ms = timed[dom]
    timing(False)
    last_out = last_run[0].download(last_run[1])
    # the intra phase's algorithmic bytes of every timed frame: the whole
    # sequence coded again, untimed, from the key frame on (the coding is
    # deterministic), the block records read back after each timed frame;
    # every kernel timed, one frame at a time, for the per-kernel table
    code_key()
    timing(True)
    intra_bytes = []
    for i, s in enumerate(order):
        step(*s, overlap=False)
        if i < args.warmup:
            continue
        eng = last_run[0]
        eng.sync()
        p = params[s[0]]
        nsb = p.sb_cols * p.sb_rows
        blk = eng.download(mav1hip.BUF_BLOCKS, np.int32, nsb * 64 * 27).reshape(-1, 27)
        sbi = eng.download(mav1hip.BUF_SBS, np.int32, nsb * 89).reshape(-1, 89)
        intra_bytes.append(intra_phase_bytes(p, blk, sbi))
    sync()
    if not np.array_equal(last_run[0].download(last_run[1]), last_out):
        raise RuntimeError("the second coding of the timed frames differs from the first")
    stats = kernel_stats()  # the second coding, frame by frame
    timing(False)
    fps = args.steps / elapsed

    # the dominant kernel's mean launch duration over the timed region
    avg_s = tot_ms / 1e3 / max(n_launch, 1)
    # average algorithmic bytes per launch over the frames of the timed region
    nref_avg = np.mean([2 if k == "leaf" else 1 for k, *_ in order[args.warmup:]])
    if dom in ("qpel_build", "dec_build"):
        per_launch = algorithmic_bytes(dom, params["arf"], 1)
    elif dom == "recon_intra":
        per_launch = np.mean(intra_bytes)  # one launch per timed frame
    else:
        per_launch = np.mean([algorithmic_bytes(dom, params[k], 2 if k == "leaf" else 1)
                              for k, *_ in order[args.warmup:]])
    achieved = per_launch / avg_s / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds, frames, params, uniq)

    line = {
        "metric": "1080p encode fps, 1 MI355X: plain HIP vs batched+pruned at ≤1% BD-rate Δ",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generator: 6 sinusoids + 12 moving rectangles + noise)",
        "config": {
            "workload": "1920x1080 8-bit 4:2:0, --turbo -q 125, turbo GOP (ARF + leaf), plain HIP path",
            "stages": ["motion_search", "analyze", "edge_types", "recon", "filter_info", "deblock", "cdef",
                       "cdef_select", "lr_select", "lr_apply"],
            "frame_kinds": "alternating hidden ALTREF (1 ref) / leaf (2 refs); leaf frames are not filtered",
            "references": "the reconstructed, filtered pictures of the referenced frames (closed loop from a coded key frame)",
            "filters": "deblocking + CDEF + loop restoration on hidden ALTREF frames (leaf frames skip them)",
            "engines": "2 (engine A: key frame + ALTREF chain; engine B: leaf frames), one context and stream each, "
                       "as the reference's driver runs a stream's mini-GOP on engine encoders; a leaf is enqueued "
                       "behind the next ALTREF's analysis and runs beside its reconstruction",
            "device": engA.device_name,
        },
        "kernels_ms_per_launch": {k: round(v[1] / max(v[0], 1), 5) for k, v in stats.items()},
        "kernels_ms_note": "per-kernel means from the untimed second coding, one frame at a time (no engine "
                           "overlap); the timed region records events around the dominant kernel only",
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": RECON_INTRA_TRAFFIC if dom == "recon_intra" else None,
            "algorithmic_bytes_per_launch": int(per_launch),
            "avg_launch_ms": round(avg_s * 1e3, 5),
            "note": ("latency-bound dependency chain (superblocks in dataflow order, blocks in decode order), "
                     "integer VALU, no MFMA: the HBM fraction is small by construction" if dom == "recon_intra" else
                     "integer-VALU bound, no MFMA: the HBM fraction is small by construction"),
        },
        "cpu_baseline": cpu,
    }
    engA.close()
    engB.close()
    print(json.dumps(line))


def cpu_baseline(seconds, frames, params, uniq):
    """The reference's own kernels (oracle/_ref: motion search, analysis,
    reconstruction, deblocking, CDEF, restoration search and apply; the edge
    filter types, filter info and CDEF choice from the oracle, as the
    reference's are in its frame driver), serial on one host core, over the
    same frame sequence until `seconds` have passed (whole mini-GOPs)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libmav1ref.so")
    kind, pre = "reference", "ref_"
    if not os.path.exists(path):
        path, kind, pre = os.path.join(ROOT, "oracle", "liboracle.so"), "port", "oracle_"
    lib = C.CDLL(path)
    me, ana, rec_fn = getattr(lib, pre + "motion_search"), getattr(lib, pre + "analyze"), getattr(lib, pre + "recon")
    lf, cdef = getattr(lib, pre + "loop_filter"), getattr(lib, pre + "cdef")
    lrsel, lrapp = getattr(lib, pre + "lr_select"), getattr(lib, pre + "lr_apply")
    nunits = sum(params["arf"].lr_unit_rows[p] * params["arf"].lr_unit_cols[p] for p in range(3))
    units = np.zeros(nunits * 12, np.int32)
    dbk, tmp, lrout = np.zeros_like(frames[0]), np.zeros_like(frames[0]), np.zeros_like(frames[0])
    # the driver's edge filter types, filter info and CDEF choice (the reference's are in its frame driver)
    orc = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    fb = frames[0].size
    info = np.zeros(params["arf"].mi_rows * params["arf"].mi_cols, np.uint8)
    cbuf = np.zeros(2 * fb, np.uint8)
    choice = np.zeros(params["arf"].sb_cols * params["arf"].sb_rows, np.int8)
    mats = mav1hip.forward_matrices()
    rec = np.zeros_like(frames[0])
    coefs = np.zeros(params["arf"].sb_cols * params["arf"].sb_rows * 6144, np.int32)
    nsb = params["arf"].sb_cols * params["arf"].sb_rows
    mv = np.zeros(nsb * 170, np.int32)
    mvg = np.zeros(nsb * 170, np.int32)
    costs = np.zeros(nsb * 1105, np.int64)
    infos = np.zeros(nsb * 1105 * 27, np.int32)
    blocks = np.zeros(nsb * 64 * 27, np.int32)
    sbs = np.zeros(nsb * 89, np.int32)
    done = 0
    t0 = time.perf_counter()
    for kind_f, f, last, alt in coding_order(64):
        fp = params[kind_f]
        s, r, g = (mav1hip.ptr(frames[f % uniq]), mav1hip.ptr(frames[last % uniq]),
                   mav1hip.ptr(frames[alt % uniq]) if alt >= 0 else None)
        me(C.byref(fp), s, r, g, mav1hip.ptr(mv), mav1hip.ptr(mvg))
        ana(C.byref(fp), s, r, g, mav1hip.ptr(mv), mav1hip.ptr(mvg), mav1hip.ptr(costs), mav1hip.ptr(infos),
            mav1hip.ptr(blocks), mav1hip.ptr(sbs))
        orc.oracle_edge_types(C.byref(fp), mav1hip.ptr(blocks), mav1hip.ptr(sbs))
        rec_fn(C.byref(fp), mav1hip.ptr(blocks), mav1hip.ptr(sbs), s, mav1hip.ptr(rec), r, g, mav1hip.ptr(coefs),
               mav1hip.ptr(mats))
        if kind_f == "arf":
            orc.oracle_filter_info(C.byref(fp), mav1hip.ptr(blocks), mav1hip.ptr(sbs), mav1hip.ptr(info))
            lf(C.byref(fp), mav1hip.ptr(rec), mav1hip.ptr(info))
            cdef(C.byref(fp), mav1hip.ptr(rec), mav1hip.ptr(cbuf), C.c_long(fb), mav1hip.ptr(info), 2)
            dbk[:] = rec
            orc.oracle_cdef_select(C.byref(fp), s, mav1hip.ptr(cbuf), C.c_int64(fb), 2, mav1hip.ptr(rec),
                                   mav1hip.ptr(choice))
            f2 = mav1hip.FrameParams.from_buffer_copy(fp)
            lrsel(C.byref(f2), s, mav1hip.ptr(rec), mav1hip.ptr(dbk), mav1hip.ptr(tmp), mav1hip.ptr(units), 0)
            lrapp(C.byref(f2), mav1hip.ptr(rec), mav1hip.ptr(dbk), mav1hip.ptr(lrout), mav1hip.ptr(units))
        done += 1
        if time.perf_counter() - t0 >= seconds and done % 2 == 0:
            break
    dt = time.perf_counter() - t0
    return {
        "value": round(done / dt, 4),
        "unit": "frames/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{done} frames of the same 1080p turbo GOP (ARF/leaf alternating), motion search + analysis + edge types + "
                  f"reconstruction + (ARF) filter info, deblocking, CDEF, CDEF choice, restoration search and apply "
                  f"(the search alone takes ~45 s per 1080p frame, so the sample is whole mini-GOPs past the time "
                  f"bound); references are source pictures, "
                  f"{dt:.1f} s on one host core ({os.cpu_count()} visible)",
    }


if __name__ == "__main__":
    main()
