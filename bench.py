#!/usr/bin/env python3
"""Benchmark of the HIP analysis backend on one MI355X.

Workload: the 1080p stream (1920x1080 8-bit 4:2:0, synthetic content with 12
objects moving up to +-48 px/frame, luma noise sigma 3), encoded with
--turbo -q 125 --keyint 50 and loop restoration on, through the sequence
driver (mav1hip_seq_*): frame 0 is a key frame, then ALTREF mini-GOPs of 2
(hidden ALTREF, then a non-reference leaf frame), and a regular frame closes
a stream of even length. A step is one frame; the default 50 steps are the
whole 50-frame stream, key frame included. Source frames are resident in
HBM when the timed region starts; every frame's outputs for the CPU side
(block and superblock records, levels, CDEF choices, restoration units, the
final picture) are copied to pinned host memory inside the timed region.
Reference pictures are the driver's own reconstructions (closed loop).

Prints one JSON line with the throughput in frames/s, the same stream with
the sources sent from host memory (PCIe included), the roofline of the
dominant kernel (per-launch durations from HIP events recorded on the
engines' streams during the timed region) and the CPU baseline (the
oracle's serial driver and stages on the host cores, on a bounded sample).
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

import autotune  # noqa: E402
import mav1hip  # noqa: E402
import synth  # noqa: E402
import train_pruner  # noqa: E402

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

# HBM traffic per launch of the dominant kernel over the timed stream of the
# default run, from the FETCH_SIZE and WRITE_SIZE passes of
# profiles/recon_intra_1080p_turbo_stream_pmc.md, scaled by the known-size
# copy's factors there (fetch x1.9992, write x1.0).
TRAFFIC = {"recon_intra": 4950880}
# VALUBusy (%) per launch over the timed stream, from the counter pass of the
# same summary (100 * SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE).
VALU_BUSY = {"recon_intra": 0.07, "lr_fit": 6.19, "recon_inter": 2.22, "me_level0": 2.46, "me_level2": 5.62,
             "me_level3": 11.50, "analyze_level0": 3.22, "qpel_build": 6.60, "cdef": 7.44}


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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=3)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched 16-stream leg")
    ap.add_argument("--batch-frames", type=int, default=60)
    ap.add_argument("--batched-only", action="store_true", help="run only the batched 16-stream leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 64 MiB device-to-device copy 5 times (the HBM counters' scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.batched_only:
        print(json.dumps({"batched": batched_leg(args.batch_frames, torch)}))
        return

    # deblocking + CDEF + loop restoration on (turbo turns restoration off unless asked for)
    cfg = mav1hip.default_config(W, H, turbo=1, qindex=Q, restoration=1, keyint=50)
    geo = synth.Geometry(W, H, cfg.tile_cols_log2)
    n = max(args.steps, args.warmup, args.cpu_frames)
    planes = generate_frames(n)
    padded = [synth.pad_frame(geo, p) for p in planes]
    sha_ok = check_frame_hashes(planes)
    dev = [torch.from_numpy(f).to("cuda:0") for f in padded]
    torch.cuda.synchronize()

    kernels = ["qpel_build", "dec_build", "me_level0", "me_level1", "me_level2", "me_level3", "analyze_level0",
               "analyze_level1", "analyze_level2", "analyze_level3", "analyze_table", "edge_types", "recon_inter",
               "intra_sched", "recon_intra", "filter_info", "deblock", "cdef", "cdef_select", "lr_fit", "lr_path",
               "lr_chain", "lr_apply"]

    devname = [""]

    def run(nframes, device=True, timed=None, one_at_a_time=False, on_frame=None):
        """The stream's first nframes frames through a new driver; returns
        (elapsed s, receive times, per-kernel (launches, ms) over both
        engines, the last frame's final picture)."""
        seq = mav1hip.Sequence(cfg)
        engines = [seq.engine(0), seq.engine(1)]
        devname[0] = engines[0].device_name
        for e in engines:
            e.timing_select(timed if timed != "all" else None)
            e.timing(timed is not None)
            e.timing_reset()
        stamps, last = [], [None]

        def take(r):
            stamps.append(time.perf_counter())
            last[0] = r
            if on_frame is not None:
                on_frame(r)

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(nframes):
            while not (seq.send_device(dev[i].data_ptr()) if device else seq.send(planes[i])):
                take(seq.receive_raw(True))
            while True:
                r = seq.receive_raw(one_at_a_time)
                if r is None:
                    break
                take(r)
        while not seq.send(None):
            take(seq.receive_raw(True))
        while True:
            r = seq.receive_raw(True)
            if r is None:
                break
            take(r)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        fb = last[0].fp.offset[2] + last[0].fp.stride[2] * last[0].fp.plane_h[2]
        final = np.ctypeslib.as_array(C.cast(last[0].recon, C.POINTER(C.c_uint8)), (fb,)).copy()
        stats = {}
        for k in kernels:
            tot = [e.timing_get(k) for e in engines]
            stats[k] = (sum(t[0] for t in tot), sum(t[1] for t in tot))
        seq.close()
        return elapsed, [s - t0 for s in stamps], stats, final

    # warm-up stream, every kernel timed: names the dominant kernel (largest total)
    _, _, warm, _ = run(args.warmup, timed="all")
    dom = max(warm, key=lambda k: warm[k][1])
    # the timed stream: events around the dominant kernel's launches only
    elapsed, stamps, timed_stats, final = run(args.steps, timed=dom)
    n_launch, tot_ms = timed_stats[dom]
    fps = args.steps / elapsed
    steady = (args.steps - 1) / (stamps[-1] - stamps[0]) if len(stamps) > 1 else None
    # the same stream again, one frame at a time with every kernel timed:
    # the per-kernel table, the block records of every frame (the intra
    # phase's algorithmic bytes) and a check that the coding is repeatable
    intra_bytes = []

    def on_frame(r):
        fp = r.fp
        nsb = fp.sb_cols * fp.sb_rows
        blk = np.ctypeslib.as_array(C.cast(r.blocks, C.POINTER(C.c_int32)), (nsb * 64 * 27,)).reshape(-1, 27)
        sbi = np.ctypeslib.as_array(C.cast(r.sbs, C.POINTER(C.c_int32)), (nsb * 89,)).reshape(-1, 89)
        intra_bytes.append(intra_phase_bytes(fp, blk, sbi))

    _, _, stats, final2 = run(args.steps, timed="all", one_at_a_time=True, on_frame=on_frame)
    if not np.array_equal(final, final2):
        raise RuntimeError("the second coding of the stream differs from the first")
    pcie = None
    if not args.no_pcie:
        e2, _, _, final3 = run(args.steps, device=False)
        if not np.array_equal(final, final3):
            raise RuntimeError("the stream sent from host memory codes differently")
        pcie = round(args.steps / e2, 3)

    # the dominant kernel's mean launch duration over the timed region
    avg_s = tot_ms / 1e3 / max(n_launch, 1)
    params = {}
    for kind in ("key", "arf", "leaf"):
        fr = mav1hip.turbo_frame(kind, 0 if kind == "key" else 1)
        params[kind], _ = mav1hip.setup_params(cfg, fr)
    if dom == "recon_intra":
        per_launch = float(np.mean(intra_bytes))  # one launch per frame
    else:
        kinds = ["key"] + [k for k, *_ in coding_order(args.steps - 1)]
        per_launch = float(np.mean([algorithmic_bytes(dom, params[k], 2 if k == "leaf" else 1) for k in kinds]))
    achieved = per_launch / avg_s / 1e9

    batched = None
    if not args.no_batched:
        del dev
        torch.cuda.empty_cache()
        batched = batched_leg(args.batch_frames, torch)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, padded, args.cpu_frames, args.cpu_runs)

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
            "workload": f"1920x1080 8-bit 4:2:0, --turbo -q 125 --keyint 50, loop restoration on: a {args.steps}-frame "
                        "stream through the sequence driver, key frame included (key frame, ALTREF / leaf mini-GOPs, "
                        "a regular frame closing an even-length stream), plain HIP path",
            "stages": ["motion_search", "analyze", "edge_types", "recon", "filter_info", "deblock", "cdef",
                       "cdef_select", "lr_select", "lr_apply"],
            "io": "sources resident in HBM; per frame the blocks, superblock records, levels, CDEF choices, "
                  "restoration units and final picture copied to pinned host memory inside the timed region",
            "references": "the driver's own reconstructed, filtered pictures (closed loop from the key frame)",
            "engines": "2 (key, regular and ALTREF frames on engine 0, leaf frames on engine 1, each leaf beside the "
                       "next ALTREF's reconstruction), as the reference's driver runs a stream's mini-GOP on engine "
                       "encoders",
            "device": devname[0],
        },
        "data_sha256_ok": sha_ok,
        "steady_fps": None if steady is None else round(steady, 3),
        "pcie_inclusive_fps": pcie,
        "kernels_ms_per_launch": {k: round(v[1] / max(v[0], 1), 5) for k, v in stats.items()},
        "kernels_launches": {k: v[0] for k, v in stats.items()},
        "kernels_ms_note": "per-kernel means from a second coding of the stream, one frame at a time with every "
                           "launch timed; the timed region records events around the dominant kernel only",
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": TRAFFIC.get(dom),
            "algorithmic_bytes_per_launch": int(per_launch),
            "avg_launch_ms": round(avg_s * 1e3, 5),
            "launches": n_launch,
            "valu_busy_pct": VALU_BUSY.get(dom),
            "valu_busy_pct_other_kernels": {k: v for k, v in VALU_BUSY.items() if k != dom},
            "counters": "profiles/recon_intra_1080p_turbo_stream_pmc.md",
            "note": ("latency-bound dependency chain (superblocks in dataflow order, blocks in decode order), "
                     "integer VALU, no MFMA: the HBM fraction is small by construction" if dom == "recon_intra" else
                     "integer-VALU bound, no MFMA: the HBM fraction is small by construction"),
        },
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def _gen_frame(t):
    return synth.config_synth(2, seed=2024).frame(t)


def generate_frames(n):
    """The stream's first n frames (config-2 generator, seed 2024), made in
    worker processes (no GPU is touched before they finish)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(8, os.cpu_count() or 1, n))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        return list(ex.map(_gen_frame, range(n)))


def check_frame_hashes(planes):
    """Whether the generated frames are the ones whose SHA-256 is committed
    (tests/golden/synth_sha256.json): the same input on every host."""
    import hashlib
    path = os.path.join(ROOT, "tests", "golden", "synth_sha256.json")
    want = json.load(open(path))["config2_seed2024"]["frames"]
    for t, p in enumerate(planes[:len(want)]):
        h = hashlib.sha256()
        for a in p:
            h.update(a.tobytes())
        if h.hexdigest() != want[t]:
            return False
    return True


def _gen_batch_frame(job):
    i, t = job
    return synth.batch_synth(i).frame(t)


def batched_leg(nframes, torch):
    """The 16 uneven streams (4 x 1920x1080, 4 x 1280x720, 8 x 854x480), each
    nframes frames from its own seed, --turbo -q 125, sources resident in
    HBM, every frame's outputs downloaded. Plain: each stream through its own
    sequence driver, one stream after another (sequences on one GPU share
    its four streams). Batched: one mav1hip_batch over all 16, fed in step.
    Aggregate frames/s of both, and whether every frame's outputs (final
    picture, levels, block records) are the same."""
    import concurrent.futures as cf
    import hashlib
    import multiprocessing as mp
    S = len(synth.BATCH_STREAMS)
    jobs = [(i, t) for i in range(S) for t in range(nframes)]
    workers = max(1, min(16, os.cpu_count() or 1))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        frames = list(ex.map(_gen_batch_frame, jobs, chunksize=4))
    cfgs = [mav1hip.default_config(w, h, turbo=1, qindex=Q) for w, h in synth.BATCH_STREAMS]
    dev = []
    for i, (w, h) in enumerate(synth.BATCH_STREAMS):
        geo = synth.Geometry(w, h, cfgs[i].tile_cols_log2)
        dev.append([torch.from_numpy(synth.pad_frame(geo, frames[i * nframes + t])).to("cuda:0")
                    for t in range(nframes)])
    ys = [[frames[i * nframes + t][0] for t in range(nframes)] for i in range(S)]
    geos = [synth.Geometry(w, h, cfgs[i].tile_cols_log2) for i, (w, h) in enumerate(synth.BATCH_STREAMS)]
    del frames
    torch.cuda.synchronize()
    psnr_log = []

    def quality(i, r):
        """luma PSNR of a coded frame against its source (the check passes)"""
        fp = r.fp
        fb = fp.offset[2] + fp.stride[2] * fp.plane_h[2]
        rec = np.ctypeslib.as_array(C.cast(r.recon, C.POINTER(C.c_uint8)), (fb,))
        psnr_log.append(mav1hip.psnr(synth.unpad_frame(geos[i], rec)[0], ys[i][r.disp]))

    def digest(r):
        fp = r.fp
        nsb = fp.sb_cols * fp.sb_rows
        fb = fp.offset[2] + fp.stride[2] * fp.plane_h[2]
        h = hashlib.sha256()
        for ptr, n in ((r.recon, fb), (r.coefs, nsb * 6144 * 4), (r.blocks, nsb * 64 * 27 * 4)):
            h.update(C.string_at(ptr, n))
        return h.hexdigest()

    def plain(n, check=False):
        take = digest if check else (lambda r: None)
        seqs = [mav1hip.Sequence(c) for c in cfgs]
        out = [[] for _ in range(S)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, seq in enumerate(seqs):
            for t in range(n):
                while not seq.send_device(dev[i][t].data_ptr()):
                    out[i].append(take(seq.receive_raw(True)))
            while not seq.send(None):
                out[i].append(take(seq.receive_raw(True)))
            while True:
                r = seq.receive_raw(True)
                if r is None:
                    break
                out[i].append(take(r))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        for seq in seqs:
            seq.close()
        return dt, out

    knobs = autotune.load_tuned()
    weights = train_pruner.load_weights()

    def batched(n, check=False, tau=0.0):
        b = mav1hip.Batch(cfgs)
        b.set_knobs(knobs)
        if tau > 0:
            b.set_pruner(weights, tau)
        def take(i, r):
            if not check:
                return None
            quality(i, r)
            return digest(r)

        out = [[] for _ in range(S)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(n):
            for i in range(S):
                while not b.send_device(i, dev[i][t].data_ptr()):
                    out[i].append(take(i, b.receive_raw(i, True)))
            # the frames whose outputs are ready, without waiting
            for i in range(S):
                while True:
                    r = b.receive_raw(i, False)
                    if r is None:
                        break
                    out[i].append(take(i, r))
        for i in range(S):
            while not b.send(i, None):
                out[i].append(take(i, b.receive_raw(i, True)))
        for i in range(S):
            while True:
                r = b.receive_raw(i, True)
                if r is None:
                    break
                out[i].append(take(i, r))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        b.close()
        return dt, out

    tau = 0.1
    warm = min(nframes, 4)
    plain(warm)
    batched(warm)
    if weights is not None:
        batched(warm, tau=tau)
    # timed codings (no hashing inside the timed region), then the same
    # codings again with every frame's outputs hashed and their luma PSNR
    tp, _ = plain(nframes)
    tb, _ = batched(nframes)
    tq = batched(nframes, tau=tau)[0] if weights is not None else None
    _, op = plain(nframes, True)
    psnr_log.clear()
    _, ob = batched(nframes, True)
    psnr_b = float(np.mean(psnr_log))
    psnr_log.clear()
    psnr_q = None
    if weights is not None:
        _, oq = batched(nframes, True, tau=tau)
        psnr_q = float(np.mean(psnr_log))
    total = S * nframes
    return {
        "workload": f"16 streams (4 x 1920x1080, 4 x 1280x720, 8 x 854x480), {nframes} frames each from its own "
                    "seed, --turbo -q 125, sources resident in HBM, every frame's outputs downloaded",
        "frames": total,
        "plain_aggregate_fps": round(total / tp, 3),
        "batched_aggregate_fps": round(total / tb, 3),
        "speedup": round(tp / tb, 3),
        "outputs_equal": op == ob and all(len(o) == nframes for o in ob),
        "luma_psnr_db": round(psnr_b, 4),
        "pruned": None if tq is None else {
            "tau": tau,
            "aggregate_fps": round(total / tq, 3),
            "speedup_vs_plain": round(tp / tq, 3),
            "luma_psnr_db": round(psnr_q, 4),
            "luma_psnr_delta_db": round(psnr_q - psnr_b, 4),
            "weights": "mav1-encoder_amd/pruner_weights.f32 (train_pruner.py)",
            "note": "BD-rate needs bit counts (no entropy coder in this build): only the distortion at equal q",
        },
        "launch_shapes": knobs or "defaults (no mav1-encoder_amd/batch_tuned.json)",
        "note": "plain: each stream through its own sequence driver, one after another; batched: one launch per "
                "stage over the 16 streams' frames at the same coding position",
    }


def copy_calibration():
    """A device-to-device copy of 64 MiB, five times: FETCH_SIZE and
    WRITE_SIZE of this kernel over its known 64 MiB read and written give
    the counters' scale."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 64 << 20, "launches": 5}))


def cpu_baseline(cfg, padded, nframes, runs):
    """The oracle's serial driver and stages (oracle/seq.py over
    oracle/liboracle.so, whose independent superblocks, rows and units run
    on OpenMP threads) on the host cores: the stream's first nframes frames
    in closed loop, the median of `runs` runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import seq as oracle_seq
    times = []
    for _ in range(runs):
        o = oracle_seq.OracleSequence(mav1hip, cfg)
        t0 = time.perf_counter()
        for f in padded[:nframes]:
            o.send(f)
        o.end()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {
        "value": round(nframes / dt, 4),
        "unit": "frames/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the stream's first {nframes} frames (key frame, ALTREF, leaf) through the oracle's driver in "
                  f"closed loop, every stage of the GPU path (the restoration search dominates: ~4 s per 1080p "
                  f"frame on 8 cores), median of {runs} runs ({', '.join(f'{t:.1f}' for t in times)} s) on "
                  f"{cores} OpenMP threads ({os.cpu_count()} CPUs visible)",
    }


if __name__ == "__main__":
    main()
