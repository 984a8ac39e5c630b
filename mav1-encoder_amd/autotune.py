#!/usr/bin/env python3
"""Autotuner of the batched form's launch shapes (mav1hip_batch_set_knob).

Codes a short version of the batched workload (the 16 uneven streams of
synth.BATCH_STREAMS, --turbo -q 125, sources resident in HBM) once per
candidate setting and keeps the fastest; coordinate descent over the knobs,
each knob's candidates timed twice (best of two). The result is written to
batch_tuned.json beside this file, which bench.py applies. The knobs change
launch shapes only: every setting codes the same bytes (checked here on the
final pictures of every stream).

    python3 mav1-encoder_amd/autotune.py [--frames 12] [--out path]
"""
import argparse
import concurrent.futures as cf
import ctypes as C
import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import mav1hip  # noqa: E402
import synth  # noqa: E402

TUNED = os.path.join(HERE, "batch_tuned.json")

CANDIDATES = {
    "intra_wg_inter": [16, 32, 64, 128],
    "intra_wg_key": [32, 64, 128, 192],
    "inter_wg": [128, 256, 512],
}


def load_tuned(path=TUNED):
    """The tuned knobs ({} when there is no tuning file)."""
    try:
        with open(path) as f:
            return dict(json.load(f).get("knobs", {}))
    except (OSError, ValueError):
        return {}


def _gen(job):
    i, t = job
    return synth.batch_synth(i).frame(t)


def make_device_frames(torch, nframes):
    S = len(synth.BATCH_STREAMS)
    jobs = [(i, t) for i in range(S) for t in range(nframes)]
    with cf.ProcessPoolExecutor(max(1, min(16, os.cpu_count() or 1)), mp_context=mp.get_context("fork")) as ex:
        frames = list(ex.map(_gen, jobs, chunksize=4))
    cfgs = [mav1hip.default_config(w, h, turbo=1, qindex=125) for w, h in synth.BATCH_STREAMS]
    dev = []
    for i, (w, h) in enumerate(synth.BATCH_STREAMS):
        geo = synth.Geometry(w, h, cfgs[i].tile_cols_log2)
        dev.append([torch.from_numpy(synth.pad_frame(geo, frames[i * nframes + t])).to("cuda:0")
                    for t in range(nframes)])
    torch.cuda.synchronize()
    return cfgs, dev


def code_batched(torch, cfgs, dev, knobs, check=False):
    """Seconds to code every stream through one batch with these knobs, and
    (check) a digest of every stream's final pictures."""
    S = len(cfgs)
    n = len(dev[0])
    b = mav1hip.Batch(cfgs)
    b.set_knobs(knobs)
    digests = [hashlib.sha256() for _ in range(S)]

    def take(i, r):
        if check and r is not None:
            fp = r.fp
            digests[i].update(C.string_at(r.recon, fp.offset[2] + fp.stride[2] * fp.plane_h[2]))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(n):
        for i in range(S):
            while not b.send_device(i, dev[i][t].data_ptr()):
                take(i, b.receive_raw(i, True))
        for i in range(S):
            while True:
                r = b.receive_raw(i, False)
                if r is None:
                    break
                take(i, r)
    for i in range(S):
        while not b.send(i, None):
            take(i, b.receive_raw(i, True))
    for i in range(S):
        while True:
            r = b.receive_raw(i, True)
            if r is None:
                break
            take(i, r)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    b.close()
    return dt, [d.hexdigest() for d in digests]


def tune(nframes=12, out=TUNED, log=print):
    import torch
    cfgs, dev = make_device_frames(torch, nframes)
    frames = len(cfgs) * nframes
    best = {}
    code_batched(torch, cfgs, dev, best)  # warm-up
    _, ref = code_batched(torch, cfgs, dev, best, check=True)
    trials = []
    for knob, values in CANDIDATES.items():
        timed = {}
        for v in values:
            k = dict(best, **{knob: v})
            dt = min(code_batched(torch, cfgs, dev, k)[0] for _ in range(2))
            timed[v] = dt
            trials.append({"knobs": k, "fps": round(frames / dt, 2)})
            log(f"{knob}={v}: {frames / dt:.1f} frames/s")
        best[knob] = min(timed, key=timed.get)
    _, dig = code_batched(torch, cfgs, dev, best, check=True)
    if dig != ref:
        raise RuntimeError("the tuned launch shapes code different pictures")
    dt = min(code_batched(torch, cfgs, dev, best)[0] for _ in range(2))
    result = {
        "knobs": best,
        "fps": round(frames / dt, 2),
        "workload": f"{len(cfgs)} streams (synth.BATCH_STREAMS), {nframes} frames each, --turbo -q 125",
        "device": torch.cuda.get_device_name(0),
        "trials": trials,
    }
    with open(out, "w") as f:
        json.dump(result, f, indent=1)
    log(json.dumps({"knobs": best, "fps": result["fps"]}))
    return result


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=12)
    ap.add_argument("--out", default=TUNED)
    a = ap.parse_args()
    tune(a.frames, a.out)
