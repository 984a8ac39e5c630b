#!/usr/bin/env python3
"""Trains the learned pruner on the GPU from the plain path's decisions.

Training data: the 64x64 superblocks of synthetic streams (seeds 900 and up,
disjoint from the tests' and the benchmarks'), coded by the sequence driver
with --turbo at q 80, 125, 170 and 215, at 1280x720 and 854x480; labels and
inputs as pruner_data.py defines them. One stream per (size, q) is held out
for the reported accuracy. The network trains with Adam on the GPU through
mav1hip_pruner_* (BF16 MFMA forward and backward kernels).

Writes pruner_weights.f32 (the FP32 parameter vector, little-endian) and
pruner_weights.json (data, training and held-out figures) beside this file.

    python3 mav1-encoder_amd/train_pruner.py [--frames 10] [--steps 3000]
"""
import argparse
import concurrent.futures as cf
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import mav1hip  # noqa: E402
import pruner_data  # noqa: E402
import synth  # noqa: E402

WEIGHTS = os.path.join(HERE, "pruner_weights.f32")
QS = (80, 125, 170, 215)
SIZES = ((1280, 720), (854, 480))


def load_weights(path=WEIGHTS):
    """The trained parameter vector, or None."""
    if not os.path.exists(path):
        return None
    return np.fromfile(path, dtype="<f4")


def _frames(job):
    w, h, seed, n = job
    g = synth.Synth(w, h, seed, n_objects=6, max_vel=12, noise_sigma=2.5)
    return [g.frame(t) for t in range(n)]


def code_stream(cfg, planes):
    seq = mav1hip.Sequence(cfg)
    coded = []
    try:
        for p in planes:
            while not seq.send(p):
                coded.append(seq.receive(True))
        while not seq.send(None):
            coded.append(seq.receive(True))
        while True:
            r = seq.receive(True)
            if r is None:
                break
            coded.append(r)
    finally:
        seq.close()
    return coded


def collect(nframes, streams_per_point, seed0=900, log=print):
    """(training x, q, labels), (held-out x, q, labels)."""
    jobs, meta = [], []
    seed = seed0
    for (w, h) in SIZES:
        for q in QS:
            for k in range(streams_per_point + 1):
                jobs.append((w, h, seed, nframes))
                meta.append((w, h, q, k == streams_per_point))
                seed += 1
    with cf.ProcessPoolExecutor(max(1, min(16, os.cpu_count() or 1)), mp_context=mp.get_context("fork")) as ex:
        frames = list(ex.map(_frames, jobs))
    train, held = [], []
    for (w, h, q, hold), planes in zip(meta, frames):
        cfg = mav1hip.default_config(w, h, turbo=1, qindex=q)
        s = pruner_data.samples_from_stream(cfg, planes, code_stream(cfg, planes))
        (held if hold else train).append(s)
    cat = lambda parts: tuple(np.concatenate([p[i] for p in parts]) for i in range(3))
    tr, ho = cat(train), cat(held)
    log(f"{len(tr[0])} training superblocks, {len(ho[0])} held out; partition classes "
        f"{np.bincount(tr[2][:, 0], minlength=4).tolist()}, mode classes {np.bincount(tr[2][:, 1], minlength=8).tolist()}")
    return tr, ho


def accuracy(net, data, batch=256):
    x, q, y = data
    pred = []
    for i in range(0, len(x), batch):
        net.load(x[i:i + batch], q[i:i + batch])
        pred.append(net.forward()[0])
    lg = np.concatenate(pred)
    return float(np.mean(lg[:, :4].argmax(1) == y[:, 0])), float(np.mean(lg[:, 4:].argmax(1) == y[:, 1])), lg


def train(nframes=10, steps=3000, streams_per_point=2, batch=256, lr=2e-3, seed=0, out=WEIGHTS, log=print):
    t0 = time.time()
    tr, ho = collect(nframes, streams_per_point, log=log)
    rng = np.random.default_rng(seed)
    net = mav1hip.Pruner(max_batch=batch)
    try:
        # He-normal convolutions and heads, zero biases (the layout of csrc/pruner.hip)
        params = []
        for shape, fan in (((32, 9, 2), 18), ((32,), 0), ((64, 9, 32), 288), ((64,), 0), ((64, 9, 64), 576), ((64,), 0)):
            params.append(rng.standard_normal(shape).ravel() * np.sqrt(2.0 / fan) if fan else np.zeros(shape))
        params += [rng.standard_normal(64) * 0.1, np.zeros(64), rng.standard_normal(256) / 8, np.zeros(4),
                   rng.standard_normal(512) / 8, np.zeros(8)]
        p0 = np.concatenate(params).astype(np.float32)
        assert p0.size == net.n
        net.set_params(p0)
        x, q, y = tr
        for step in range(1, steps + 1):
            idx = rng.integers(0, len(x), batch)
            net.load(x[idx], q[idx], y[idx])
            net.train_step(step, lr=lr * (0.1 ** (step / steps)))
            if step % 500 == 0 or step == steps:
                net.load(x[idx], q[idx], y[idx])
                log(f"step {step}: batch loss {net.forward()[1]:.3f}")
        pa, ma, _ = accuracy(net, ho)
        prior = float(np.mean(ho[2][:, 0] == np.bincount(y[:, 0], minlength=4).argmax()))
        params = net.get_params()
    finally:
        net.close()
    params.astype("<f4").tofile(out)
    meta = {
        "samples": int(len(tr[0])), "held_out": int(len(ho[0])), "frames_per_stream": nframes,
        "sizes": [list(s) for s in SIZES], "q": list(QS), "steps": steps, "batch": batch,
        "held_out_partition_accuracy": round(pa, 4), "held_out_mode_accuracy": round(ma, 4),
        "held_out_most_frequent_partition_class": round(prior, 4),
        "partition_classes": np.bincount(tr[2][:, 0], minlength=4).tolist(),
        "seconds": round(time.time() - t0, 1),
    }
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))
    return meta


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    train(a.frames, a.steps, out=a.out)
