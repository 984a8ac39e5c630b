"""The pruner network's GPU kernels (BF16 MFMA convolutions, FP32 heads and
optimizer) against the FP64 oracle, and a training run on decisions of the
plain path.

Tolerance (floating point): the oracle rounds to BF16 where the GPU stores
BF16, so the differences come from FP32 accumulation order and the BF16
roundings it can flip (one BF16 ulp, 2^-8 relative, on a few elements):
logits within 1e-2 (1 + max |logit|); the loss within 1e-3 relative; each
parameter tensor's gradient within 2e-2 relative L2 error."""
import os
import sys

import numpy as np
import pytest

import mav1hip
import pruner_data
import synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pruner as orc  # noqa: E402

pytestmark = pytest.mark.gpu


def batch(rng, B, unlabeled=False):
    x = np.stack([rng.uniform(0, 1, (B, 64, 64)), rng.normal(0, 0.1, (B, 64, 64))], axis=-1).astype(np.float32)
    q = rng.uniform(0.1, 0.9, B).astype(np.float32)
    labels = np.stack([rng.integers(0, 4, B), rng.integers(0, 8, B)], axis=1).astype(np.int32)
    if unlabeled:
        labels[::3, 0] = -1
        labels[1::4, 1] = -1
    return x, q, labels


def check_against_oracle(net, params, x, q, labels):
    net.set_params(params)
    net.load(x, q, labels)
    logits, loss = net.forward()
    g = net.backward()
    olog, oloss, og = orc.forward_backward(params.astype(np.float64), x.astype(np.float64), q.astype(np.float64),
                                           labels, bf=True)
    assert np.abs(logits - olog).max() <= 1e-2 * (1 + np.abs(olog).max()), "logits differ"
    assert abs(loss - oloss) <= 1e-3 * abs(oloss), f"loss {loss} vs {oloss}"
    off = 0
    for name, shape in orc.LAYOUT:
        n = int(np.prod(shape))
        a, b = g[off:off + n].astype(np.float64), og[off:off + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"{name}: relative L2 error {err:.2e}", flush=True)
        assert err <= 2e-2, f"gradient of {name}: relative L2 error {err}"
        off += n


@pytest.mark.parametrize("B,seed", [(16, 1), (7, 2)])
def test_pruner_gradients_match_oracle(gpu, B, seed):
    print(f"\n[pruner] gradient check batch {B} seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    net = mav1hip.Pruner(max_batch=32)
    try:
        params = orc.init_params(rng)
        P = orc.unpack(params)
        for k in ("B1", "B2", "B3", "BQ", "BP", "BM"):
            P[k] = rng.normal(0, 0.05, P[k].shape)
        check_against_oracle(net, orc.pack(P).astype(np.float32), *batch(rng, B, unlabeled=seed == 2))
    finally:
        net.close()


def test_pruner_rejects_bad_labels(gpu):
    net = mav1hip.Pruner(max_batch=4)
    try:
        rng = np.random.default_rng(3)
        x, q, labels = batch(rng, 2)
        labels[0, 1] = 8
        with pytest.raises(mav1hip.HipError, match="labels"):
            net.load(x, q, labels)
    finally:
        net.close()


def stream_samples(w, h, seed, nframes, q):
    cfg = mav1hip.default_config(w, h, turbo=1, qindex=q, keyint=0)
    gen = synth.Synth(w, h, seed, n_objects=6, max_vel=8, noise_sigma=2.0)
    planes = [gen.frame(t) for t in range(nframes)]
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
    return pruner_data.samples_from_stream(cfg, planes, coded)


def test_pruner_training_on_plain_decisions(gpu):
    """Samples from the plain path's decisions on synthetic streams (two
    seeds for training, a third held out); 300 Adam steps on the GPU. The
    training loss must fall well below its start, and on the held-out
    stream the trained network's partition class must beat always
    predicting the training set's most frequent class."""
    print("\n[pruner] training on decisions of 640x360 turbo streams", flush=True)
    parts = [stream_samples(640, 360, s, 8, q) for s, q in ((11, 80), (12, 170))]
    x = np.concatenate([p[0] for p in parts])
    q = np.concatenate([p[1] for p in parts])
    y = np.concatenate([p[2] for p in parts])
    hx, hq, hy = stream_samples(640, 360, 13, 8, 125)
    print(f"{len(x)} training samples, {len(hx)} held out; partition classes {np.bincount(y[:, 0], minlength=4)}",
          flush=True)
    rng = np.random.default_rng(0)
    net = mav1hip.Pruner(max_batch=64)
    try:
        net.set_params(orc.init_params(rng).astype(np.float32))
        net.load(x[:64], q[:64], y[:64])
        _, first = net.forward()
        step = 0
        for epoch in range(300 // max(1, len(x) // 64)):
            order = rng.permutation(len(x))
            for i in range(0, len(x) - 63, 64):
                idx = order[i:i + 64]
                net.load(x[idx], q[idx], y[idx])
                step += 1
                net.train_step(step, lr=2e-3)
        net.load(x[:64], q[:64], y[:64])
        _, last = net.forward()
        print(f"{step} steps: loss on the first batch {first:.3f} -> {last:.3f}", flush=True)
        assert last < 0.7 * first
        pred = []
        for i in range(0, len(hx), 64):
            net.load(hx[i:i + 64], hq[i:i + 64])
            pred.append(net.forward()[0][:, :4].argmax(axis=1))
        pred = np.concatenate(pred)
        acc = float(np.mean(pred == hy[:, 0]))
        prior = float(np.mean(hy[:, 0] == np.bincount(y[:, 0], minlength=4).argmax()))
        print(f"held-out partition accuracy {acc:.3f}, most-frequent-class baseline {prior:.3f}", flush=True)
        assert acc > prior
    finally:
        net.close()
