"""The pruner network's FP64 oracle (oracle/pruner.py), pinned three ways:
its analytic gradient against central finite differences and against torch
autograd (float64, an independent implementation of the same network), and
its BF16 rounding against torch's bfloat16 conversion. Also the training
labels' definitions on hand-built superblocks, and that the library's
parameter count matches the oracle's layout."""
import os
import sys

import numpy as np
import pytest

import mav1hip
import pruner_data

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import pruner as orc  # noqa: E402


def sample_batch(rng, B):
    x = np.stack([rng.uniform(0, 1, (B, 64, 64)), rng.normal(0, 0.1, (B, 64, 64))], axis=-1)
    q = rng.uniform(0.1, 0.9, B)
    labels = np.stack([rng.integers(0, 4, B), rng.integers(0, 8, B)], axis=1)
    return x, q, labels


def test_param_count_matches_library():
    assert orc.NUM_PARAMS == mav1hip.lib().mav1hip_pruner_num_params()


def test_gradient_finite_differences():
    rng = np.random.default_rng(5)
    params = orc.init_params(rng)
    P = orc.unpack(params)
    for k in ("B1", "B2", "B3"):  # non-zero biases so every bias gradient is exercised
        P[k] = rng.normal(0, 0.05, P[k].shape)
    params = orc.pack(P)
    x, q, labels = sample_batch(rng, 2)
    labels[1, 0] = -1  # a sample without a partition label
    _, _, g = orc.forward_backward(params, x, q, labels, bf=False)
    eps = 1e-6
    off = 0
    for name, shape in orc.LAYOUT:
        n = int(np.prod(shape))
        for i in rng.choice(n, size=min(n, 5), replace=False):
            j = off + int(i)
            pp, pm = params.copy(), params.copy()
            pp[j] += eps
            pm[j] -= eps
            fd = (orc.forward_backward(pp, x, q, labels, bf=False, grads=False)[1] -
                  orc.forward_backward(pm, x, q, labels, bf=False, grads=False)[1]) / (2 * eps)
            assert abs(fd - g[j]) <= 1e-5 * max(1.0, abs(fd)) + 1e-8, f"{name}[{i}]: fd {fd} analytic {g[j]}"
        off += n


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(6)
    params = orc.init_params(rng)
    x, q, labels = sample_batch(rng, 3)
    logits, loss, g = orc.forward_backward(params, x, q, labels, bf=False)
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in orc.unpack(params).items()}

    def conv_w(w):  # [co][tap][ci] -> [co][ci][ky][kx]
        co, _, ci = w.shape
        return w.reshape(co, 3, 3, ci).permute(0, 3, 1, 2)

    t = torch.tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)
    for k in ("1", "2", "3"):
        t = torch.relu(F.conv2d(t, conv_w(P["W" + k]), P["B" + k], stride=2, padding=1))
    h = t.mean(dim=(2, 3)) + torch.tensor(q, dtype=torch.float64)[:, None] * P["WQ"] + P["BQ"]
    lp, lm = h @ P["WP"] + P["BP"], h @ P["WM"] + P["BM"]
    lab = torch.tensor(labels)
    tl = (F.cross_entropy(lp, lab[:, 0], reduction="sum") + F.cross_entropy(lm, lab[:, 1], reduction="sum")) / 3
    tl.backward()
    assert abs(tl.item() - loss) < 1e-10
    np.testing.assert_allclose(logits, torch.cat([lp, lm], 1).detach().numpy(), rtol=1e-10, atol=1e-12)
    gt = orc.pack({k: v.grad.numpy() for k, v in P.items()})
    np.testing.assert_allclose(g, gt, rtol=1e-8, atol=1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-3, 1000), [0.0, 1.0, -2.5, 1 + 2 ** -8, 1 + 3 * 2 ** -9]])
    v = v.astype(np.float32)
    want = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(orc.bf16(v), want)


def _blk(bsize, is_inter=0, mv=(0, 0), ref=1, ref1=0):
    b = np.zeros(27, np.int32)
    b[2], b[3], b[8], b[9], b[23], b[24] = bsize, is_inter, mv[0], mv[1], ref, ref1
    return b


def test_labels_on_known_superblocks():
    # one 64x64 inter block on LAST with a motion vector
    one = np.stack([_blk(12, 1, (8, 0))])
    assert pruner_data.partition_class(one) == 0
    assert pruner_data.mode_class(one) == 3  # LAST, mv != 0
    # 32x32 quads: two intra, two compound with zero mv: tie -> the first class (intra)
    quads = np.stack([_blk(9), _blk(9), _blk(9, 1, ref1=7), _blk(9, 1, ref1=7)])
    assert pruner_data.partition_class(quads) == 1
    assert pruner_data.mode_class(quads) == 0
    # 16x16 and 8x8 blocks; the second reference covers most
    mix = np.stack([_blk(6, 1, (0, 4), ref=7)] * 12 + [_blk(3)] * 16)
    assert pruner_data.partition_class(mix) == 3
    assert pruner_data.mode_class(mix) == 5  # second reference, mv != 0
    assert pruner_data.partition_class(np.stack([_blk(6)] * 16)) == 2
