"""ORACLE — test infrastructure only.

The learned pruner's network (mav1-encoder_amd/csrc/pruner.hip) restated in
float64 numpy: forward pass, loss and the analytic gradient of every
parameter. Only tests/ may use it.

`bf16=True` rounds (to nearest even, through float32) exactly where the GPU
stores BF16: the input, the conv weights, each post-ReLU activation and
each pre-activation gradient; everything else is float64, so the GPU's only
remaining differences are its FP32 accumulation order (the tests' tolerance).
`bf16=False` is the plain float64 network, whose gradient
tests/test_pruner_oracle.py checks against central finite differences and
against torch autograd.

There is no reference implementation (the reference has no pruner); the
network is this build's own design.
"""
import numpy as np

C0, C1, C2, C3, NP, NM, NF = 2, 32, 64, 64, 4, 8, 64
LAYOUT = [("W1", (C1, 9, C0)), ("B1", (C1,)), ("W2", (C2, 9, C1)), ("B2", (C2,)), ("W3", (C3, 9, C2)), ("B3", (C3,)),
          ("WQ", (NF,)), ("BQ", (NF,)), ("WP", (NF, NP)), ("BP", (NP,)), ("WM", (NF, NM)), ("BM", (NM,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(vec):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(vec[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in LAYOUT])


def init_params(rng):
    """He-normal convolutions and heads, zero biases."""
    d = {}
    for name, shape in LAYOUT:
        if name[0] == "W" and len(shape) == 3:
            d[name] = rng.standard_normal(shape) * np.sqrt(2.0 / (9 * shape[2]))
        elif name in ("WP", "WM"):
            d[name] = rng.standard_normal(shape) * np.sqrt(1.0 / NF)
        elif name == "WQ":
            d[name] = rng.standard_normal(shape) * 0.1
        else:
            d[name] = np.zeros(shape)
    return pack(d)


def bf16(a):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _cols(x):
    """x [B, H, W, C] -> the 3x3 stride-2 pad-1 patches [B, H/2, W/2, 9, C],
    taps in (ky, kx) order."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C))
    xp[:, 1:H + 1, 1:W + 1] = x
    Ho, Wo = H // 2, W // 2
    cols = np.empty((B, Ho, Wo, 9, C))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, 3 * ky + kx] = xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2]
    return cols


def _conv(x, w, b):
    return np.einsum("bhwtc,otc->bhwo", _cols(x), w) + b


def _wgrad(x, dpre):
    return np.einsum("bhwo,bhwtc->otc", dpre, _cols(x))


def _dgrad(dpre, w, H, W):
    """Gradient with respect to the convolution's input [B, H, W, Ci]."""
    B = dpre.shape[0]
    dcols = np.einsum("bhwo,otc->bhwtc", dpre, w)
    Ho, Wo = H // 2, W // 2
    dxp = np.zeros((B, H + 2, W + 2, w.shape[2]))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2] += dcols[:, :, :, 3 * ky + kx]
    return dxp[:, 1:H + 1, 1:W + 1]


def _ce(logits, y):
    """Softmax cross-entropy per row and its gradient (rows with y < 0: 0)."""
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    s = e.sum(axis=1, keepdims=True)
    p = e / s
    valid = y >= 0
    yy = np.where(valid, y, 0)
    loss = np.where(valid, np.log(s[:, 0]) + m[:, 0] - logits[np.arange(len(y)), yy], 0.0)
    g = p.copy()
    g[np.arange(len(y)), yy] -= 1.0
    g[~valid] = 0.0
    return loss, g


def forward_backward(params, x, q, labels, bf=True, grads=True):
    """params: the flat parameter vector; x [B, 64, 64, 2]; q [B];
    labels [B, 2]. Returns (logits [B, 12], mean loss, flat gradient or
    None)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    B = x.shape[0]
    x0 = r(x)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    y1 = r(np.maximum(_conv(x0, W1, P["B1"]), 0))
    y2 = r(np.maximum(_conv(y1, W2, P["B2"]), 0))
    y3 = r(np.maximum(_conv(y2, W3, P["B3"]), 0))
    h = y3.mean(axis=(1, 2)) + np.asarray(q, np.float64)[:, None] * P["WQ"] + P["BQ"]
    lp = h @ P["WP"] + P["BP"]
    lm = h @ P["WM"] + P["BM"]
    labels = np.asarray(labels)
    Lp, gp = _ce(lp, labels[:, 0])
    Lm, gm = _ce(lm, labels[:, 1])
    loss = float((Lp + Lm).sum() / B)
    logits = np.concatenate([lp, lm], axis=1)
    if not grads:
        return logits, loss, None
    gp /= B
    gm /= B
    G = {"WP": h.T @ gp, "BP": gp.sum(0), "WM": h.T @ gm, "BM": gm.sum(0)}
    dh = gp @ P["WP"].T + gm @ P["WM"].T
    G["WQ"] = (dh * np.asarray(q, np.float64)[:, None]).sum(0)
    G["BQ"] = dh.sum(0)
    d3 = r(np.broadcast_to((dh / 64.0)[:, None, None, :], y3.shape) * (y3 > 0))
    G["W3"], G["B3"] = _wgrad(y2, d3), d3.sum((0, 1, 2))
    d2 = r(_dgrad(d3, W3, 16, 16) * (y2 > 0))
    G["W2"], G["B2"] = _wgrad(y1, d2), d2.sum((0, 1, 2))
    d1 = r(_dgrad(d2, W2, 32, 32) * (y1 > 0))
    G["W1"], G["B1"] = _wgrad(x0, d1), d1.sum((0, 1, 2))
    return logits, loss, pack(G)
