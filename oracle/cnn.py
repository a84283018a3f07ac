"""CPU FP64 restatement of the learned-initializer CNN (test infrastructure).

The network has no counterpart in the reference; this module is the oracle
its device kernels (sitk-rs_amd/csrc/initializer.hip) are checked against:
3-D convolutions with 3^3 kernels and zero 'same' padding (forward, input
gradient, weight gradient, bias gradient), ReLU, global average pool and a
two-layer MLP head, all in float64. It is pinned by central finite
differences on small configurations (tests/test_oracle_cnn.py).

Layouts follow the device: activations channels-last (n, z, y, x, c); conv
weights (co, 27, ci) with tap = kz*9 + ky*3 + kx.
"""

from __future__ import annotations

import numpy as np

CHANNELS = (8, 16, 32, 64, 64, 128)   # input padded to 8 channels (2 used)
STRIDES = (1, 2, 2, 2, 2)
HIDDEN = 128
OUT = 12


def bf16(a) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layout(channels=CHANNELS, hidden=HIDDEN, out=OUT):
    """Offsets of every tensor in the flat parameter vector."""
    offs = {}
    o = 0
    for l in range(len(channels) - 1):
        ci, co = channels[l], channels[l + 1]
        offs[f"w{l}"] = (o, (co, 27, ci))
        o += co * 27 * ci
        offs[f"b{l}"] = (o, (co,))
        o += co
    feat = channels[-1]
    offs["fw1"] = (o, (hidden, feat)); o += hidden * feat
    offs["fb1"] = (o, (hidden,)); o += hidden
    offs["fw2"] = (o, (out, hidden)); o += out * hidden
    offs["fb2"] = (o, (out,)); o += out
    return offs, o


def unflatten(flat, offs):
    return {k: np.asarray(flat[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for k, (o, s) in offs.items()}


def im2col(x, stride):
    """x (n, d, d, d, c) -> patches (n, do, do, do, 27, c) with zero padding 1."""
    n, d, _, _, c = x.shape
    do = d // stride
    xp = np.zeros((n, d + 2, d + 2, d + 2, c))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    cols = np.empty((n, do, do, do, 27, c))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                tap = kz * 9 + ky * 3 + kx
                cols[:, :, :, :, tap] = xp[:, kz:kz + stride * do:stride, ky:ky + stride * do:stride,
                                           kx:kx + stride * do:stride]
    return cols


def conv_forward(x, w, b, stride):
    """Pre-activation output (n, do, do, do, co)."""
    cols = im2col(x, stride)
    n, do = cols.shape[0], cols.shape[1]
    y = cols.reshape(-1, 27 * x.shape[-1]) @ w.reshape(w.shape[0], -1).T + b
    return y.reshape(n, do, do, do, w.shape[0])


def conv_backward(x, w, dy, stride):
    """Given dL/d(pre-activation) dy: (dw, db, dx)."""
    cols = im2col(x, stride)
    ci = x.shape[-1]
    co = w.shape[0]
    dy2 = dy.reshape(-1, co)
    dw = (dy2.T @ cols.reshape(-1, 27 * ci)).reshape(co, 27, ci)
    db = dy2.sum(axis=0)
    dcols = (dy2 @ w.reshape(co, -1)).reshape(cols.shape)
    n, d = x.shape[0], x.shape[1]
    dxp = np.zeros((n, d + 2, d + 2, d + 2, ci))
    do = dy.shape[1]
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                tap = kz * 9 + ky * 3 + kx
                dxp[:, kz:kz + stride * do:stride, ky:ky + stride * do:stride,
                    kx:kx + stride * do:stride] += dcols[:, :, :, :, tap]
    return dw, db, dxp[:, 1:-1, 1:-1, 1:-1]


def head_forward(feat, p):
    h = np.maximum(feat @ p["fw1"].T + p["fb1"], 0.0)
    return h, h @ p["fw2"].T + p["fb2"]


def head_backward(feat, h, dout, p):
    gw2 = dout.T @ h
    gb2 = dout.sum(axis=0)
    dh = (dout @ p["fw2"]) * (h > 0)
    gw1 = dh.T @ feat
    gb1 = dh.sum(axis=0)
    dfeat = dh @ p["fw1"]
    return gw1, gb1, gw2, gb2, dfeat


def forward(x, p, strides=STRIDES):
    """Full FP64 forward. Returns (outputs, cache)."""
    acts = [x]
    for l, s in enumerate(strides):
        acts.append(np.maximum(conv_forward(acts[-1], p[f"w{l}"], p[f"b{l}"], s), 0.0))
    last = acts[-1]
    feat = last.reshape(last.shape[0], -1, last.shape[-1]).mean(axis=1)
    h, out = head_forward(feat, p)
    return out, (acts, feat, h)


def backward(dout, p, cache, strides=STRIDES):
    """Gradients of every parameter for dL/dout."""
    acts, feat, h = cache
    g = {}
    g["fw1"], g["fb1"], g["fw2"], g["fb2"], dfeat = head_backward(feat, h, dout, p)
    last = acts[-1]
    npos = int(np.prod(last.shape[1:4]))
    dy = np.broadcast_to(dfeat[:, None, None, None, :] / npos, last.shape).copy()
    for l in range(len(strides) - 1, -1, -1):
        dy = dy * (acts[l + 1] > 0)
        g[f"w{l}"], g[f"b{l}"], dx = conv_backward(acts[l], p[f"w{l}"], dy, strides[l])
        dy = dx
    return g


def mse(out, target):
    e = out - target
    return float(np.mean(e * e)), 2.0 * e / e.size
