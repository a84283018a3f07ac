"""Pins the CNN oracle (oracle/cnn.py) by what the mathematics fixes: central
finite differences of the loss for every parameter tensor on a small network,
the convolution against a direct sum, and the bfloat16 rounding against known
bit patterns."""

import numpy as np

from oracle import cnn as N

SMALL_CH = (2, 3, 4, 4, 3, 5)
SMALL_STRIDES = (1, 2, 2, 2, 1)


def _small_net(seed=0):
    offs, n = N.layout(SMALL_CH, hidden=6, out=4)
    rng = np.random.default_rng(seed)
    flat = rng.normal(0, 0.4, n)
    return offs, N.unflatten(flat, offs), flat


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3, 0.0], dtype=np.float32)
    got = N.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0          # tie -> even (1.0)
    assert got[2] == 1.0 + 2 ** -7                  # above the tie -> up
    assert got[3] == -2.5 and got[5] == 0.0
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_conv_forward_is_the_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 6, 6, 6, 3))
    w = rng.normal(size=(2, 27, 3))
    b = rng.normal(size=2)
    for stride in (1, 2):
        y = N.conv_forward(x, w, b, stride)
        do = 6 // stride
        for (z, yy, xx, o) in [(0, 0, 0, 0), (do - 1, 1, do - 1, 1), (2, do - 1, 0, 0)]:
            s = b[o]
            for kz in range(3):
                for ky in range(3):
                    for kx in range(3):
                        zi, yi, xi = z * stride + kz - 1, yy * stride + ky - 1, xx * stride + kx - 1
                        if 0 <= zi < 6 and 0 <= yi < 6 and 0 <= xi < 6:
                            s += w[o, kz * 9 + ky * 3 + kx] @ x[0, zi, yi, xi]
            assert abs(y[0, z, yy, xx, o] - s) < 1e-12


def test_backward_matches_finite_differences():
    offs, p, flat = _small_net(2)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 16, 16, 16, SMALL_CH[0]))
    target = rng.normal(size=(2, 4))

    def loss_of(fl):
        out, _ = N.forward(x, N.unflatten(fl, offs), SMALL_STRIDES)
        return N.mse(out, target)[0]

    out, cache = N.forward(x, p, SMALL_STRIDES)
    _, dout = N.mse(out, target)
    g = N.backward(dout, p, cache, SMALL_STRIDES)
    for name, (o, shape) in offs.items():
        size = int(np.prod(shape))
        for k in rng.choice(size, min(size, 3), replace=False):
            h = 1e-6
            fp, fm = flat.copy(), flat.copy()
            fp[o + k] += h
            fm[o + k] -= h
            fd = (loss_of(fp) - loss_of(fm)) / (2 * h)
            an = g[name].reshape(-1)[k]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(an), (name, k, fd, an)
