"""Learned-initializer CNN kernels against the FP64 oracle (oracle/cnn.py).

Each layer is checked on the device's own inputs (its BF16 activations and the
BF16-rounded weights), so the comparison isolates that layer's kernel:
  forward conv  : |device - oracle| <= 1 bfloat16 step of the output (2^-8 relative)
                  plus 1e-3 of the layer's largest magnitude (FP32 accumulation order)
  weight / bias : relative 1e-3 of the gradient's norm (same BF16 inputs, FP32 vs
  gradients       FP64 accumulation)
  input gradient: 2^-8 relative + 1e-3 of the largest magnitude (BF16 store)
  MLP head      : FP32 vs FP64, 1e-4 relative
Training is checked to lower the loss and to be deterministic run to run.
"""

import numpy as np
import pytest

from helpers import record_error
from oracle import cnn as N

pytestmark = pytest.mark.gpu

B = 2


def _volumes(seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.arange(64, dtype=np.float32),) * 3, indexing="ij")
    out = []
    for _ in range(2 * B):
        v = np.zeros((64, 64, 64), np.float32)
        for _ in range(6):
            c = rng.uniform(8, 56, 3)
            s = rng.uniform(3, 8)
            v += rng.uniform(50, 200) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
                                               / (2 * s * s))
        v += rng.normal(0, 3, v.shape).astype(np.float32)
        out.append(v)
    return out


@pytest.fixture(scope="module")
def net_run(hip):
    net = hip.CNN(max_batch=4, seed=7)
    vols = _volumes(1)
    for b in range(B):
        net.set_input_raw(b, vols[2 * b], vols[2 * b + 1])
    out = net.forward(B)
    acts = [net.activation(i, B) for i in range(6)]
    offs, n = N.layout()
    assert n == net.nparams
    p = N.unflatten(net.params().astype(np.float64), offs)
    rng = np.random.default_rng(2)
    dout = rng.normal(0, 0.1, (B, 12)).astype(np.float32)
    net.backward(B, dout)
    grads = N.unflatten(net.grads().astype(np.float64), offs)
    dacts = [net.activation(i, B, gradient=True) for i in range(1, 6)]
    return dict(net=net, out=out, acts=acts, p=p, dout=dout, grads=grads, dacts=dacts)


def _close(dev, ref, rel_step=2.0 ** -8, floor=1e-3):
    dev = np.asarray(dev, np.float64)
    ref = np.asarray(ref, np.float64)
    tol = rel_step * np.abs(ref) + floor * max(np.max(np.abs(ref)), 1e-30)
    bad = np.abs(dev - ref) > tol
    return not bad.any(), float(np.max(np.abs(dev - ref) - tol))


def test_batched_inputs_are_the_per_pair_inputs(hip):
    """set_inputs packs n pairs in four launches; every slot must hold what
    set_input writes for that pair alone, and that is the general resampler's
    output (identity map onto the 64^3 grid spanning the image) z-scored, as
    set_input_raw packs it from host volumes."""
    rng = np.random.default_rng(5)
    shapes = [((40, 36, 30), (1.1, 0.9, 1.3)), ((33, 47, 29), (0.8, 1.0, 1.2)),
              ((64, 64, 64), (1.0, 1.0, 1.0))]
    fixed, moving = [], []
    for shp, sp in shapes:
        for lst in (fixed, moving):
            v = rng.normal(100.0, 20.0, shp[::-1]).astype(np.float32)
            lst.append(hip.DeviceImage.upload(v, spacing=sp, origin=tuple(rng.uniform(-5, 5, 3))))
    n = len(shapes)
    net = hip.CNN(max_batch=n + 1, seed=3)
    for b in range(n):
        net.set_input(b, fixed[b], moving[b])
    one = net.activation(0, n)
    net.set_inputs(0, fixed, moving)
    many = net.activation(0, n)
    assert np.array_equal(one.view(np.uint32), many.view(np.uint32))
    # a set starting at slot 1
    net.set_inputs(1, fixed[:1], moving[:1])
    assert np.array_equal(net.activation(0, 2)[1].view(np.uint32), many[0].view(np.uint32))
    # the general resampler's volumes packed from the host
    vols = []
    for img in (fixed[2], moving[2]):
        g = img.geometry()
        sp64 = [g.spacing[d] * g.size[d] / 64.0 for d in range(3)]
        org = [g.origin[d] + 0.5 * (sp64[d] - g.spacing[d]) for d in range(3)]
        r = hip.resample_linear(img, hip.Geometry.make((64, 64, 64), sp64, org))
        vols.append(r.to_host())
        r.free()
    net.set_input_raw(0, vols[0], vols[1])
    assert np.array_equal(net.activation(0, 1)[0].view(np.uint32), many[2].view(np.uint32))
    for img in fixed + moving:
        img.free()


def test_input_packing_is_z_scored_bf16(net_run):
    x = net_run["acts"][0]
    assert x.shape == (B, 64, 64, 64, 8)
    assert not x[..., 2:].any()
    for b in range(B):
        for c in range(2):
            ch = x[b, ..., c].astype(np.float64)
            assert abs(ch.mean()) < 1e-2 and abs(ch.std() - 1.0) < 1e-2


@pytest.mark.parametrize("layer", range(5))
def test_conv_forward_layer(net_run, layer):
    p, acts = net_run["p"], net_run["acts"]
    w = N.bf16(p[f"w{layer}"])
    ref = np.maximum(N.conv_forward(acts[layer].astype(np.float64), w, p[f"b{layer}"],
                                    N.STRIDES[layer]), 0.0)
    ok, excess = _close(acts[layer + 1], ref)
    assert ok, excess


def test_head_forward(net_run):
    p, acts = net_run["p"], net_run["acts"]
    last = acts[5].astype(np.float64)
    feat = last.reshape(B, -1, 128).mean(axis=1)
    _, out = N.head_forward(feat, p)
    assert np.allclose(net_run["out"], out, rtol=1e-4, atol=1e-4 * np.abs(out).max())


def test_head_backward(net_run):
    p, acts, g = net_run["p"], net_run["acts"], net_run["grads"]
    last = acts[5].astype(np.float64)
    feat = last.reshape(B, -1, 128).mean(axis=1)
    h, _ = N.head_forward(feat, p)
    gw1, gb1, gw2, gb2, dfeat = N.head_backward(feat, h, net_run["dout"].astype(np.float64), p)
    for dev, ref in [(g["fw1"], gw1), (g["fb1"], gb1), (g["fw2"], gw2), (g["fb2"], gb2)]:
        assert np.linalg.norm(dev - ref) <= 1e-4 * max(np.linalg.norm(ref), 1e-30)
    # the pooled-feature gradient reaches conv5's output as dfeat / 64 (before its mask)
    d5 = net_run["dacts"][4].astype(np.float64)
    want = np.broadcast_to(dfeat[:, None, None, None, :] / 64.0, d5.shape) * (acts[5] > 0)
    ok, excess = _close(d5, want)
    assert ok, excess


@pytest.mark.parametrize("layer", range(5))
def test_conv_backward_layer(net_run, layer):
    p, acts, g, dacts = net_run["p"], net_run["acts"], net_run["grads"], net_run["dacts"]
    dy = dacts[layer].astype(np.float64)   # gradient at conv `layer`'s output, masked
    w = N.bf16(p[f"w{layer}"])
    dw, db, dx = N.conv_backward(acts[layer].astype(np.float64), w, dy, N.STRIDES[layer])
    assert np.linalg.norm(g[f"w{layer}"] - dw) <= 1e-3 * max(np.linalg.norm(dw), 1e-30)
    assert np.linalg.norm(g[f"b{layer}"] - db) <= 1e-3 * max(np.linalg.norm(db), 1e-30)
    if layer > 0:
        # the device stores dx as BF16 and masks it with the previous ReLU
        dev = dacts[layer - 1].astype(np.float64)
        ref = dx * (acts[layer] > 0)
        ok, excess = _close(dev, ref)
        assert ok, excess


def test_training_lowers_the_loss_and_is_deterministic(hip):
    vols = _volumes(5)
    rng = np.random.default_rng(6)
    targets = rng.normal(0, 0.3, (B, 12)).astype(np.float32)
    runs = []
    for _ in range(2):
        net = hip.CNN(max_batch=B, seed=11)
        for b in range(B):
            net.set_input_raw(b, vols[2 * b], vols[2 * b + 1])
        losses = [net.train_step(B, targets, lr=1e-3) for _ in range(60)]
        runs.append((losses, net.params()))
    losses, params = runs[0]
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
    assert np.array_equal(params, runs[1][1])
    assert losses == runs[1][0]


def test_the_trained_network_beats_the_centered_start(hip):
    """What the network adds: on held-out affine pairs (+-30 deg, +-10%,
    +-20 mm) the same 30-iteration Mattes refinement succeeds (TRE <= 0.5 mm)
    far more often from the network's prediction than from the centered
    initializer. Weights: sitk-rs_amd/weights/initializer_cnn.npy, trained by
    sitk-rs_amd/train_initializer.py with this library's kernels."""
    import os

    import initializer as I

    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "sitk-rs_amd", "weights", "initializer_cnn.npy")
    net = hip.CNN(max_batch=16, seed=0)
    I.load(net, path)
    rec = I.evaluate(net, 16, seed=777, plain_pairs=2)
    assert rec["prediction_tre_median_mm"] < 15.0, rec
    assert rec["learned"]["success_fraction"] >= 0.75, rec
    assert rec["learned"]["success_fraction"] >= rec["centered_refine_only"]["success_fraction"] + 0.5, rec


def test_empty_cache_keeps_live_buffers(hip):
    img = hip.DeviceImage.upload(np.arange(4096, dtype=np.float32).reshape(16, 16, 16))
    tmp = hip.DeviceImage.upload(np.ones((8, 8, 8), np.float32))
    tmp.free()
    hip.empty_cache()
    assert np.array_equal(img.to_host(), np.arange(4096, dtype=np.float32).reshape(16, 16, 16))


def test_adam_steps_match_numpy(hip):
    """Two Adam steps (Kingma & Ba 2015) of the device against the same float32
    arithmetic in numpy: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
    p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps), with the device's
    own gradients of two different backward passes. The update of every
    parameter agrees to 1e-5 of the learning rate (float32 rounding)."""
    net = hip.CNN(max_batch=B, seed=13)
    vols = _volumes(9)
    for b in range(B):
        net.set_input_raw(b, vols[2 * b], vols[2 * b + 1])
    rng = np.random.default_rng(10)
    lr, b1, b2, eps = np.float32(2e-3), np.float32(0.9), np.float32(0.999), np.float32(1e-8)
    p = net.params()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2):
        net.forward(B)
        net.backward(B, rng.normal(0, 0.1, (B, 12)).astype(np.float32))
        g = net.grads()
        # most parameters have a gradient (zeros: the first layer's padded
        # input channels, units the ReLU switched off for every voxel)
        assert np.count_nonzero(g) > 0.5 * g.size
        net.adam(lr=float(lr), beta1=float(b1), beta2=float(b2), eps=float(eps))
        c1 = np.float32(1.0) - b1 ** np.float32(t)
        c2 = np.float32(1.0) - b2 ** np.float32(t)
        m = b1 * m + (np.float32(1.0) - b1) * g
        v = b2 * v + (np.float32(1.0) - b2) * g * g
        want = p - lr * (m / c1) / (np.sqrt(v / c2) + eps)
        got = net.params()
        err = np.max(np.abs((got - p) - (want - p)))
        assert err <= 1e-5 * lr, (t, err)
        # parameters with a zero gradient still move by the momentum of the
        # first step on the second
        assert np.count_nonzero(got - p) >= np.count_nonzero(g)
        p = got


def test_warp_inverse_labels_are_where_the_plain_optimizer_converges(hip):
    """The training labels are the synthetic warps' inverses. Registered by
    the plain refinement from those labels, 64 pairs of the training
    distribution move by a median of well under the prediction error the
    network reaches (3.3 mm median on held-out pairs): the plain optimizer's
    converged parameters and the warp inverses are the same labels to that
    precision."""
    import initializer as I

    gaps = I.label_gap(64)
    record_error("learned initializer: warp-inverse label vs plain-optimizer convergence (64 pairs)",
                 median_mm=float(np.median(gaps)), max_mm=float(np.max(gaps)))
    assert np.median(gaps) < 0.1, np.median(gaps)
    assert np.max(gaps) < 0.5, np.max(gaps)
