"""The learned initializer around the library's CNN: seeded training pairs made
on the device, training with the library's own forward/backward kernels,
prediction, and the refine-or-fall-back registration step.

A pair: fixed = sum of 40 Gaussian blobs (amplitude U[50, 200], sigma U[2, 8]
mm, centres uniform over the volume) + N(0, 5) noise on a 128^3 grid of 1 mm;
moving = fixed resampled (linear, 0 outside) through a seeded affine T_gt
(rotations U[-30, 30] deg, scales 1 + U[-0.1, 0.1], translations U[-20, 20] mm
about the volume centre). The registration transform that aligns the pair is
T = T_gt^-1 (fixed points to moving points), the network's label:
[(A - I) row-major (9), t / 10 mm (3)] of T about the fixed centre.
"""

from __future__ import annotations

import math
import time

import numpy as np

import sitk_hip as S
import workload as W

EDGE = 128
T_SCALE = 10.0


def pair_params(rng, edge=EDGE, max_deg=30.0, max_scale=0.1, max_mm=20.0, nblobs=40):
    ext = float(edge)
    centres = rng.uniform(0.0, ext, (nblobs, 3))
    amps = rng.uniform(50.0, 200.0, nblobs)
    sigmas = rng.uniform(2.0, 8.0, nblobs)
    seed = int(rng.integers(0, 2 ** 62))
    centre = np.full(3, 0.5 * (edge - 1))
    T_gt = W.random_affine(rng, max_deg, max_scale, max_mm, centre)
    return dict(edge=edge, centres=centres, amps=amps, sigmas=sigmas, seed=seed, T_gt=T_gt,
                centre=centre)


def make_pair(pp):
    """Device fixed and moving images of a pair."""
    e = pp["edge"]
    fixed = S.synth_blobs((e, e, e), centres=pp["centres"], amplitudes=pp["amps"],
                          sigmas=pp["sigmas"], noise=5.0, seed=pp["seed"])
    moving = S.resample_linear(fixed, fixed.geometry(), pp["T_gt"].stage().reshape(1, 12))
    return fixed, moving


def label_of(T: W.AffineMap) -> np.ndarray:
    return np.concatenate([(T.A - np.eye(3)).reshape(9), T.t / T_SCALE]).astype(np.float32)


def affine_of(y, centre) -> W.AffineMap:
    y = np.asarray(y, dtype=np.float64)
    return W.AffineMap(np.eye(3) + y[:9].reshape(3, 3), y[9:] * T_SCALE, np.asarray(centre, float))


def train(net: S.CNN, steps: int, batch: int, lr: float = 1e-3, seed: int = 0, log=None,
          log_every: int = 100):
    """Online training: every step draws `batch` fresh pairs. Returns the losses."""
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        targets = np.empty((batch, 12), np.float32)
        for b in range(batch):
            pp = pair_params(rng)
            f, m = make_pair(pp)
            net.set_input(b, f, m)
            targets[b] = label_of(pp["T_gt"].inverse())
            f.free()
            m.free()
        losses.append(net.train_step(batch, targets, lr=lr))
        if log and (step + 1) % log_every == 0:
            log(f"step {step + 1}/{steps} loss {np.mean(losses[-log_every:]):.5f} "
                f"({time.perf_counter() - t0:.1f}s)")
    return losses


def predict(net: S.CNN, pairs) -> list:
    """pairs: list of (fixed, moving, centre). Returns one AffineMap per pair."""
    out = []
    for i in range(0, len(pairs), net.max_batch):
        chunk = pairs[i:i + net.max_batch]
        net.set_inputs(0, [f for f, _, _ in chunk], [m for _, m, _ in chunk])
        y = net.forward(len(chunk))
        out.extend(affine_of(y[b], chunk[b][2]) for b in range(len(chunk)))
    return out


def affine_params(T: W.AffineMap):
    return list(T.A.reshape(9)) + list(T.t)


def refine_config(iterations=30, bins=50, sampling=0, percentage=1.0, seed=0):
    """The plain Mattes refinement: RegularStepGradientDescent with
    physical-shift scales, one level at full resolution."""
    return S.registration_config(metric=S.MATTES, bins=bins, sampling=sampling,
                                 sampling_percentage=percentage, seed=seed,
                                 optimizer=S.REGULAR_STEP, minimum_step=1e-4,
                                 iterations=iterations, gradient_tolerance=1e-8,
                                 learning_rate_mode=S.LR_ONCE)


def tre(T_est: W.AffineMap, pp) -> float:
    """Mean target registration error (mm) against T_gt^-1 over every 8th voxel
    of the pair's grid (1 mm spacing, origin 0)."""
    e = pp["edge"]
    g = np.stack(np.meshgrid(*(np.arange(0, e, 8),) * 3, indexing="ij"), -1).reshape(-1, 3)
    g = g.astype(np.float64)
    return float(np.linalg.norm(T_est.apply(g) - pp["T_gt"].inverse().apply(g), axis=1).mean())


def label_gap(n_pairs: int = 64, seed: int = 777, iterations: int = 30) -> np.ndarray:
    """How far the plain optimizer moves a training label: each pair of the
    training distribution registered by the batched plain refinement
    (refine_config_fast: Mattes 50 bins at 25%, RSGD) starting from the
    warp-inverse label; returns the mean TRE (mm, every 8th voxel) between the
    label and the transform the optimizer converges to, per pair. Small gaps
    mean the warp inverse is the label the plain optimizer would give."""
    rng = np.random.default_rng(seed)
    pps = [pair_params(rng) for _ in range(n_pairs)]
    pairs = [make_pair(pp) for pp in pps]
    labels = [pp["T_gt"].inverse() for pp in pps]
    got, _ = S.register_batch([f for f, _ in pairs], [m for _, m in pairs],
                              [refine_config_fast(iterations)] * n_pairs, [S.AFFINE] * n_pairs,
                              [affine_params(T) for T in labels], [pp["centre"] for pp in pps])
    e = pps[0]["edge"]
    g = np.stack(np.meshgrid(*(np.arange(0, e, 8),) * 3, indexing="ij"), -1).reshape(-1, 3)
    g = g.astype(np.float64)
    gaps = []
    for p, T, pp in zip(got, labels, pps):
        opt = W.AffineMap(np.asarray(p[:9]).reshape(3, 3), np.asarray(p[9:]), pp["centre"])
        gaps.append(float(np.linalg.norm(opt.apply(g) - T.apply(g), axis=1).mean()))
    for f, m in pairs:
        f.free()
        m.free()
    return np.asarray(gaps)


def save(net: S.CNN, path: str) -> None:
    np.save(path, net.params())


def load(net: S.CNN, path: str) -> None:
    net.set_params(np.load(path, allow_pickle=False))


def train_schedule(net: S.CNN, steps: int, batch: int, lr: float = 1e-3, final_lr: float = 1e-4,
                   seed: int = 0, log=None, log_every: int = 250):
    """Online training with a cosine learning-rate decay from lr to final_lr."""
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        a = final_lr + 0.5 * (lr - final_lr) * (1.0 + math.cos(math.pi * step / max(steps - 1, 1)))
        targets = np.empty((batch, 12), np.float32)
        for b in range(batch):
            pp = pair_params(rng)
            f, m = make_pair(pp)
            net.set_input(b, f, m)
            targets[b] = label_of(pp["T_gt"].inverse())
            f.free()
            m.free()
        losses.append(net.train_step(batch, targets, lr=a))
        if log and (step + 1) % log_every == 0:
            log(f"step {step + 1}/{steps} lr {a:.2e} loss {np.mean(losses[-log_every:]):.5f} "
                f"({time.perf_counter() - t0:.1f}s)")
    return losses


def plain_config(iterations=100):
    """The plain method for these pairs: three-level pyramid (shrink 4/2/1,
    sigma 2/1/0 mm), Mattes 50 bins at 25% sampling, RegularStepGradientDescent
    with physical-shift scales from the centered initializer."""
    return S.registration_config(metric=S.MATTES, bins=50, sampling=S.SAMPLING_RANDOM,
                                 sampling_percentage=0.25, seed=1, optimizer=S.REGULAR_STEP,
                                 minimum_step=1e-4, iterations=iterations, gradient_tolerance=1e-8,
                                 learning_rate_mode=S.LR_ONCE,
                                 shrink_factors=((4, 4, 4), (2, 2, 2), (1, 1, 1)),
                                 sigmas=(2.0, 1.0, 0.0))


def refine_config_fast(iterations=30):
    """The refinement after the prediction: Mattes 50 bins at 25% sampling, one
    level at full resolution, `iterations` RegularStepGradientDescent steps."""
    return S.registration_config(metric=S.MATTES, bins=50, sampling=S.SAMPLING_RANDOM,
                                 sampling_percentage=0.25, seed=1, optimizer=S.REGULAR_STEP,
                                 minimum_step=1e-4, iterations=iterations, gradient_tolerance=1e-8,
                                 learning_rate_mode=S.LR_ONCE)


def evaluate(net: S.CNN, n_pairs: int, seed: int = 12345, refine_iterations: int = 30,
             plain_pairs: int = 16, log=None):
    """Held-out evaluation of learned initialisation + batched refinement
    (with the centered-initializer fallback) against the plain method one
    pair at a time. Returns a dict of rates and success fractions."""
    rng = np.random.default_rng(seed)
    pps = [pair_params(rng) for _ in range(n_pairs)]
    pairs = [make_pair(pp) for pp in pps]
    centres = [pp["centre"] for pp in pps]
    cfg = refine_config_fast(refine_iterations)
    S.backend_init(0)
    t0 = time.perf_counter()
    pred = predict(net, [(f, m, c) for (f, m), c in zip(pairs, centres)])
    t_pred = time.perf_counter() - t0
    p0 = [affine_params(T) for T in pred]
    # the refinement, also probing the centered initializer's value (identity
    # about the centre) on the same prepared level
    ident = [list(np.eye(3).reshape(9)) + [0.0, 0.0, 0.0]] * n_pairs
    got, res, v_ident = S.register_batch([f for f, _ in pairs], [m for _, m in pairs],
                                         [cfg] * n_pairs, [S.AFFINE] * n_pairs, p0, centres,
                                         probe=ident)
    worse = [i for i in range(n_pairs) if res[i].value > v_ident[i]]
    if worse:
        got_f, res_f = S.register_batch([pairs[i][0] for i in worse], [pairs[i][1] for i in worse],
                                        [cfg] * len(worse), [S.AFFINE] * len(worse),
                                        [ident[i] for i in worse], [centres[i] for i in worse])
        for k, i in enumerate(worse):
            got[i] = got_f[k]
    t_learned = time.perf_counter() - t0

    def tre_params(p, pp):
        return tre(W.AffineMap(np.asarray(p[:9]).reshape(3, 3), np.asarray(p[9:]), pp["centre"]), pp)

    tre_l = np.array([tre_params(g, pp) for g, pp in zip(got, pps)])
    tre_pred = np.array([tre(T, pp) for T, pp in zip(pred, pps)])
    k = min(plain_pairs, n_pairs)
    pc = plain_config()
    t1 = time.perf_counter()
    plain = [S.register(pairs[i][0], pairs[i][1], pc, S.AFFINE, ident[i], center=centres[i])
             for i in range(k)]
    t_plain = time.perf_counter() - t1
    tre_p = np.array([tre_params(g, pp) for (g, _), pp in zip(plain, pps[:k])])
    # the plain method's fixed-budget refinement alone, batched (what the network adds)
    got_c, _ = S.register_batch([f for f, _ in pairs], [m for _, m in pairs], [cfg] * n_pairs,
                                [S.AFFINE] * n_pairs, ident, centres)
    tre_c = np.array([tre_params(g, pp) for g, pp in zip(got_c, pps)])
    out = {
        "pairs": n_pairs,
        "prediction_tre_median_mm": float(np.median(tre_pred)),
        "learned": {"seconds": t_learned, "predict_seconds": t_pred,
                    "registrations_per_s": n_pairs / t_learned,
                    "success_fraction": float(np.mean(tre_l <= 0.5)),
                    "fallbacks": len(worse)},
        "centered_refine_only": {"success_fraction": float(np.mean(tre_c <= 0.5))},
        "plain": {"pairs": k, "seconds": t_plain, "registrations_per_s": k / t_plain,
                  "success_fraction": float(np.mean(tre_p <= 0.5))},
    }
    out["learned"]["successful_per_s"] = out["learned"]["success_fraction"] * n_pairs / t_learned
    out["plain"]["successful_per_s"] = out["plain"]["success_fraction"] * k / t_plain
    out["speedup_successful"] = (out["learned"]["successful_per_s"] /
                                 max(out["plain"]["successful_per_s"], 1e-12))
    for f, m in pairs:
        f.free()
        m.free()
    if log:
        log(str(out))
    return out
