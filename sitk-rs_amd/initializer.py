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
        for b, (f, m, _) in enumerate(chunk):
            net.set_input(b, f, m)
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
