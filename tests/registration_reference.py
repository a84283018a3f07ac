"""Test-side restatement of the reference's host registration loop for one level,
driving the CPU oracle's metric: RegularStepGradientDescent with physical-shift
scales and a once-estimated learning rate
(crates/sitk/src/registration/optimizer.rs:353-450,
crates/sitk/src/registration/scales.rs:659-826,
crates/sitk/src/registration/method.rs:3207-3233, 3357-3383).

Test infrastructure only: checks the device driver in sitk-rs_amd/csrc/registration.hip.
"""

import math

import numpy as np

import oracle as O


def _kahan(values):
    s, c = 0.0, 0.0
    for x in values:
        ci = x - c
        t = s + ci
        c = (t - s) - ci
        s = t
    return s


def corner_points(size, spacing, origin):
    pts = []
    for i in range(8):
        idx = [(size[d] - 1) if (i >> d) & 1 else 0 for d in range(3)]
        pts.append([origin[d] + spacing[d] * idx[d] for d in range(3)])
    return pts


class Estimator:
    def __init__(self, t, size, spacing, origin, delta=0.01):
        self.delta = delta
        self.max_step = min(spacing)
        self.jac = [O.transform_jacobian(t, p) for p in corner_points(size, spacing, origin)]
        self.np = t.nparams

    def scales(self):
        shifts = [0.0] * self.np
        for j in self.jac:
            for p in range(self.np):
                sq = sum(j[r, p] * j[r, p] for r in range(3))
                shifts[p] = max(shifts[p], self.delta * math.sqrt(sq))
        nz = [s for s in shifts if s > np.finfo(float).eps]
        if not nz:
            return [1.0] * self.np
        mn = min(nz)
        return [((mn * mn) if s <= np.finfo(float).eps else s * s) / (self.delta ** 2)
                for s in shifts]

    def learning_rate(self, step):
        mx = max(abs(v) for v in step)
        if mx <= np.finfo(float).eps:
            return 1.0
        f = self.delta / mx
        small = np.asarray(step) * f
        m = max(float(np.sum((j @ small) ** 2)) for j in self.jac)
        scale = math.sqrt(m) / f
        return 1.0 if scale <= np.finfo(float).eps else self.max_step / scale


def register_rsgd(pb, geom, kind, p0, center, size, spacing, origin, minimum_step, iterations,
                  grad_tol, relaxation=0.5):
    """Returns (params, iterations, stop reason name)."""
    t = O.make_transform(kind, p0, center=center)
    est = Estimator(t, size, spacing, origin)
    scales = est.scales()

    def evaluate(p):
        tt = O.make_transform(kind, p, center=center)
        v, d, _ = pb.mattes_evaluate(geom, tt)
        return v, d

    p = list(p0)
    v, g = evaluate(p)
    s0 = [g[k] / scales[k] for k in range(len(p))]
    gm0 = math.sqrt(sum(x * x for x in s0))
    lr = est.learning_rate(s0) * gm0
    v, g = evaluate(p)
    prev = [0.0] * len(p)
    relax = 1.0
    taken = 0
    stop = "MaxIterations"
    while True:
        scaled = [g[k] / scales[k] for k in range(len(p))]
        gm = math.sqrt(_kahan([x * x for x in scaled]))
        if gm < grad_tol:
            stop = "GradientConverged"
            break
        if _kahan([a * b for a, b in zip(scaled, prev)]) < 0.0:
            relax *= relaxation
        step = relax * lr
        if step < minimum_step:
            stop = "StepTooSmall"
            break
        if taken >= iterations:
            break
        factor = step / gm
        p = [p[k] - factor * scaled[k] for k in range(len(p))]
        prev = scaled
        taken += 1
        v, g = evaluate(p)
    return np.array(p), taken, stop
