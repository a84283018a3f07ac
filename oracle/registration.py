"""Host restatement of the reference's registration loop for one level,
driving the CPU oracle's metric: RegularStepGradientDescent with physical-shift
scales and a once-estimated learning rate
(crates/sitk/src/registration/optimizer.rs:353-450,
crates/sitk/src/registration/scales.rs:659-826,
crates/sitk/src/registration/method.rs:3207-3233, 3357-3383), and gradient
descent with the golden-section line search (optimizer.rs:458-785). The two
optimizers are generic over the objective (rsgd, gd_line_search) so that the
reference's own optimizer fixtures pin them (tests/test_optimizer_reference.py).

Test infrastructure only, like the rest of the oracle: it checks the device
driver in sitk-rs_amd/csrc/registration.hip, and the benchmark's CPU-baseline
leg times it; nothing in the product path imports it.
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


def rsgd(evaluate, p0, learning_rate_of, minimum_step, iterations, grad_tol, scales=None,
         relaxation=0.5):
    """RegularStepGradientDescentOptimizer::optimize_with_lr (optimizer.rs:353-450):
    evaluate(p) -> (value, gradient); learning_rate_of(scaled gradient) -> the
    step-length scale. Returns (params, value, iterations, stop reason name)."""
    n = len(p0)
    scales = [1.0] * n if scales is None else list(scales)
    p = list(p0)
    v, g = evaluate(p)
    prev = [0.0] * n
    relax = 1.0
    taken = 0
    stop = "MaxIterations"
    while True:
        scaled = [g[k] / scales[k] for k in range(n)]
        gm = math.sqrt(_kahan([x * x for x in scaled]))
        if gm < grad_tol:
            stop = "GradientConverged"
            break
        if _kahan([a * b for a, b in zip(scaled, prev)]) < 0.0:
            relax *= relaxation
        step = relax * learning_rate_of(scaled)
        if step < minimum_step:
            stop = "StepTooSmall"
            break
        if taken >= iterations:
            break
        factor = step / gm
        p = [p[k] - factor * scaled[k] for k in range(n)]
        prev = scaled
        taken += 1
        v, g = evaluate(p)
    return p, v, taken, stop


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

    v, g = evaluate(list(p0))
    s0 = [g[k] / scales[k] for k in range(len(p0))]
    gm0 = math.sqrt(sum(x * x for x in s0))
    lr = est.learning_rate(s0) * gm0
    p, _, taken, stop = rsgd(evaluate, p0, lambda _s: lr, minimum_step, iterations, grad_tol,
                             scales, relaxation)
    return np.array(p), taken, stop


# GradientDescentLineSearchOptimizer (crates/sitk/src/registration/optimizer.rs:
# 587-785) with the golden-section search (optimizer.rs:458-545) and the window
# convergence monitor (registration/convergence.rs:35-95), restated on the host.
GOLDEN_PHI = 1.618034
GOLDEN_RESPHI = 2.0 - GOLDEN_PHI


def golden_section(eps, max_it, a, b, c, fb, counter, f):
    if counter[0] > max_it:
        return (c + a) / 2.0
    counter[0] += 1
    x = b + GOLDEN_RESPHI * (c - b) if c - b > b - a else b - GOLDEN_RESPHI * (b - a)
    if abs(c - a) < eps * (abs(b) + abs(x)):
        return (c + a) / 2.0
    fx = f(x)
    if fb is None:
        fb = f(b)
    if fx < fb:
        if c - b > b - a:
            return golden_section(eps, max_it, b, x, c, fx, counter, f)
        return golden_section(eps, max_it, a, x, b, fx, counter, f)
    if c - b > b - a:
        return golden_section(eps, max_it, a, b, x, fb, counter, f)
    return golden_section(eps, max_it, x, b, c, fb, counter, f)


class WindowConvergence:
    def __init__(self, window):
        self.window = window
        self.values = []
        self.total = 0.0

    def add(self, v):
        self.values.append(v)
        if len(self.values) > self.window:
            self.values.pop(0)
        self.total += abs(v)

    def value(self):
        if len(self.values) < self.window:
            return None
        if self.total == 0.0:
            return 0.0
        w = float(self.window)
        o0 = o1 = d0 = d1 = 0.0
        for n, e in enumerate(self.values):
            x = n / (w - 1.0)
            a, b = 1.0 - x, x
            y = e / self.total
            d = a * a + b * b
            o0 += a * a
            o1 += b * b
            d0 += a * a * a * y / d
            d1 += b * b * b * y / d
        c0 = d0 / o0 if o0 != 0.0 else 0.0
        c1 = d1 / o1 if o1 != 0.0 else 0.0
        return c0 - c1


def gd_line_search(evaluate, value_of, p0, learning_rate, iterations, scales=None, window=None,
                   min_convergence=None, epsilon=0.01, max_line_iterations=20, lower=0.0,
                   upper=5.0):
    """GradientDescentLineSearchOptimizer::run (optimizer.rs:688-785): the
    golden-section rate over [lr lower, lr, lr upper], carried to the next
    iteration; the optional window convergence monitor; the lowest-value
    iterate is returned. Returns (best params, best value, iterations, stop)."""
    n = len(p0)
    scales = [1.0] * n if scales is None else list(scales)
    p = list(p0)
    value, grad = evaluate(p)
    best, best_p = value, list(p)
    mon = WindowConvergence(window) if window is not None else None
    lr = learning_rate
    taken, stop = 0, "MaxIterations"
    while True:
        if value < best:
            best, best_p = value, list(p)
        if mon is not None:
            mon.add(value)
            cv = mon.value()
            if cv is not None and cv <= min_convergence:
                stop = "Converged"
                break
        if taken >= iterations:
            break
        d = [grad[k] / scales[k] for k in range(n)]
        base = p
        lr = golden_section(epsilon, max_line_iterations, lr * lower, lr, lr * upper, None, [0],
                            lambda x: value_of([base[k] - x * d[k] for k in range(n)]))
        p = [p[k] - lr * d[k] for k in range(n)]
        taken += 1
        value, grad = evaluate(p)
    return best_p, best, taken, stop


def register_gd_line_search(pb, geom, kind, p0, center, size, spacing, origin, iterations,
                            window=10, min_convergence=1e-6):
    """The estimated-rate line-search method (method.rs:1047-1057, 3361-3373):
    rate from the physical-shift estimator at the start, physical-shift
    scales. Returns (best params, iterations, stop reason name)."""
    t = O.make_transform(kind, p0, center=center)
    est = Estimator(t, size, spacing, origin)
    scales = est.scales()

    def evaluate(p):
        tt = O.make_transform(kind, p, center=center)
        v, d, _ = pb.mattes_evaluate(geom, tt)
        return v, d

    def value_of(p):
        return pb.mattes_value(geom, O.make_transform(kind, p, center=center))

    _, g = evaluate(list(p0))
    lr = est.learning_rate([g[k] / scales[k] for k in range(len(p0))])
    best_p, _, taken, stop = gd_line_search(evaluate, value_of, p0, lr, iterations, scales,
                                            window, min_convergence)
    return np.array(best_p), taken, stop
