"""Benchmark of the registration metric-evaluation path on one MI355X.

Workload (one step): one Mattes mutual-information value + derivative
evaluation -- histogram pass, device tail, derivative pass -- at the finest level
of the single-large-pair configuration: fixed and moving 256x256x160 float32
(spacing 0.9x0.9x1.5 mm, 40 seeded blobs, +-20% bias field, moving = fixed
through a seeded affine), 50 histogram bins, 20% random sampling (2,097,152
SplitMix64 draws), affine transform (12 parameters). Inputs are resident in HBM
before the timed region; every step ends with the 13 result doubles on the host.

value = whole-step throughput in GB/s of algorithmic (unique) bytes: each of the
two sample passes reads, per sample, its uint32 fixed index and uint8 fixed
Parzen bin (5 B) and reads the moving volume once (4 B per voxel):
    bytes/step = 2 * (n_samples * 5 + moving_voxels * 4).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32|f64]
                       [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sitk-rs_amd"))
sys.path.insert(0, ROOT)

METRIC = "registrations/s at ≤0.5mm TRE, learned-init+batched vs plain; metric-eval GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    """(CPU model, physical cores, logical CPUs this process may use)."""
    model, cores, cur = None, set(), {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in list(fh) + [""]:
                if not line.strip():
                    if "physical id" in cur and "core id" in cur:
                        cores.add((cur["physical id"], cur["core id"]))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if model is None and k.strip() == "model name":
                    model = v.strip()
    except OSError:
        pass
    return model, (len(cores) or None), len(os.sched_getaffinity(0))


def cpu_baseline(fixed, moving, idx, stage_affine, bins=50, evals=2):
    """The oracle (serial FP64 restatement of the reference's host Mattes,
    crates/sitk/src/registration/mattes.rs:522-676) on the same level: `evals`
    full value+derivative evaluations, timed on one host core. Returns the
    median time, the count, the oracle's (value, derivative, valid) and the
    transform it evaluated."""
    import oracle as O

    pb = O.MetricProblem(fixed.data, fixed.spacing, fixed.origin, moving.data, moving.spacing,
                         moving.origin, indices=idx)
    g = pb.mattes_geometry(bins)
    A, off = stage_affine
    t = O.make_transform(O.AFFINE, list(A.reshape(9)) + [0.0, 0.0, 0.0])
    t.offset[:] = [float(v) for v in off]
    times, res = [], None
    for _ in range(evals):
        t0 = time.perf_counter()
        res = pb.mattes_evaluate(g, t)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times), res, t


def parity_record(S, O, t, value, mom, valid, want):
    """Device value and derivative at the bench's evaluation point against the
    oracle's: relative value error, relative derivative error (vector norm),
    and whether the valid counts agree."""
    wv, wd, wn = want
    j0 = O.transform_jacobian(t, [0, 0, 0])
    c = np.stack([O.transform_jacobian(t, e) - j0 for e in np.eye(3)])
    deriv = S.contract_mattes(mom, j0, c, t.nparams)
    return {"value_rel_err": abs(value - wv) / abs(wv),
            "derivative_rel_err": float(np.linalg.norm(deriv - wd) / np.linalg.norm(wd)),
            "valid_equal": int(valid) == int(wn), "oracle_value": wv,
            "tolerance": {"value": 1e-6, "derivative": 1e-4}}


def cpu_threads():
    """Threads for the reference's parallel CPU legs: the physical cores this
    process may use, at most 16 (the host share of one GPU)."""
    _, phys, avail = cpu_info()
    return max(1, min(phys or avail, avail, 16))


def gaussian_cpu_record(S, O, fixed, df, sigma, device_ms, reps=3):
    """The recursive Gaussian of one 256x256x160 volume by the oracle, serial
    (t1) and parallel over lines (tN, as the reference runs it,
    crates/sitk/src/core/parallel.rs:990), against the device's volume bit for
    bit."""
    v = fixed.data
    size = tuple(int(x) for x in fixed.size)
    n = cpu_threads()
    t1, tn, ref = [], [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        ref = O.smoothing_recursive_gaussian(v, size, fixed.spacing, sigma, threads=1)
        t1.append(time.perf_counter() - t0)
    for _ in range(2 * reps):
        t0 = time.perf_counter()
        O.smoothing_recursive_gaussian(v, size, fixed.spacing, sigma, threads=n)
        tn.append(time.perf_counter() - t0)
    img = S.recursive_gaussian(df, sigma)
    dev = img.to_host().reshape(-1)
    img.free()
    same = bool(np.array_equal(dev.view(np.uint32), np.asarray(ref, np.float32).reshape(-1).view(np.uint32)))
    return {"t1_ms": float(np.median(t1)) * 1e3, "tN_ms": float(np.median(tn)) * 1e3, "N": n,
            "device_ms": device_ms, "device_bit_exact": same,
            "sample": f"{size[0]}x{size[1]}x{size[2]} f32, sigma {sigma[0]} mm, median of "
                      f"{reps} (t1) and {2 * reps} (tN) runs"}


def ms_record(S, W, O, steps=50, evals_n=3):
    """Mean squares at 256^3, every voxel (16,777,216 samples), Euler3D: the
    device's value + derivative evaluation, and the oracle's serial (t1) and
    row-parallel with an in-order fold (tN, crates/sitk/src/core/parallel.rs:
    766-818) evaluation of the same, with the device's error against it. The
    reference publishes its CUDA kernel at this size
    (crates/sitk/src/cuda/ops/mean_squares.rs:239-244: 6.088 ms f32/f32 on an
    RTX 5000 Ada)."""
    fixed = W.blob_volume((256, 256, 256), (1.0, 1.0, 1.0), 11, nblobs=40)
    rng = np.random.default_rng(511)
    _, T = W.random_euler(rng, 5.0, 3.0, fixed.center())
    moving = W.resample_linear(fixed, T)
    df = S.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = S.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    met = S.ResidentMetric(df, dm)
    p = np.array([0.01, -0.02, 0.015, 0.5, -0.3, 0.2])
    stage = W.AffineMap(W.euler_matrix(*p[:3]), p[3:], fixed.center()).stage().reshape(1, 12)
    t = O.make_transform(O.EULER3D, list(p), center=list(fixed.center()))
    for _ in range(5):
        met.evaluate(stage)
    t0 = time.perf_counter()
    for _ in range(steps):
        mom = met.evaluate(stage)
    dev_ms = (time.perf_counter() - t0) * 1e3 / steps
    pb = O.MetricProblem(fixed.data, fixed.spacing, fixed.origin, moving.data, moving.spacing,
                         moving.origin)
    n = cpu_threads()
    t0 = time.perf_counter()
    wv, wd, wn = pb.ms_evaluate(t)
    t1 = time.perf_counter() - t0
    tn = []
    for _ in range(evals_n):
        t0 = time.perf_counter()
        pb.ms_evaluate(t, threads=n)
        tn.append(time.perf_counter() - t0)
    j0 = O.transform_jacobian(t, [0, 0, 0])
    c = np.stack([O.transform_jacobian(t, e) - j0 for e in np.eye(3)])
    value, deriv, valid = S.contract_ms(mom, j0, c, 6)
    met.free()
    df.free()
    dm.free()
    nvox = fixed.data.size
    return {"workload": "256^3 f32 1 mm, 40 blobs, moving through a seeded Euler3D "
                        "(+-5 deg, +-3 mm); mean squares on every voxel; Euler3D point map",
            "device_ms": dev_ms, "device_GBps": 2 * nvox * 4 / (dev_ms * 1e-3) / 1e9,
            "cpu_t1_ms": t1 * 1e3, "cpu_tN_ms": float(np.median(tn)) * 1e3, "N": n,
            "value_rel_err": abs(value - wv) / abs(wv),
            "derivative_rel_err": float(np.linalg.norm(deriv - wd) / np.linalg.norm(wd)),
            "valid_equal": int(valid) == int(wn),
            "published_cuda_kernel_ms": {"rtx5000ada_f32_f32": 6.088, "rtx5000ada_f64_f64": 5.755,
                                         "source": "crates/sitk/src/cuda/ops/mean_squares.rs:239-244"}}


def cpu_pair_record(S, W, O):
    """The CPU-runnable rigid pair (64^3, every voxel, Euler3D, Mattes 50 bins,
    RSGD with physical-shift scales, at most 100 iterations, one level)
    registered by the host restatement of the reference's loop driving the
    oracle (one core, as the reference's Mattes) and by the device."""
    from oracle.registration import register_rsgd

    fixed, moving, T_gt = W.config_cpu_pair()
    c = list(fixed.center())
    pb = O.MetricProblem(fixed.data, fixed.spacing, fixed.origin, moving.data, moving.spacing,
                         moving.origin)
    g = pb.mattes_geometry(50)
    t0 = time.perf_counter()
    want, want_it, _ = register_rsgd(pb, g, O.EULER3D, [0.0] * 6, c, fixed.size, fixed.spacing,
                                     fixed.origin, minimum_step=1e-4, iterations=100,
                                     grad_tol=1e-6)
    t_cpu = time.perf_counter() - t0
    df = S.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = S.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    cfg = S.registration_config(metric=S.MATTES, bins=50, optimizer=S.REGULAR_STEP,
                                minimum_step=1e-4, iterations=100, gradient_tolerance=1e-6,
                                learning_rate_mode=S.LR_ONCE)
    S.register(df, dm, cfg, S.EULER3D, [0.0] * 6, center=fixed.center())
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        got, res = S.register(df, dm, cfg, S.EULER3D, [0.0] * 6, center=fixed.center())
        times.append(time.perf_counter() - t0)
    df.free()
    dm.free()

    def tre(p):
        est = W.AffineMap(W.euler_matrix(*p[:3]), np.asarray(p[3:]), fixed.center())
        return float(W.tre(est, T_gt.inverse(), fixed))

    t_dev = float(np.median(times))
    return {"workload": "64^3 1 mm, 40 blobs + N(0,5), Euler3D +-5 deg / +-3 mm; Mattes 50 bins, "
                        "every voxel (262,144 samples), RSGD (physical-shift scales, rate "
                        "estimated once, minimum step 1e-4), at most 100 iterations, one level",
            "cpu": {"seconds": t_cpu, "registrations_per_s": 1.0 / t_cpu, "cores": 1,
                    "iterations": int(want_it), "tre_mm": tre(want)},
            "device": {"seconds": t_dev, "registrations_per_s": 1.0 / t_dev,
                       "iterations": int(res.iterations), "tre_mm": tre(got),
                       "parameter_max_abs_diff_vs_cpu": float(np.max(np.abs(got - want)))}}


def traffic_probe() -> int:
    """Child of the traffic measurement (run under rocprofv3 --pmc): the
    bench's evaluation a few times, then a plain copy of known size."""
    import sitk_hip as S
    import workload as W

    S.backend_init(0)
    fixed, moving, T_gt = W.config_large_pair(seed=2)
    idx = S.draw_samples(fixed.data.size, S.SAMPLING_RANDOM, 0.2, 7)
    df = S.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = S.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    metric = S.ResidentMattes(df, dm, bins=50, indices=idx)
    stage = T_gt.inverse().stage().reshape(1, 12)
    for _ in range(6):
        metric.value_and_moments(stage)
    for _ in range(3):
        S.recursive_gaussian(df, [0.0, 0.0, 0.0]).free()
    return 0


def measure_traffic(kernels, copy_bytes):
    """HBM bytes per launch of `kernels`, from two rocprofv3 --pmc runs of the
    traffic probe (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled
    by what the counter reads for copy_f32 of `copy_bytes` read and written.
    Returns ({kernel: {read, written, total}}, calibration) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    per = {}
    cal = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
            cmd = ["timeout", "-s", "KILL", "120", exe, "--pmc", counter, "-f", "csv", "-d", d,
                   "-o", "run", "--", sys.executable, os.path.abspath(__file__), "--traffic-probe"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                return None, f"rocprofv3 --pmc {counter} exited {r.returncode}"
            rows = []
            for root, _, files in os.walk(d):
                for fn in files:
                    if fn.endswith("counter_collection.csv"):
                        with open(os.path.join(root, fn)) as fh:
                            rows += list(csv.DictReader(fh))
        vals = {}
        for row in rows:
            if row.get("Counter_Name") != counter:
                continue
            name = row.get("Kernel_Name", "")
            for k in list(kernels) + ["copy_f32"]:
                if f"::{k}<" in name or name.endswith(f"::{k}") or f"::{k}(" in name:
                    vals.setdefault(k, []).append(float(row["Counter_Value"]))
        if "copy_f32" not in vals:
            return None, f"no copy_f32 dispatch in the {counter} run"
        factor = copy_bytes / (float(np.median(vals["copy_f32"])) * 1024.0)
        cal[counter] = {"copy_KiB": float(np.median(vals["copy_f32"])), "factor": factor}
        for k in kernels:
            if k in vals:
                # skip the warm-up dispatches' first (cold) launch
                v = vals[k][1:] if len(vals[k]) > 1 else vals[k]
                per.setdefault(k, {})[counter] = float(np.median(v)) * 1024.0 * factor
    out = {}
    for k, c in per.items():
        rd, wr = c.get("FETCH_SIZE"), c.get("WRITE_SIZE")
        out[k] = {"read": rd, "written": wr,
                  "total": (rd + wr) if rd is not None and wr is not None else None}
    return out, cal


def batched_pair_configs(S, pairs):
    """Per-pair registration of the batched rigid workload: Euler3D about the
    fixed centre from the identity, RegularStepGradientDescent with
    physical-shift scales and a once-estimated rate, a two-level pyramid
    (shrink 2 / 1, sigma 1 / 0 mm; three levels 4 / 2 / 1, sigma 2 / 1 / 0 for
    pairs whose shortest edge is >= 64). Mean-squares pairs use every voxel,
    Mattes pairs 32 bins at 25% random sampling."""
    cfgs = []
    for i, p in enumerate(pairs):
        if min(p["fixed"].size) >= 64:
            sh, sg = ((4, 4, 4), (2, 2, 2), (1, 1, 1)), (2.0, 1.0, 0.0)
        else:
            sh, sg = ((2, 2, 2), (1, 1, 1)), (1.0, 0.0)
        mattes = p["metric"] == "mattes"
        cfgs.append(S.registration_config(
            metric=S.MATTES if mattes else S.MEAN_SQUARES, bins=32,
            sampling=S.SAMPLING_RANDOM if mattes else S.SAMPLING_NONE,
            sampling_percentage=0.25, seed=17 + i, minimum_step=1e-3, iterations=100,
            gradient_tolerance=1e-8, learning_rate_mode=S.LR_ONCE, shrink_factors=sh, sigmas=sg))
    return cfgs


def registrations_record(S, W, n_pairs, n_plain):
    """Registrations/s of the batched rigid workload: every pair registered in
    one batched run (one batched evaluation per iteration, default tiling: the
    autotuner's 6-18 ms per level does not pay back over these levels' 30-101
    evaluations of 0.27-0.66 ms)
    against the same registrations run one at a time; both from volumes
    resident on the device, success = TRE <= 0.5 mm against the inverse of the
    generating warp."""
    t0 = time.perf_counter()
    pairs = W.config_batched_pairs(n_pairs, seed=3)
    log(f"{n_pairs} pairs generated in {time.perf_counter() - t0:.1f}s")
    fixed, moving = [], []
    for p in pairs:
        v = p["fixed"]
        f = S.DeviceImage.upload(v.data, spacing=v.spacing, origin=v.origin)
        fixed.append(f)
        moving.append(S.resample_linear(f, f.geometry(), p["T_gt"].stage().reshape(1, 12)))
    cfgs = batched_pair_configs(S, pairs)
    centres = [p["fixed"].center() for p in pairs]
    kinds = [S.EULER3D] * n_pairs
    p0 = [[0.0] * 6] * n_pairs

    def tre_of(params, p):
        v = p["fixed"]
        est = W.AffineMap(W.euler_matrix(*params[:3]), params[3:], v.center())
        return W.tre(est, p["T_gt"].inverse(), v)

    t0 = time.perf_counter()
    got, res = S.register_batch(fixed, moving, cfgs, kinds, p0, centres, autotune=False)
    t_batch = time.perf_counter() - t0
    tre_b = np.array([tre_of(g, p) for g, p in zip(got, pairs)])
    log(f"batched: {n_pairs} registrations in {t_batch:.2f}s, "
        f"{np.mean(tre_b <= 0.5) * 100:.1f}% at <= 0.5 mm")
    k = min(n_plain, n_pairs)
    t0 = time.perf_counter()
    plain = [S.register(fixed[i], moving[i], cfgs[i], kinds[i], p0[i], center=centres[i])
             for i in range(k)]
    t_plain = time.perf_counter() - t0
    tre_p = np.array([tre_of(g, p) for (g, _), p in zip(plain, pairs[:k])])
    log(f"plain: {k} registrations in {t_plain:.2f}s, {np.mean(tre_p <= 0.5) * 100:.1f}% at <= 0.5 mm")
    ok_b = float(np.mean(tre_b <= 0.5))
    ok_p = float(np.mean(tre_p <= 0.5))
    its = np.array([r.iterations for r in res])
    roof = batch_roofline(S, pairs, fixed, moving)
    return {
        "workload": f"{n_pairs} rigid pairs, edges U{{32..128}} per axis, 1 mm, 40 blobs, "
                    "Euler3D +-15 deg / +-15 mm; half mean squares (every voxel), "
                    "half Mattes (32 bins, 25% sampling)",
        "batched": {"pairs": n_pairs, "seconds": t_batch, "registrations_per_s": n_pairs / t_batch,
                    "success_fraction": ok_b, "successful_per_s": ok_b * n_pairs / t_batch,
                    "iterations_mean": float(its.mean())},
        "plain": {"pairs": k, "seconds": t_plain, "registrations_per_s": k / t_plain,
                  "success_fraction": ok_p, "successful_per_s": ok_p * k / t_plain},
        "speedup": (n_pairs / t_batch) / (k / t_plain),
        "batch_roofline": roof,
    }


def batch_roofline(S, pairs, fixed, moving, reps=20):
    """The batched engine's evaluation of every pair's finest level (mean
    squares on every voxel, Mattes 32 bins on 25% random draws) at the inverse
    of each generating warp, timed per kernel with the library's profiler.
    Algorithmic bytes of the first pass (all pairs' values, histograms and
    moments): mean squares N*4 (fixed value) + M*4 (moving volume), Mattes
    n*(4 + 1) (sample index, fixed bin) + M*4; the second pass (the Mattes
    derivative) reads the Mattes part again."""
    metrics, stages = [], []
    pass1 = pass2 = 0
    for i, p in enumerate(pairs):
        nvox = int(np.prod(p["fixed"].size))
        if p["metric"] == "mattes":
            idx = S.draw_samples(nvox, S.SAMPLING_RANDOM, 0.25, 17 + i)
            m = S.ResidentMattes(fixed[i], moving[i], bins=32, indices=idx)
            b = len(idx) * 5 + nvox * 4
            pass1 += b
            pass2 += b
        else:
            m = S.ResidentMetric(fixed[i], moving[i])
            pass1 += nvox * 8
        metrics.append(m)
        stages.append(p["T_gt"].inverse().stage())
    batch = S.Batch(metrics)
    stages = np.stack(stages)
    for _ in range(3):
        batch.evaluate(stages)
    # per-kernel times from direct launches (a graph replay is timed whole)
    tile, grid = batch.tiling()
    batch.set_tiling(tile, grid, use_graph=False)
    S.profile_reset()
    S.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.evaluate(stages)
    wall_ms = (time.perf_counter() - t0) * 1e3 / reps
    S.profile_enable(False)
    kern = {}
    for name in ("batch_pass1", "batch_tail", "batch_pass2", "batch_fold"):
        cnt, tot = S.profile_query(name)
        kern[name] = tot / cnt * 1e3 if cnt else None
    batch.set_tiling(tile, grid, use_graph=True)
    for _ in range(3):
        batch.evaluate(stages)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.evaluate(stages)
    graph_ms = (time.perf_counter() - t0) * 1e3 / reps
    batch.free()
    for m in metrics:
        m.free()
    p1_us = kern["batch_pass1"]
    return {
        "workload": f"{len(pairs)} pairs' finest level, one batched value + derivative evaluation",
        "bytes_pass1": pass1,
        "bytes_pass2": pass2,
        "kernels_us": kern,
        "tiling": {"tile_samples": tile, "grid": grid},
        "ms_per_evaluation_graph": graph_ms,
        "ms_per_evaluation_direct_profiled": wall_ms,
        "pass1_GBps": pass1 / (p1_us * 1e-6) / 1e9 if p1_us else None,
        "pass1_frac": pass1 / (p1_us * 1e-6) / 1e9 / 8000.0 if p1_us else None,
        "evaluation_GBps": (pass1 + pass2) / (graph_ms * 1e-3) / 1e9,
    }


def learned_init_record(S, n_pairs, plain_pairs):
    """Learned initialisation + batched refinement against the plain method on
    held-out affine pairs (128^3, +-30 deg, +-10% scale, +-20 mm), with the
    network trained by this library's kernels (sitk-rs_amd/train_initializer.py)."""
    import initializer as I

    path = os.path.join(ROOT, "sitk-rs_amd", "weights", "initializer_cnn.npy")
    if not os.path.exists(path):
        return None
    net = S.CNN(max_batch=32, seed=0)
    I.load(net, path)
    rec = I.evaluate(net, n_pairs, plain_pairs=plain_pairs)
    net.free()
    rec["workload"] = (f"{n_pairs} held-out pairs, 128^3 1 mm, 40 blobs, affine +-30 deg / "
                       "+-10% / +-20 mm; CNN prediction then 30 RSGD iterations of Mattes "
                       "(50 bins, 25%) batched, centered fallback; plain = centered start, "
                       "3-level RSGD one pair at a time")
    return rec


def large_pair_record(S, W, fixed, moving, T_gt, reps=3):
    """The single large pair registered end to end by the plain method on the
    device: affine from the identity about the centre, three-level pyramid
    (shrink 4 / 2 / 1, sigma 2 / 1 / 0), Mattes 50 bins at 20% sampling,
    gradient descent with the golden-section line search (100 iterations per
    level, physical-shift scales, rate estimated once). Median wall clock of
    `reps` runs from resident volumes; TRE against the inverse generating
    affine."""
    df = S.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = S.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    ident = list(np.eye(3).reshape(9)) + [0.0, 0.0, 0.0]
    cfg = S.registration_config(metric=S.MATTES, bins=50, sampling=S.SAMPLING_RANDOM,
                                sampling_percentage=0.2, seed=1, optimizer=S.LINE_SEARCH,
                                iterations=100, learning_rate_mode=S.LR_ONCE,
                                shrink_factors=((4, 4, 4), (2, 2, 2), (1, 1, 1)),
                                sigmas=(2.0, 1.0, 0.0))
    times, res, p = [], None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        p, res = S.register(df, dm, cfg, S.AFFINE, ident, center=fixed.center())
        times.append(time.perf_counter() - t0)
    est = W.AffineMap(np.asarray(p[:9]).reshape(3, 3), np.asarray(p[9:]), fixed.center())
    tre = W.tre(est, T_gt.inverse(), fixed)
    df.free()
    dm.free()
    return {"workload": "256x256x160 pair (0.9x0.9x1.5 mm, +-20% bias), affine +-10 deg / +-5% / "
                        "+-10 mm; 3-level pyramid 4/2/1, sigma 2/1/0; Mattes 50 bins, 20%; "
                        "gradient descent with line search, affine from the identity",
            "seconds": float(np.median(times)), "registrations_per_s": 1.0 / float(np.median(times)),
            "tre_mm": float(tre), "iterations": int(res.iterations),
            "value_and_gradient_evaluations": int(res.evaluations),
            "value_evaluations": int(res.value_evaluations)}


def deformable_record(S, W, n_starts, n_plain, iterations=100):
    """The deformable multi-start workload: a 192^3 pair related by a seeded
    Euler3D offset and a 20-bump Gaussian-RBF displacement (peak 4 mm); a
    learned-initializer-seeded affine pre-alignment (CNN prediction refined by
    30 RSGD iterations, centered fallback); `n_starts` starts = that
    pre-alignment perturbed (+-2 deg, +-2 mm; start 0 unperturbed) as the
    moving-initial transform of a cubic B-spline (10^3 control points, 3000
    parameters), Mattes 32 bins at 10% sampling, L-BFGS-B. Batched: every
    start optimised together, one set of launches per round; plain: the same
    starts one at a time. Success = mean TRE <= 0.5 mm over the interior."""
    import initializer as I

    t0 = time.perf_counter()
    f, m, truth = W.config_deformable()
    log(f"deformable pair generated in {time.perf_counter() - t0:.1f}s")
    df = S.DeviceImage.upload(f.data, spacing=f.spacing, origin=f.origin)
    dm = S.DeviceImage.upload(m.data, spacing=m.spacing, origin=m.origin)
    c = f.center()
    mesh, npar = (7, 7, 7), 3 * 10 ** 3
    zero = np.zeros(npar)
    t0 = time.perf_counter()
    path = os.path.join(ROOT, "sitk-rs_amd", "weights", "initializer_cnn.npy")
    seeds = [W.AffineMap(np.eye(3), np.zeros(3), c)]
    if os.path.exists(path):
        net = S.CNN(max_batch=32, seed=0)
        I.load(net, path)
        g128 = S.Geometry.make((128,) * 3, (1.0,) * 3, tuple(c - 63.5))
        cf, cm = S.resample_linear(df, g128), S.resample_linear(dm, g128)
        seeds.insert(0, I.predict(net, [(cf, cm, c)])[0])
        net.free()
    cfg = I.refine_config_fast(30)
    got, res = S.register_batch([df] * len(seeds), [dm] * len(seeds), [cfg] * len(seeds),
                                [S.AFFINE] * len(seeds), [I.affine_params(T) for T in seeds],
                                [c] * len(seeds))
    k = int(np.argmin([r.value for r in res]))
    pre = W.AffineMap(np.asarray(got[k][:9]).reshape(3, 3), np.asarray(got[k][9:]), c)
    t_pre = time.perf_counter() - t0
    tre_pre = W.deformable_tre(truth, f, mesh, zero, pre.stage())
    log(f"pre-alignment ({'network' if k == 0 and len(seeds) == 2 else 'centered'} seed) "
        f"in {t_pre:.2f}s, TRE {tre_pre:.3f} mm")
    rng = np.random.default_rng(17)
    stages = [pre.stage()]
    for _ in range(n_starts - 1):
        _, P = W.random_euler(rng, 2.0, 2.0, c)
        stages.append(W.AffineMap(pre.A @ P.A, pre.A @ P.t + pre.t, c).stage())
    stages = np.stack(stages)
    kw = dict(bins=32, sampling=S.SAMPLING_RANDOM, sampling_percentage=0.10, seed=1,
              iterations=iterations)
    t0 = time.perf_counter()
    xs, rs, best = S.register_bspline_multistart(df, dm, mesh, n_starts, stages=stages, **kw)
    t_batch = time.perf_counter() - t0
    tre_b = np.array([W.deformable_tre(truth, f, mesh, xs[s], stages[s]) for s in range(n_starts)])
    log(f"batched: {n_starts} starts in {t_batch:.2f}s, best start TRE {tre_b[best]:.3f} mm, "
        f"{np.mean(tre_b <= 0.5) * 100:.1f}% of starts at <= 0.5 mm")
    kp = min(n_plain, n_starts)
    t0 = time.perf_counter()
    plain = [S.register_bspline_multistart(df, dm, mesh, 1, stages=stages[s:s + 1], **kw)
             for s in range(kp)]
    t_plain = time.perf_counter() - t0
    tre_p = np.array([W.deformable_tre(truth, f, mesh, x[0], stages[s])
                      for s, (x, _, _) in enumerate(plain)])
    log(f"plain: {kp} starts one at a time in {t_plain:.2f}s")
    ok_b, ok_p = float(np.mean(tre_b <= 0.5)), float(np.mean(tre_p <= 0.5))
    return {
        "workload": "192^3 1 mm pair, 135 blobs, Euler3D +-5 deg / +-5 mm offset plus 20 "
                    "Gaussian-RBF bumps (peak 4 mm); B-spline 10^3 control points (3000 "
                    f"parameters) after a learned-init affine pre-alignment; {n_starts} starts "
                    f"(pre-alignment perturbed +-2 deg / +-2 mm); Mattes 32 bins, 10% sampling; "
                    f"L-BFGS-B, {iterations} iterations",
        "pre_alignment": {"seconds": t_pre, "tre_mm": tre_pre,
                          "seed": "network" if (k == 0 and len(seeds) == 2) else "centered"},
        "batched": {"starts": n_starts, "seconds": t_batch, "registrations_per_s": n_starts / t_batch,
                    "success_fraction": ok_b, "successful_per_s": ok_b * n_starts / t_batch,
                    "best_start": int(best), "best_start_tre_mm": float(tre_b[best]),
                    "median_tre_mm": float(np.median(tre_b)),
                    "evaluations_mean": float(np.mean([r.evaluations for r in rs]))},
        "plain": {"starts": kp, "seconds": t_plain, "registrations_per_s": kp / t_plain,
                  "success_fraction": ok_p, "successful_per_s": ok_p * kp / t_plain},
        "speedup": (n_starts / t_batch) / (kp / t_plain),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-evals", type=int, default=16)
    ap.add_argument("--pairs", type=int, default=512,
                    help="batched rigid pairs for the registrations/s record (0 skips it)")
    ap.add_argument("--plain-pairs", type=int, default=512,
                    help="pairs registered one at a time for the plain rate")
    ap.add_argument("--lp-reps", type=int, default=3,
                    help="end-to-end registrations of the large pair (0 skips the record)")
    ap.add_argument("--starts", type=int, default=64,
                    help="starts of the deformable multi-start record (0 skips it)")
    ap.add_argument("--plain-starts", type=int, default=64,
                    help="deformable starts registered one at a time for the plain rate")
    ap.add_argument("--init-pairs", type=int, default=1024,
                    help="held-out pairs for the learned-initializer record (0 skips it)")
    ap.add_argument("--init-plain", type=int, default=1024,
                    help="of those, pairs registered one at a time by the plain method")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the HBM-traffic measurement (two rocprofv3 --pmc child runs)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        return traffic_probe()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        return 2

    import sitk_hip as S
    import workload as W

    S.backend_init(0)
    log(f"device: {S.device_name()}")
    t0 = time.perf_counter()
    fixed, moving, T_gt = W.config_large_pair(seed=2)
    log(f"workload generated in {time.perf_counter() - t0:.1f}s")
    nvox = fixed.data.size
    idx = S.draw_samples(nvox, S.SAMPLING_RANDOM, 0.2, 7)
    n = idx.size
    prec = S.PRECISION_F32 if args.precision == "f32" else S.PRECISION_F64
    df = S.DeviceImage.upload(fixed.data, spacing=fixed.spacing, origin=fixed.origin)
    dm = S.DeviceImage.upload(moving.data, spacing=moving.spacing, origin=moving.origin)
    metric = S.ResidentMattes(df, dm, bins=50, indices=idx, precision=prec)
    # the evaluation point: the registration's solution (the inverse of the
    # warp that made the moving image), where a converging optimizer spends its
    # iterations
    I = T_gt.inverse()
    stage = I.stage().reshape(1, 12)

    # secondary: the level-2 pyramid smoothing of one volume (sigma 2 mm)
    sig = [2.0, 2.0, 2.0]
    for _ in range(3):
        S.recursive_gaussian(df, sig).free()

    for _ in range(args.warmup):
        metric.value_and_moments(stage)
    # The timed steps run without per-kernel events (they add tens of us per
    # step); the kernel durations come from a separate profiled loop below.
    # The K steps are made inside the library, one after another, each
    # waiting for its result as the reference's compiled optimizer loop does
    # (a Python call per step would add the interpreter's own turnaround);
    # the same K steps from Python are reported beside them.
    metric.value_and_moments_repeat(stage, args.warmup)
    t0 = time.perf_counter()
    value, mom, valid = metric.value_and_moments_repeat(stage, args.steps)
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        metric.value_and_moments(stage)
    ms_step_py = (time.perf_counter() - t0) * 1e3 / args.steps
    S.profile_reset()
    S.profile_enable(True)
    for _ in range(min(args.steps, 100)):
        metric.value_and_moments(stage)
    S.profile_enable(False)

    # The evaluation's bytes: the reference's two passes each read the moving
    # volume once and the sample stream (index u32 + fixed bin u8) once. The
    # headline GB/s is that fixed amount per evaluation over the step time.
    per_pass = n * 5 + nvox * 4
    bytes_step = 2 * per_pass
    gbs = bytes_step / (ms_step * 1e-3) / 1e9
    # The roofline of the dominant kernel, the value pass, on the algorithm's
    # bytes for one pass (moving volume once + sample stream); the records it
    # also writes for its own derivative pass are reported beside it.
    rec_bytes = 16 if prec == S.PRECISION_F32 else 32

    kern = {}
    for name in ("mattes_hist", "mattes_tail", "mattes_deriv", "fold_partials"):
        cnt, tot = S.profile_query(name)
        kern[name] = {"count": cnt, "avg_us": (tot / cnt * 1e3) if cnt else None}
    dom = "mattes_hist"
    dom_us = kern[dom]["avg_us"]
    achieved = per_pass / (dom_us * 1e-6) / 1e9
    with_records = per_pass + n * rec_bytes

    # Gaussian, timed the same way (4 events around its 3 axis launches)
    S.profile_reset()
    S.profile_enable(True)
    gsteps = 10
    tg = time.perf_counter()
    for _ in range(gsteps):
        S.recursive_gaussian(df, sig).free()
    g_ms = (time.perf_counter() - tg) * 1e3 / gsteps
    S.profile_enable(False)
    gk = {}
    for ax in range(3):
        cnt, tot = S.profile_query(f"gauss_iir_axis{ax}")
        gk[f"axis{ax}_us"] = tot / cnt * 1e3 if cnt else None
    g_bytes = nvox * 40  # f32 in, two f64 intermediates, f32 out

    traffic, traffic_cal = None, None
    if not args.no_traffic:
        t0 = time.perf_counter()
        traffic, traffic_cal = measure_traffic(["mattes_hist", "mattes_deriv_rec"], nvox * 4)
        log(f"traffic measured in {time.perf_counter() - t0:.1f}s: {traffic or traffic_cal}")

    cpu, parity = None, None
    if not args.no_cpu:
        import oracle as O

        model, phys, avail = cpu_info()
        t_cpu, k, want, t_eval = cpu_baseline(fixed, moving, idx, (I.A, I.offset),
                                              evals=args.cpu_evals)
        parity = parity_record(S, O, t_eval, value, mom, valid, want)
        log(f"parity vs oracle: {parity}")
        cpu = {
            "value": bytes_step / t_cpu / 1e9,
            "unit": "GB/s",
            "cores": 1,
            "kind": "port",
            "sample": f"median of {k} full value+derivative evaluations of the same level "
                      f"({n} samples, 12-parameter affine) by the serial FP64 oracle, "
                      f"{t_cpu * 1e3:.0f} ms each (the reference's host Mattes is serial)",
            "ms_per_eval": t_cpu * 1e3,
            "cpu_model": model,
            "physical_cores": phys,
            "cpus_available": avail,
        }
        t0 = time.perf_counter()
        cpu["gaussian"] = gaussian_cpu_record(S, O, fixed, df, sig, g_ms)
        cpu["mean_squares_256"] = ms_record(S, W, O)
        cpu["cpu_pair_registration"] = cpu_pair_record(S, W, O)
        log(f"CPU legs in {time.perf_counter() - t0:.1f}s")

    out = {
        "metric": METRIC,
        "value": gbs,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "ms_per_step_python_loop": ms_step_py,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": "Mattes MI value+derivative, finest level of the single large pair: "
                        "256x256x160 f32, 50 bins, 20% random sampling, affine",
            "samples": int(n),
            "moving_voxels": int(nvox),
            "bytes_per_step": int(bytes_step),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (traffic or {}).get(dom, {}).get("total") if traffic else None,
            "bytes_per_launch": int(per_pass),
            "avg_us": dom_us,
            "with_records": {"bytes_per_launch": int(with_records),
                             "achieved": with_records / (dom_us * 1e-6) / 1e9,
                             "frac": with_records / (dom_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
            "step_frac": gbs / HBM_PEAK_GBS,
        },
        "traffic": {"per_kernel": traffic, "calibration": traffic_cal},
        "kernels": kern,
        "gaussian_level": {
            "ms_per_volume": g_ms,
            "GBps": g_bytes / (g_ms * 1e-3) / 1e9,
            **gk,
        },
        "result": {"value": value, "valid": int(valid)},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    if args.lp_reps > 0:
        out["large_pair_registration"] = large_pair_record(S, W, fixed, moving, T_gt, args.lp_reps)
    if args.pairs > 0:
        out["registrations"] = registrations_record(S, W, args.pairs, args.plain_pairs)
    if args.init_pairs > 0:
        rec = learned_init_record(S, args.init_pairs, min(args.init_plain, args.init_pairs))
        if rec is not None:
            out["learned_init"] = rec
    if args.starts > 0:
        out["deformable_multistart"] = deformable_record(S, W, args.starts, args.plain_starts)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
