"""Test-side helpers: build point-map stages and the Jacobian's affine
decomposition from the oracle's transforms, the way the reference's
registration/cuda.rs (affine_form :294-363, point_stages :248-283) does."""

import numpy as np

import oracle as O


def stages_of(t) -> np.ndarray:
    return O.transform_stage(t).reshape(1, 12)


def affine_form(t):
    """(J0, C) with J(x) = J0 + sum_e x_e C_e; C has shape (3 e, 3 d, nparams)."""
    j0 = O.transform_jacobian(t, [0.0, 0.0, 0.0])
    c = np.empty((3, 3, t.nparams))
    for e in range(3):
        basis = [0.0, 0.0, 0.0]
        basis[e] = 1.0
        c[e] = O.transform_jacobian(t, basis) - j0
    return j0, c


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    return np.max(np.abs(a - b) / den)


def rel_vec(a, b):
    """Relative error of a vector against its norm."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def blob_volume(shape, seed, nblobs=12, spacing=(1.0, 1.0, 1.0)):
    """Small seeded blob volume (sum of Gaussian blobs + a gentle ramp) for
    parity cases; float32, (nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) * spacing[2], np.arange(ny) * spacing[1],
                          np.arange(nx) * spacing[0], indexing="ij")
    v = np.zeros(shape)
    ext = np.array([nx * spacing[0], ny * spacing[1], nz * spacing[2]])
    for _ in range(nblobs):
        c = rng.uniform(0.1, 0.9, 3) * ext
        s = rng.uniform(2.0, 6.0)
        a = rng.uniform(50, 200)
        v += a * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    v += 0.3 * x + 0.2 * y
    v += rng.normal(0, 2.0, shape)
    return v.astype(np.float32)


def record_error(config: str, **fields) -> None:
    """Appends one JSON line {config, **fields} to the file named by the
    SITKHIP_ERROR_LOG environment variable (no-op when it is unset): the
    measured errors against the oracle, per configuration."""
    import json
    import os

    path = os.environ.get("SITKHIP_ERROR_LOG")
    if not path:
        return
    with open(path, "a") as fh:
        fh.write(json.dumps({"config": config, **{k: (float(v) if isinstance(v, (np.floating, float))
                                                       else v) for k, v in fields.items()}}) + "\n")
