"""Train the learned initializer on the GPU with the library's own kernels and
write its parameters (float32 .npy).

    python sitk-rs_amd/train_initializer.py OUT.npy [--steps N] [--batch B]
                                            [--lr LR] [--eval N]
"""

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import numpy as np  # noqa: E402

import initializer as I  # noqa: E402
import sitk_hip as S  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--final-lr", type=float, default=5e-5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--eval", type=int, default=0, help="held-out pairs to evaluate after training")
    args = ap.parse_args()
    S.backend_init(0)
    net = S.CNN(max_batch=max(args.batch, 32), seed=args.seed)

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    losses = I.train_schedule(net, args.steps, args.batch, lr=args.lr, final_lr=args.final_lr,
                              seed=args.seed + 1, log=log)
    log(f"trained {args.steps} steps x {args.batch} pairs in {time.perf_counter() - t0:.1f}s, "
        f"final loss {np.mean(losses[-100:]):.5f}")
    I.save(net, args.out)
    if args.eval:
        log(json.dumps(I.evaluate(net, args.eval)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
