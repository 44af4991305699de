"""Train the warm-start network on the GPU and write its weights.

    python -m slam_amd.train_warmstart [--pairs 16384] [--epochs 30] [--out PATH]

Renders seeded frame pairs at 160x120 with the project's renderer (motions up
to 8 cm / 6 deg around varied poses in varied rooms), trains with the
library's own bf16 MFMA forward/backward kernels and Adam, reports the
validation error of the predicted motion against the zero-motion guess, and
saves the fp32 parameters (``slam_amd/weights/warmstart.npz`` by default).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import warmstart as ws


def log(msg: str) -> None:
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=16384)
    ap.add_argument("--val", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--wd", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--max-trans", type=float, default=0.08)
    ap.add_argument("--max-rot-deg", type=float, default=6.0)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=ws.WEIGHTS)
    args = ap.parse_args()

    t0 = time.time()
    tr = ws.render_pairs_parallel(args.pairs, args.seed, args.max_trans, args.max_rot_deg, args.workers, log=log)
    va = ws.render_pairs_parallel(args.val, args.seed + 1, args.max_trans, args.max_rot_deg, args.workers)
    log(f"rendered {args.pairs} + {args.val} pairs in {time.time() - t0:.0f}s")
    net = ws.WarmStartNet(max_batch=args.batch)
    net.init_params(args.seed)
    t1 = time.time()
    hist = ws.train(net, ws.to_device(tr), ws.to_device(va), epochs=args.epochs, batch=args.batch, lr=args.lr,
                    weight_decay=args.wd, seed=args.seed, log=log)
    log(f"trained in {time.time() - t1:.1f}s")
    net.save(args.out, {"pairs": args.pairs, "epochs": args.epochs, "seed": args.seed,
                        "max_trans": args.max_trans, "max_rot_deg": args.max_rot_deg})
    print(json.dumps({"final": hist[-1], "train_s": round(time.time() - t1, 1), "out": args.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
