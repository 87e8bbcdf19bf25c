"""Train the learned coarse-grid network on the GPU and save its weights.

    python -m topoptlab_amd.train_coarse_net --nelx 96 --runs 40 --iters 10 \
        --epochs 30 --out topoptlab_amd/data/coarse_net.npz

Pairs come from plain GMG-PCG solves along OC runs of nelx x nelx/2 x nelx/2
cantilevers with seeded point-load positions (``learned.generate_pairs``) and,
with ``--field-pairs N``, on seeded random density fields under the line load
of the 3D cantilever benchmark or point loads (``learned.generate_field_pairs``);
training is Adam on the MSE of the max-normalised level-1 displacement
(``learned.train``).  Progress lines go to stderr; the saved file holds the
fp32 parameters, the loss history and the data sizes.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nelx", type=int, default=96)
    ap.add_argument("--runs", type=int, default=40)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--field-pairs", type=int, default=0,
                    help="additional pairs on seeded random density fields (learned.generate_field_pairs)")
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "data" / "coarse_net.npz"))
    args = ap.parse_args(argv)
    import torch
    from . import learned
    dev = torch.device("cuda", 0)
    dims = (args.nelx, args.nelx // 2, args.nelx // 2)
    t0 = time.perf_counter()
    X, T = learned.generate_pairs(*dims, runs=args.runs, iters=args.iters, seed=args.seed,
                                  device=dev, log=log)
    nOC = len(X)
    if args.field_pairs:
        Xf, Tf = learned.generate_field_pairs(*dims, count=args.field_pairs, seed=args.seed + 1,
                                              device=dev, log=log)
        X, T = X + Xf, T + Tf
    t1 = time.perf_counter()
    log(f"{nOC} pairs from {args.runs} OC runs and {len(X) - nOC} random-field pairs in "
        f"{t1 - t0:.1f} s")
    net = learned.CoarseNet(dev, seed=args.seed)
    hist = learned.train(net, X, T, epochs=args.epochs, lr=args.lr, seed=args.seed, log=log)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"training: {args.epochs} epochs in {t2 - t1:.1f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(out, w=net.state_dict()["w"], loss=np.array(hist),
                        dims=np.array(dims), pairs=len(X), runs=args.runs, iters=args.iters,
                        field_pairs=len(X) - nOC,
                        data_seconds=t1 - t0, train_seconds=t2 - t1)
    log(f"saved {out}")


if __name__ == "__main__":
    main()
