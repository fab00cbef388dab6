"""Trains the learned direction/queue policy on the GPU and writes
xbfs_amd/data/policy_mlp.f32 (+ .json with the training record).

    python -m xbfs_amd.train_policy [--graphs 200] [--sources 4] [--steps 4000]
    python -m xbfs_amd.train_policy --data SAMPLES.npz --init WEIGHTS.f32 --steps 500 --lr 1e-3

--data trains on samples recorded earlier (--out-data of a collecting run)
instead of collecting them (--graphs, --sources and --reps then only describe
those samples in the training record); --init starts from existing weights
(fine-tuning a model to a changed engine) instead of a fresh initialisation.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import policy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out-data", default="")
    ap.add_argument("--data", default="", help="train on these recorded samples (.npz)")
    ap.add_argument("--init", default="", help="start from these weights (.f32)")
    args = ap.parse_args()
    t0 = time.time()
    if args.data:
        d = np.load(args.data)
        Y, M = d["Y"], d["M"]
        R = d["R"] if "R" in d.files else None
        # recorded raw level states are featurised by the library's current
        # definition; samples without them keep their stored features
        X = policy.features_of(R) if R is not None else d["X"]
    else:
        specs = policy.train_specs(args.graphs)
        X, Y, M, R = policy.collect(specs, sources=args.sources, reps=args.reps,
                                    log=lambda s: print(s, flush=True))
    t_collect = time.time() - t0
    if args.out_data:
        extra = {"R": R} if R is not None else {}
        np.savez_compressed(args.out_data, X=X, Y=Y, M=M, **extra)
    init = policy.load(args.init) if args.init else None
    params, curve = policy.train(X, Y, M, steps=args.steps, lr=args.lr, init=init)
    pred_mode = None
    meta = {"graphs": args.graphs, "sources_per_graph": args.sources, "reps": args.reps,
            "samples": int(X.shape[0]), "steps": args.steps, "lr": args.lr,
            "loss_first": float(curve[0]), "loss_last": float(curve[-1]),
            "collect_s": round(t_collect, 1), "spec_rule": "train_specs(): rng(i) for i < graphs",
            "data": os.path.basename(args.data) if args.data else "collected",
            "init": os.path.basename(args.init) if args.init else "fresh"}
    # agreement of the model's argmin with the measured fastest mode
    from ._capi import mlp_forward
    pred = mlp_forward(params, X)
    meta["argmin_agreement"] = float((pred.argmin(1) == Y.argmin(1)).mean())
    policy.save(params, meta)
    print("[policy] " + " ".join(f"{k}={v}" for k, v in meta.items()), flush=True)
    del pred_mode
    return 0


if __name__ == "__main__":
    sys.exit(main())
