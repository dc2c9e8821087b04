#!/usr/bin/env python3
"""Trains the learned Davidson preconditioner on samples harvested from the library's
own sweeps and writes its parameters.

Training problems: the seeded synthetic ab-initio generator (chemps2_amd.problems.
random_abinitio) at L = 24, N = 24, 2S = 0, random D2h irreps, seeds --seeds (default
0..7), virtual dimension --D (default 1000). For each: a random left-normalised MPS at
the SyBookkeeper dimensions, the operators of every boundary, then one left and one
right sweep with the diagonal preconditioner (Davidson tolerance 1e-5), harvesting
--per-iter elements of every preconditioner application. The network then trains on
the device (Adam) and its parameters go to --out (float32 .npy) with a JSON summary
beside it.

Usage: python -m chemps2_amd.precond_train [--seeds 0-7] [--D 1000] [--out PATH]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "precond_mlp.npy")


def parse_seeds(text):
    if "-" in text:
        a, b = text.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(v) for v in text.split(",")]


def harvest(L, D, seed, per_iter, rtol=1e-5):
    import chemps2_amd
    from chemps2_amd.problems import mx_table, random_abinitio

    irr, T, V = random_abinitio(L, seed, group=7)
    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
    c.harvest(per_iter)
    c.dmrg_init(D, seed + 1)
    c.dmrg_sweep(False, D, rtol, 0.0, False)
    e = c.dmrg_sweep(True, D, rtol, 0.0, True)
    data = c.harvest_take()
    c.close()
    return data, e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--D", type=int, default=1000)
    ap.add_argument("--seeds", default="0-7")
    ap.add_argument("--val-seed", type=int, default=1000)
    ap.add_argument("--per-iter", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()

    import chemps2_amd

    seeds = parse_seeds(args.seeds)
    parts = []
    t0 = time.time()
    for s in seeds:
        t = time.time()
        data, e = harvest(args.L, args.D, s, args.per_iter)
        parts.append(data)
        print(f"seed {s}: {len(data[1])} samples, sweep energy {e:.10f} ({time.time() - t:.1f} s)", file=sys.stderr,
              flush=True)
    X, p, e, w = (np.concatenate([d[k] for d in parts]) for k in range(4))
    val, _ = harvest(args.L, args.D, args.val_seed, args.per_iter)
    t = time.time()
    c = chemps2_amd.Context(0)
    c.mlp_set(chemps2_amd.mlp_init(0))
    base = c.mlp_loss_grad(X, p, e, w)[0]
    vbase = c.mlp_loss_grad(*val)[0]
    l0, l1 = c.mlp_train(X, p, e, w, batch=args.batch, epochs=args.epochs, lr=args.lr, seed=1)
    final = c.mlp_loss_grad(X, p, e, w)[0]
    vfinal = c.mlp_loss_grad(*val)[0]
    params = c.mlp_get()
    c.close()
    print(f"training: {len(p)} samples, loss of the plain preconditioner {base:.4e}, after training {final:.4e}; "
          f"held-out seed {args.val_seed}: {vbase:.4e} -> {vfinal:.4e} ({time.time() - t:.1f} s)", file=sys.stderr,
          flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)
    summary = {
        "L": args.L, "D": args.D, "seeds": seeds, "per_iter": args.per_iter, "samples": int(len(p)),
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "loss_plain": base, "loss_first_epoch": l0, "loss_last_epoch": l1, "loss_trained": final,
        "val_seed": args.val_seed, "val_samples": int(len(val[1])), "val_loss_plain": vbase, "val_loss_trained": vfinal,
        "seconds": round(time.time() - t0, 1),
    }
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
