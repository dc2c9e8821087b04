#!/usr/bin/env python3
"""Times the initialisation and the first left+right sweep of DMRG::Solve on one GPU for
the problems the design notes quote, and prints one JSON object.

  python -m chemps2_amd.sweep_run --problem abinitio --L 24 --D 1000 [--seed 0]
  python -m chemps2_amd.sweep_run --problem cylinder --D 3000        (4 x 8 Hubbard, N = 28, U = 8)
  python -m chemps2_amd.sweep_run --problem hubbard --L 12 --D 200   (1D chain, N = L, U = 4)

Per-site progress goes to stderr with CHM_VERBOSE=1 (set by --verbose).
"""
import argparse
import json
import os
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problem", choices=["abinitio", "cylinder", "hubbard"], default="abinitio")
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--D", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rtol", type=float, default=1e-5)
    ap.add_argument("--form", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.verbose:
        os.environ["CHM_VERBOSE"] = "1"

    import chemps2_amd
    from chemps2_amd.problems import hubbard, hubbard_cylinder, mx_table, random_abinitio

    if args.problem == "abinitio":
        L, N, group = args.L, args.L, 7
        irr, T, V = random_abinitio(L, args.seed, group=7)
        desc = f"random ab-initio D2h L={L} N={N} 2S=0 seed {args.seed}"
    elif args.problem == "cylinder":
        T, V, _ = hubbard_cylinder()
        L, N, group, irr = 32, 28, 0, [0] * 32
        desc = "Hubbard 4x8 cylinder (periodic around 4, open along 8, snaked) N=28 2S=0 U=8 C1"
    else:
        L, N, group, irr = args.L, args.L, 0, [0] * args.L
        T, V = hubbard(L, t=1.0, U=4.0)
        desc = f"1D Hubbard L={L} N={N} 2S=0 U=4 C1"
    c = chemps2_amd.Context(0)
    c.set_heff_form(chemps2_amd.FORM_MIXED if args.form == "mixed" else chemps2_amd.FORM_FP64)
    c.set_problem(L, group, irr, N, 0, 0, mx_table(T, V, N))
    t0 = time.time()
    c.dmrg_init(args.D, args.seed + 1)
    t_init = time.time() - t0
    print(f"initialisation {t_init:.1f} s", file=sys.stderr, flush=True)
    t0 = time.time()
    e_left = c.dmrg_sweep(False, args.D, args.rtol, 0.0, False)
    t_left = time.time() - t0
    print(f"left half {t_left:.1f} s, energy {e_left:.10f}", file=sys.stderr, flush=True)
    t0 = time.time()
    e_right = c.dmrg_sweep(True, args.D, args.rtol, 0.0, True)
    t_right = time.time() - t0
    phases, nmv = c.dmrg_stats()
    flops, ms = c.dmrg_heff_stats()
    c.close()
    print(json.dumps({
        "problem": desc, "D": args.D, "form": args.form, "davidson_rtol": args.rtol,
        "init_seconds": round(t_init, 3), "left_seconds": round(t_left, 3), "right_seconds": round(t_right, 3),
        "sweep_seconds": round(t_left + t_right, 3), "energy_left_half": e_left, "energy": e_right,
        "heff_applications": nmv, "heff_tflops_sweep_average": round(flops / max(ms, 1e-9) / 1e9, 3),
        "phase_seconds": {k: round(v, 3) for k, v in phases.items()},
    }), flush=True)


if __name__ == "__main__":
    main()
