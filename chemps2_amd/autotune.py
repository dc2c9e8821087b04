"""Autotuner of the Heff routing knobs, run on the GPU:

    python -m chemps2_amd.autotune [--out chemps2_amd/data/autotune.json]

For two-site steps of real DMRG states (the L = 32, D = 2000 and L = 24, D = 1000
random D2h Hamiltonians after a warm-up schedule, and the 4 x 8 Hubbard cylinder at
D = 1000), every candidate setting of the knobs that chm_heff_prepare routes by
(explicit_max: the largest |kappa||kappa'| stored as explicit sector matrix;
fchain_min: chained blocks whose largest dimension is above it go to the fused MFMA
kernel k_fchain; item_terms / group_terms: terms per work item and per input group of
the chained kernels; chain_mfma: the one-wavefront MFMA chained kernel for blocks up to
48) is prepared and timed with HIP events (chm_timing_*). Sites are grouped by their
largest block dimension; each group keeps the setting with the least summed product
time over its sites (a coordinate search: explicit_max first, then fchain_min, then
item_terms, then chain_mfma).
The table, with the measurements it came from, goes to data/autotune.json; Context
loads it at creation (chm_set_tuning), and chm_heff_prepare looks up each site in it.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "autotune.json")

EXPLICIT = [0, 1024, 4096, 16384]
ITEMS = [32, 64, 128]
MFMA = [0, 1]
FCHAIN = [0, 16, 32, 48, 64]
# largest block dimension classes of the table
DMAX_EDGES = [1, 8, 16, 24, 32, 48, 64, 128, 1 << 30]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_site(c, site, knobs, reps):
    c.set_heff_knobs(*knobs)
    c.prepare(site)
    x = np.random.default_rng(site).normal(size=c.V)
    c.apply(x)  # first application outside the timing
    c.timing(True)
    for _ in range(reps):
        c.apply(x)
    ms, cnt = c.timing_read()
    c.timing(False)
    return sum(ms) / max(cnt, 1)


def tune_state(c, L, D, sites, reps, records, label):
    """Left half sweep of the loaded state, stopping at each site of `sites` (descending)
    to time every candidate there."""
    c.dmrg_presolve()
    first = L - 2
    for s in sorted(sites, reverse=True):
        if s + 1 <= first:
            c.dmrg_sweep_range(False, first, s + 1, D, 1e-5, 0.0, False)
        first = s
        c.join(s)
        base = [4096, 64, 32, 0, 32]
        res = {}

        def run(k):
            key = tuple(k)
            if key not in res:
                res[key] = time_site(c, s, k, reps)
            return res[key]

        def best(cands, at):
            for v in cands:
                k = list(base)
                k[at] = v
                run(k)
            return min(cands, key=lambda v: res[tuple(base[:at] + [v] + base[at + 1:])])

        base[0] = best(EXPLICIT, 0)
        base[4] = best(FCHAIN, 4)
        for it in ITEMS:
            run([base[0], it, max(8, it // 2), base[3], base[4]])
        best_it = min(ITEMS, key=lambda it: res[(base[0], it, max(8, it // 2), base[3], base[4])])
        base[1], base[2] = best_it, max(8, best_it // 2)
        base[3] = best(MFMA, 3)
        used = c.tuning_used()
        rec = {"workload": label, "site": s, "V": int(c.V), "dmax": int(used["dmax"]),
               "times_ms": [[list(k), round(v, 4)] for k, v in res.items()]}
        records.append(rec)
        top = min(res, key=res.get)
        log(f"{label} site {s}: V={c.V} dmax={used['dmax']}: best {top} {res[top]:.3f} ms "
            f"(default {res.get((4096, 64, 32, 0, 32), float('nan')):.3f} ms)")


def build_table(records):
    rows = []
    for lo, hi in zip(DMAX_EDGES[:-1], DMAX_EDGES[1:]):
        grp = [r for r in records if lo <= r["dmax"] < hi]
        if not grp:
            continue
        # candidates measured on every site of the class; the least summed time wins
        keys = None
        for r in grp:
            ks = {tuple(k) for k, _ in r["times_ms"]}
            keys = ks if keys is None else keys & ks
        if not keys:
            continue
        tot = {k: sum(dict((tuple(kk), v) for kk, v in r["times_ms"])[k] for r in grp) for k in keys}
        best = min(tot, key=tot.get)
        rows.append([lo, hi - 1, 0, 1 << 62, best[0], best[1], best[2], best[3], best[4]])
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="one small workload (checks the script)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import chemps2_amd
    from chemps2_amd.problems import hubbard_cylinder, mx_table, random_abinitio

    workloads = []
    if args.quick:
        workloads.append(("abinitio L=16 D=300", 16, 300, 7, 16, lambda: random_abinitio(16, 2, group=7),
                          [(150, 1), (300, 1)], [3, 7, 11]))
    else:
        workloads.append(("abinitio L=32 D=2000", 32, 2000, 7, 32, lambda: random_abinitio(32, 0, group=7),
                          [(250, 2), (500, 1), (1000, 1), (2000, 1)], [3, 6, 9, 12, 15, 18, 22, 26]))
        workloads.append(("abinitio L=24 D=1000", 24, 1000, 7, 24, lambda: random_abinitio(24, 0, group=7),
                          [(250, 1), (500, 1), (1000, 1)], [4, 8, 11, 15, 19]))

        def cyl():
            T, V, _ = hubbard_cylinder(8, 4, U=8.0)
            return np.zeros(32, dtype=np.int32), T, V

        workloads.append(("cylinder 4x8 D=1000", 32, 1000, 0, 28, cyl, [(250, 1), (1000, 1)], [8, 15, 22]))
    records = []
    t0 = time.time()
    for label, L, D, group, N, mk, sched, sites in workloads:
        irr, T, V = mk()
        mx = mx_table(T, V, N)
        c = chemps2_amd.Context(0, tuning=False)
        c.set_problem(L, group, irr, N, 0, 0, mx)
        c.dmrg_init(sched[0][0], 1)
        for Dw, ns in sched:
            c.dmrg_solve([(Dw, 1e-12, ns, 0.0, 1e-5)])
        with tempfile.TemporaryDirectory() as tmp:
            ck = os.path.join(tmp, "s.h5")
            c.mps_save(ck)
            c.close()
            c = chemps2_amd.Context(0, tuning=False)
            c.set_problem(L, group, irr, N, 0, 0, mx)
            c.set_sectors_auto(D)
            c.mps_load_dims(ck)
            c.mps_load(ck)
        tune_state(c, L, D, sites, args.reps, records, label)
        c.close()
    rows = build_table(records)
    out = {"columns": ["dmax_lo", "dmax_hi", "V_lo", "V_hi", "explicit_max", "item_terms", "group_terms",
                       "chain_mfma", "fchain_min"],
           "rows": rows, "measurements": records, "seconds": round(time.time() - t0, 1)}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log(f"table ({len(rows)} rows) -> {args.out}")
    for r in rows:
        log("  ", r)


if __name__ == "__main__":
    main()
