#!/usr/bin/env python3
"""Benchmark of the two-site DMRG effective-Hamiltonian path on one MI355X.

Workload (the metric's configuration): random ab-initio Hamiltonian with D2h symmetry
(chemps2_amd.problems.random_abinitio, seed 0), L = 32 orbitals, N = 32 electrons,
2S = 0, virtual dimension D = 2000.

1. Warm-up (FP64, untimed): DMRG from a random MPS with the schedule D = 250 (2 sweeps),
   500, 1000, 2000 (one sweep each; Davidson tolerance 1e-5, no noise), the way
   ConvergenceScheme instructions run in DMRG::Solve (DMRG.cpp:268-354). The state is
   saved in the reference's MPS checkpoint layout.
2. Timed unit: one full left+right sweep at D = 2000 from that state (DMRG::Solve with
   one instruction after DMRG::PreSolve: Join, Heff preparation, Davidson, Split and the
   operator updates of every site on the device), in each Heff form: fp64, mixed and
   mixed+learned (the learned preconditioner of chemps2_amd/data/precond_mlp.npy on the
   mixed form). Each form starts from the same checkpoint. Reported per form: sweep
   seconds and phases, Heff applications, the Heff.x rate over the sweep
   (sum n F_H / sum of event-timed application time), TotalMinEnergy and its difference
   to the fp64 form, per-site energies and product counts.
3. value = the Heff.x TFLOP/s of the fp64 form over every application of the timed
   sweep (F_H: the GEMM flops of the reference's makeHeff at each site); steps = the
   number of those applications, ms_per_step their mean device time.
4. Site benchmark: --warmup + --steps applications at sites 15-16 of the warmed-up state
   (x, y and all operators resident in HBM), event-timed; its Heff roofline follows
   F_H / t against min(FP64 peak, F_H / B_H x HBM bandwidth).
5. CPU baseline: the reference CheMPS2 library built from its sources (oracle/_ref) on
   the host cores: its first left-sweep steps from the same checkpoint (the same
   Heff matrices), Davidson products x F_H over the Davidson time of those steps.

One JSON line goes to stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
WARM_SCHEDULE = [(250, 2), (500, 1), (1000, 1), (2000, 1)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per application from a committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("bytes_per_launch"), "profiles/" + name
    except Exception:
        return None, None


def roofline_record(F_H, B_H, ms, traffic, traffic_source):
    """The roofline of one Heff application: F_H flops in ms milliseconds of kernels
    against min(FP64 matrix-core peak, F_H / B_H x HBM bandwidth)."""
    tflops = F_H / (ms * 1e-3) / 1e12
    ai = F_H / B_H
    peak = min(PEAK_FP64_TFLOPS, ai * PEAK_HBM_GBS / 1e3)
    return {
        "bound": "mfma" if ai * PEAK_HBM_GBS / 1e3 >= PEAK_FP64_TFLOPS else "hbm",
        "achieved": round(tflops, 4),
        "peak": round(peak, 2),
        "unit": "TFLOP/s",
        "frac": round(tflops / peak, 4),
        "traffic": traffic,
        "kernel": "Heff x at sites 15-16 (all its kernels), F_H / t",
        "F_H": F_H,
        "B_H": B_H,
        "traffic_over_B_H": round(traffic / B_H, 2) if traffic else None,
        "traffic_source": traffic_source,
    }


def traffic_from(fetch_csv, write_csv, steps, out_name):
    """traffic_numbers() of the two counter files, written to profiles/out_name."""
    out = traffic_numbers(fetch_csv, write_csv, steps)
    with open(os.path.join(ROOT, "profiles", out_name), "w") as f:
        json.dump(out, f, indent=1)
    return out


def traffic_numbers(fetch_csv, write_csv, steps):
    """HBM bytes per Heff application from two rocprofv3 --pmc runs of --site-from (one
    with FETCH_SIZE, one with WRITE_SIZE, in KiB per dispatch): the counted products are
    the dispatches between the two 1 GiB copy marks, and each counter is scaled by what it
    reads for the mark itself (1 GiB read, 1 GiB written)."""
    import csv

    def per_dispatch(path, counter):
        rows = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                d = int(r["Dispatch_Id"])
                rows[d] = (r["Kernel_Name"], float(r["Counter_Value"]))
        return [rows[d] for d in sorted(rows)]

    def window(seq):
        marks = [i for i, (n, _) in enumerate(seq) if "copyBuffer" in n or "elementwise" in n]
        big = sorted(marks, key=lambda i: -seq[i][1])[:2]
        i0, i1 = sorted(big)
        return seq[i0][1], [v for _, v in seq[i0 + 1:i1]], seq[i1][1]

    f0, fk, f1 = window(per_dispatch(fetch_csv, "FETCH_SIZE"))
    w0, wk, w1 = window(per_dispatch(write_csv, "WRITE_SIZE"))
    gib_kib = COPY_MARK_DOUBLES * 8 / 1024.0
    fscale = gib_kib / (0.5 * (f0 + f1))
    wscale = gib_kib / (0.5 * (w0 + w1))
    rd = sum(fk) * fscale * 1024.0 / steps
    wr = sum(wk) * wscale * 1024.0 / steps
    out = {
        "kernel": "every kernel of one Heff application (k_bspmv, k_tiles_stream1, k_tiles_wave1, k_tiles_wave2, "
                  "k_tiles_fit, k_fchain, k_chain and their reductions)",
        "workload": "Heff x at sites 15-16 of the warmed-up L=32 D=2000 random D2h state",
        "bytes_per_launch": rd + wr, "read_bytes": rd, "write_bytes": wr,
        "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4),
        "dispatches_per_application": len(fk) / steps, "applications": steps,
    }
    return out


def new_ctx(L, irr, mx, D, form, params, ckpt):
    import chemps2_amd

    c = chemps2_amd.Context(0)
    c.set_heff_form(form)
    if params is not None:
        c.mlp_set(params)
        c.precond_mode(chemps2_amd.PRECOND_LEARNED)
    c.set_problem(L, 7, irr, L, 0, 0, mx)
    c.set_sectors_auto(D)  # FCI dimensions for Split
    c.mps_load_dims(ckpt)
    c.mps_load(ckpt)
    return c


def warm_up(L, D, seed, irr, mx, rtol, ckpt):
    import chemps2_amd

    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, L, 0, 0, mx)
    t0 = time.time()
    c.dmrg_init(WARM_SCHEDULE[0][0], seed + 1)
    out = []
    for Dw, nsw in WARM_SCHEDULE:
        t1 = time.time()
        e = c.dmrg_solve([(min(Dw, D), 1e-12, nsw, 0.0, rtol)])
        out.append({"D": min(Dw, D), "sweeps": nsw, "energy": e, "seconds": round(time.time() - t1, 2)})
        log(f"warm-up D={min(Dw, D)} x{nsw}: E={e:.10f} ({time.time() - t1:.1f}s)")
    c.mps_save(ckpt)
    dims = [int(sum(s[3] for s in c.get_sectors(b))) for b in range(L + 1)]
    c.close()
    return {"seconds": round(time.time() - t0, 2), "schedule": out, "total_dim_per_boundary": dims}


COPY_MARK_DOUBLES = 1 << 27  # 1 GiB: the plain copy that brackets the counted products


def copy_mark():
    """A plain device copy of 1 GiB (torch, hipMemcpy device to device): the counter
    profiles use it to scale FETCH_SIZE and WRITE_SIZE and to find the counted products."""
    import torch

    a = torch.empty(COPY_MARK_DOUBLES, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def site_bench(L, D, irr, mx, ckpt, site, steps, warmup, marks=False):
    """Heff.x at sites (site, site+1) of the warmed-up state: the left half sweep down to
    site+1 (fixed dimensions) builds the right operators, then steps timed applications.
    marks: a 1 GiB copy before and after the timed applications (counter profiles)."""
    import numpy as np
    import torch

    import chemps2_amd

    c = new_ctx(L, irr, mx, D, chemps2_amd.FORM_FP64, None, ckpt)
    c.dmrg_presolve()
    c.dmrg_sweep_range(False, L - 2, site + 1, D, 1e-5, 0.0, False)
    c.join(site)
    n = c.prepare(site)
    F_H, B_H = c.flops()
    st = c.plan_stats()
    torch.cuda.set_device(0)
    x = torch.from_numpy(np.random.default_rng(0).normal(size=n)).to("cuda")
    y = torch.empty_like(x)
    stream = torch.cuda.ExternalStream(c.stream())
    for _ in range(warmup):
        c.apply(x, y)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if marks:
        copy_mark()
    e0.record(stream)
    for _ in range(steps):
        c.apply(x, y)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    if marks:
        copy_mark()
    c.timing(True)
    for _ in range(steps):
        c.apply(x, y)
    kms, cnt = c.timing_read()
    c.timing(False)
    c.close()
    # device time of the product's kernels (events around its kernel groups on the
    # library stream); ms above also holds the host gaps between the calls
    ms_k = sum(kms) / max(cnt, 1)
    traffic, tsrc = read_traffic("heff_site15_L32_D2000_traffic.json")
    roof = roofline_record(F_H, B_H, ms_k, traffic, tsrc)
    tflops, ai = roof["achieved"], F_H / B_H
    log(f"site {site}: V={n} F_H={F_H:.4e} B_H={B_H:.4e} AI={ai:.1f} flop/B: {ms_k:.3f} ms of kernels "
        f"({ms:.3f} ms per call), {tflops:.2f} TFLOP/s "
        f"(kernels: tiles {kms[0] / max(cnt, 1):.3f}, explicit {kms[1] / max(cnt, 1):.3f}, "
        f"chained {kms[2] / max(cnt, 1):.3f} ms)")
    return {
        "site": site, "V": n, "F_H": F_H, "B_H": B_H, "ms_kernels": round(ms_k, 4), "ms_per_call": round(ms, 4),
        "tflops": round(tflops, 4),
        "kernel_ms": [round(v / max(cnt, 1), 4) for v in kms], "plan_stats": st,
    }, roof


def timed_sweep(name, L, D, irr, mx, ckpt, rtol, form, params):
    c = new_ctx(L, irr, mx, D, form, params, ckpt)
    t0 = time.time()
    c.dmrg_presolve()
    t_init = time.time() - t0
    t0 = time.time()
    e = c.dmrg_solve([(D, 1e-12, 1, 0.0, rtol)])
    secs = time.time() - t0
    phases, nmv = c.dmrg_stats()
    hf, hms = c.dmrg_heff_stats()
    sites = c.site_log()
    low = c.mixed_stats()[0]
    c.close()
    rate = hf / max(hms, 1e-9) / 1e9
    log(f"sweep ({name}): {secs:.1f}s after PreSolve {t_init:.1f}s; E={e:.12f}; {nmv} Heff applications "
        f"({low} hi-only), {rate:.2f} TFLOP/s over the sweep; phases {dict((k, round(v, 2)) for k, v in phases.items())}")
    return {
        "seconds": round(secs, 3),
        "presolve_seconds": round(t_init, 3),
        "total_min_energy": e,
        "heff_applications": nmv,
        "heff_applications_hi_only": low,
        "heff_flops": hf,
        "heff_device_seconds": round(hms / 1e3, 4),
        "heff_tflops_sweep_average": round(rate, 4),
        "phase_seconds": {k: round(v, 3) for k, v in phases.items()},
        "sites": [{"site": s["site"], "energy": s["energy"], "products": s["products"],
                   "heff_ms": round(s["heff_ms"], 3), "F_H": s["F_H"]} for s in sites],
    }


def cpu_baseline(L, D, irr, T, V, ckpt, rtol, budget_s, fp64_sites):
    """The reference library (oracle/_ref, OpenMP + MKL) from the same checkpoint: its
    first left-sweep steps until about budget_s seconds of steps."""
    from oracle import refdrv

    if not refdrv.available():
        log("cpu: oracle/_ref is not built; no CPU baseline")
        return None
    threads = os.cpu_count()
    with tempfile.TemporaryDirectory() as tmp:
        prob = os.path.join(tmp, "problem.bin")
        refdrv.write_problem(prob, L, 7, list(irr), L, 0, 0, T, V)
        # the first left-sweep sites of the timed fp64 sweep, with their F_H (same matrices)
        fh = {s["site"]: s["F_H"] for s in fp64_sites[:L - 2]}
        gpu_prod = {s["site"]: s["products"] for s in fp64_sites[:L - 2]}
        t0 = time.time()
        res = refdrv.steps(prob, D, L - 2, 1, rtol, budget_s=budget_s, threads=threads, chkpt_file=ckpt)
        log(f"cpu: reference PreSolve {res['init_seconds']:.1f}s, {len(res['steps'])} steps in "
            f"{sum(st[2] for st in res['steps']):.1f}s (wall {time.time() - t0:.1f}s)")
        flops, secs, e_diff = 0.0, 0.0, 0.0
        gpu_e = {s["site"]: s["energy"] for s in fp64_sites[:L - 2]}
        for idx, e, s, nit in res["steps"]:
            if nit is None or idx not in fh:
                continue
            flops += nit * fh[idx]
            secs += s
            e_diff = max(e_diff, abs(e - gpu_e[idx]))
        steps_desc = ", ".join(f"{idx}:{nit}" for idx, _, _, nit in res["steps"])
    return {
        "value": flops / max(secs, 1e-9) / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference CheMPS2 (oracle/_ref, OpenMP + MKL, {threads} threads): the first {len(res['steps'])} "
                   f"steps of the D={D} left sweep from the same checkpoint (site:Davidson products {steps_desc}; "
                   f"GPU products on those sites {[gpu_prod.get(i[0]) for i in res['steps']]}); Davidson products x "
                   f"F_H over {secs:.1f} s of solve+update time; reference PreSolve {res['init_seconds']:.1f} s not "
                   f"counted; largest |E_ref - E_gpu| over the steps {e_diff:.2e}"),
        "seconds_per_step": [round(st[2], 3) for st in res["steps"]],
        "presolve_seconds": res["init_seconds"],
    }


N2_PUBLISHED = {"D": 2000, "seconds_per_sweep": 219.0, "energy": -109.28216077, "discarded_weight": 1.8910e-07,
                "hardware": "16 Intel Xeon E5-2670 cores @ 2.6 GHz, OpenMP", "source": "sphinx/resources.rst:45-56"}


def n2_ccpvdz(rtol):
    """The reference's published sweep point: N2/cc-pVDZ (tests/golden/N2.CCPVDZ.FCIDUMP,
    14 electrons in 28 orbitals, D2h, orbitals reordered as Problem::SetupReorderD2h does,
    Davidson tolerance 1e-5). Warm-up D = 250, 500, 1000, 1500 (one sweep each), then one
    timed left+right sweep at D = 2000, beside the published 219 s per sweep."""
    import chemps2_amd
    from chemps2_amd.problems import molpro_to_xor, mx_table, permute, read_fcidump, reorder_d2h

    d = read_fcidump(os.path.join(ROOT, "tests", "golden", "N2.CCPVDZ.FCIDUMP"))
    irr = [molpro_to_xor(7, m) for m in d["molpro_irreps"]]
    irr2, T2, V2 = permute(irr, d["T"], d["V"], reorder_d2h(irr))
    L, N = d["L"], d["N"]
    mx = mx_table(T2, V2, N)
    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, [int(v) for v in irr2], N, 0, 0, mx)
    t0 = time.time()
    c.dmrg_init(250, 1)
    for Dw in (250, 500, 1000, 1500):
        c.dmrg_solve([(Dw, 1e-12, 1, 0.0, rtol)])
    warm = time.time() - t0
    t0 = time.time()
    e = c.dmrg_solve([(2000, 1e-12, 1, 0.0, rtol)]) + d["econst"]
    secs = time.time() - t0
    phases, nmv = c.dmrg_stats()
    hf, hms = c.dmrg_heff_stats()
    c.close()
    log(f"N2/cc-pVDZ: D=2000 sweep {secs:.1f}s (warm-up {warm:.1f}s), E={e:.8f}; published "
        f"{N2_PUBLISHED['seconds_per_sweep']} s, E={N2_PUBLISHED['energy']}")
    return {"seconds_per_sweep": round(secs, 3), "energy": e, "warm_up_seconds": round(warm, 2),
            "heff_applications": nmv, "heff_tflops_sweep_average": round(hf / max(hms, 1e-9) / 1e9, 4),
            "phase_seconds": {k: round(v, 3) for k, v in phases.items()},
            "published": N2_PUBLISHED,
            "speedup_vs_published": round(N2_PUBLISHED["seconds_per_sweep"] / secs, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--D", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rtol", type=float, default=1e-5)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-n2", action="store_true", help="skip the N2/cc-pVDZ D=2000 sweep beside the published one")
    ap.add_argument("--save-warm", metavar="H5", help="only the warm-up: save its state to H5 and exit")
    ap.add_argument("--site-from", metavar="H5", help="only the site benchmark, from a state --save-warm wrote, "
                    "its timed applications between two 1 GiB copy marks (for rocprofv3 --pmc)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/" + "heff_site15_L32_D2000_traffic.json from the counter_collection.csv "
                    "of two --site-from runs under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (--steps as in those runs)")
    args = ap.parse_args()
    if args.traffic_from:
        print(json.dumps(traffic_from(args.traffic_from[0], args.traffic_from[1], args.steps,
                                      "heff_site15_L32_D2000_traffic.json")))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    # torch's HIP context first: it carries the site benchmark's device vectors, and it
    # must not start after the library has taken most of the device memory
    torch.cuda.init()
    torch.empty(1, device="cuda:0")

    import chemps2_amd
    from chemps2_amd.problems import mx_table, random_abinitio

    L, D, seed = args.L, args.D, args.seed
    irr, T, V = random_abinitio(L, seed, group=7)
    mx = mx_table(T, V, L)
    site = L // 2 - 1
    if args.site_from:
        sb, roof = site_bench(L, D, irr, mx, args.site_from, site, args.steps, args.warmup, marks=True)
        print(json.dumps({"site_bench": sb, "roofline": roof}), flush=True)
        return
    tmp = tempfile.mkdtemp(prefix="chm_bench_")
    ckpt = args.save_warm or os.path.join(tmp, "warm.h5")
    warm = warm_up(L, D, seed, irr, mx, args.rtol, ckpt)
    if args.save_warm:
        print(json.dumps({"warm_up": warm, "checkpoint": ckpt}), flush=True)
        return

    sb, roof = site_bench(L, D, irr, mx, ckpt, site, args.steps, args.warmup)

    wpath = os.path.join(ROOT, "chemps2_amd", "data", "precond_mlp.npy")
    params = np.load(wpath, allow_pickle=False) if (not args.no_learned and os.path.exists(wpath)) else None
    forms = [("fp64", chemps2_amd.FORM_FP64, None)]
    if not args.no_mixed:
        forms.append(("mixed", chemps2_amd.FORM_MIXED, None))
        if params is not None:
            forms.append(("mixed+learned", chemps2_amd.FORM_MIXED, params))
    sweeps = {}
    for name, form, prm in forms:
        sweeps[name] = timed_sweep(name, L, D, irr, mx, ckpt, args.rtol, form, prm)
    ref = sweeps["fp64"]
    for name, s in sweeps.items():
        if name == "fp64":
            continue
        s["delta_E_vs_fp64"] = abs(s["total_min_energy"] - ref["total_min_energy"])
        s["speedup_vs_fp64"] = round(ref["seconds"] / s["seconds"], 3)
        log(f"{name}: |dE| vs fp64 {s['delta_E_vs_fp64']:.2e}, sweep {s['seconds']:.1f}s vs {ref['seconds']:.1f}s")

    n2 = None
    if not args.no_n2:
        n2 = n2_ccpvdz(args.rtol)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(L, D, irr, T, V, ckpt, args.rtol, args.cpu_seconds, ref["sites"])
        except Exception as e:  # the baseline is reported, never the product
            log(f"cpu: reference baseline failed: {e}")
    try:
        os.remove(ckpt)
        os.rmdir(tmp)
    except OSError:
        pass

    out = {
        "metric": "Heff·x TFLOP/s + Davidson sweep time at L=32 D=2000; |ΔE| vs FP64 ≤1e-8 Ha",
        "value": ref["heff_tflops_sweep_average"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": ref["heff_applications"],
        "warmup": args.warmup,
        "ms_per_step": round(ref["heff_device_seconds"] * 1e3 / max(ref["heff_applications"], 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("Heff x over one left+right DMRG sweep at D=2000 after a D=250/500/1000/2000 warm-up, "
                         "random ab-initio D2h Hamiltonian (seed 0), L=32 N=32 2S=0"),
            "L": L, "D": D, "davidson_rtol": args.rtol,
            "sweep_seconds_fp64": ref["seconds"],
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "warm_up": warm,
        "site_bench": sb,
        "sweep": sweeps,
        "n2_ccpvdz_published_point": n2,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
