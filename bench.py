#!/usr/bin/env python3
"""Benchmark of the two-site effective-Hamiltonian product Heff x on one MI355X.

Workload: random ab-initio Hamiltonian with D2h symmetry (chemps2_amd.problems.
random_abinitio, seed 0), L = 32 orbitals, N = 32 electrons, 2S = 0, virtual dimension
D = 2000 distributed over the sectors as SyBookkeeper does; the two-site step at the
middle of the chain (sites 15, 16) with pseudo-random operator tensors of the
reference's shapes. A step is one application y = Heff x with x, y and all operators
already resident in HBM. value = F_H / time, where F_H is the GEMM flop count of the
reference's makeHeff at this site.

The second half of the metric, the sweep time, is measured on the same problem: a
random left-normalised MPS with the SyBookkeeper dimensions of D = 2000, the operators
of every boundary built on the GPU (DMRG::PreSolve), then the first left+right sweep
of DMRG::Solve at D = 2000 (sweepleft over sites 30..1 with fixed dimensions, then
sweepright over sites 0..29 with truncation to D): Join, Heff preparation, Davidson
(tolerance 1e-5), Split and the operator update of every site, all on the device. It
is reported under "sweep", with the phase times and the Heff.x rate averaged over
every application of the sweep (sum of n_matvec * F_H over the summed device time).

Both are also measured in the mixed form ("mixed"): the explicit sector matrix stored as
float pairs, hi-only products in the Davidson until its residual is below 1e-5, hi + lo
products after that, and once more with the learned preconditioner on top of the mixed
form ("mixed+learned", parameters from chemps2_amd/data/precond_mlp.npy, trained on
other Hamiltonians). The sweeps at Davidson tolerance 1e-5 follow different Davidson
trajectories from the random start, so |dE| between the forms is checked separately
("same_site_check"): in a separate context, the FP64 left half runs up to sites 15-16
and the Davidson of that step runs from the joined two-site tensor at tolerance 1e-8 in
all three forms.

One JSON line goes to stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprof PMC runs."""
    path = os.path.join(ROOT, "profiles", "bspmv_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(L, D, site, seed, irr, T, V, x, y_gpu, budget_s):
    """The oracle (FP64 restatement of the reference, OpenMP over output blocks like
    Heff.cpp:54-64) on a bounded sample of output blocks of the same step; also checks
    the GPU result on that sample."""
    import numpy as np

    from oracle.wrap import OracleDMRG, threads

    o = OracleDMRG(L, 7, irr, L, 0, 0, T, V)
    t0 = time.time()
    n = o.synthetic(D, site, seed + 1, seed + 2)
    log(f"cpu: oracle workload built in {time.time() - t0:.1f}s (V={n})")
    lab, k2i = o.kappa()
    nk = len(lab)
    # grow the sample geometrically until it costs about budget_s seconds
    k1, flops, secs = 0, 0.0, 0.0
    step = max(1, nk // 512)
    ys = np.zeros(n)
    while k1 < nk and secs < budget_s:
        k_next = min(nk, k1 + step)
        y, f, s = o.apply_range(x, k1, k_next)
        ys[k2i[k1]:k2i[k_next]] = y[k2i[k1]:k2i[k_next]]
        flops += f
        secs += s
        k1 = k_next
        step *= 2
    a, b = k2i[0], k2i[k1]
    err = float(np.abs(ys[a:b] - y_gpu[a:b]).max() / max(1e-300, np.abs(ys[a:b]).max()))
    # a whole step cheaper than the budget: repeat whole steps until it is spent
    passes = 1
    if k1 == nk:
        while secs < budget_s:
            _, f, s = o.apply_range(x, 0, nk)
            flops += f
            secs += s
            passes += 1
    log(f"cpu: {k1}/{nk} output blocks x {passes} in {secs:.1f}s, {flops:.3e} flop; GPU vs oracle on the sample: {err:.2e}")
    what = (f"{passes} whole Heff x steps ({nk} output blocks)" if k1 == nk else
            f"Heff x on output blocks 0..{k1 - 1} of {nk}")
    return {
        "value": flops / secs / 1e12,
        "unit": "TFLOP/s",
        "cores": threads(),
        "kind": "port",
        "sample": f"oracle {what} at the same site and data ({flops:.3e} flop, {secs:.1f} s, {threads()} threads)",
    }, err


def timed_apply(ctx, x, y, steps, warmup, f32=False):
    """ms per application (events on the library stream) and per-kernel ms."""
    import torch

    stream = torch.cuda.ExternalStream(ctx.stream())
    torch.cuda.synchronize()
    for _ in range(warmup):
        ctx.apply(x, y, f32=f32)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        ctx.apply(x, y, f32=f32)
    e1.record(stream)
    e1.synchronize()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    ctx.timing(True)
    for _ in range(steps):
        ctx.apply(x, y, f32=f32)
    kms, cnt = ctx.timing_read()
    ctx.timing(False)
    return ms, [v / max(cnt, 1) for v in kms]


def same_site_check(sw, site, rtol, forms):
    """Davidson at `site` from the joined two-site tensor in every form, on the operators
    left by the sweep (same matrix for all forms up to their product precision)."""
    import chemps2_amd

    sw.join(site)
    guess = sw.s_get()
    out = {"site": site, "davidson_rtol": rtol}
    for name, form, prm in forms:
        sw.set_heff_form(form)
        sw.precond_mode(chemps2_amd.PRECOND_DIAGONAL)
        if prm is not None:
            sw.mlp_set(prm)
            sw.precond_mode(chemps2_amd.PRECOND_LEARNED)
        sw.prepare(site)
        g = guess.copy()
        t0 = time.time()
        e, nm = sw.davidson(g, rtol)
        out[name] = {"energy": e, "products": nm, "seconds": round(time.time() - t0, 3)}
    sw.set_heff_form(chemps2_amd.FORM_FP64)
    sw.precond_mode(chemps2_amd.PRECOND_DIAGONAL)
    for name in [f[0] for f in forms[1:]]:
        out[name]["delta_E_vs_fp64"] = abs(out[name]["energy"] - out["fp64"]["energy"])
        log(f"same-site check ({name}, site {site}, rtol {rtol:g}): |dE| vs FP64 = {out[name]['delta_E_vs_fp64']:.2e}, "
            f"{out[name]['products']} products (FP64: {out['fp64']['products']})")
    return out


def run_check(L, D, seed, irr, mx, rtol, check_rtol, forms):
    """A separate context runs the first left half up to sites L/2 - 1, L/2 and does the
    same-site check there (so the timed sweeps are not interrupted)."""
    import chemps2_amd

    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, L, 0, 0, mx)
    c.dmrg_init(D, seed + 1)
    c.dmrg_sweep_range(False, L - 2, L // 2, D, rtol, 0.0, False)
    out = same_site_check(c, L // 2 - 1, check_rtol, forms)
    c.close()
    return out


def run_sweep(L, D, seed, irr, mx, rtol, form, precond=None):
    """Initialisation and the first left+right sweep of DMRG::Solve in one Heff form
    (precond: parameters of the learned preconditioner, or None for the diagonal)."""
    import chemps2_amd

    name = ("mixed" if form == chemps2_amd.FORM_MIXED else "fp64") + ("+learned" if precond is not None else "")
    sw = chemps2_amd.Context(0)
    sw.set_heff_form(form)
    if precond is not None:
        sw.mlp_set(precond)
        sw.precond_mode(chemps2_amd.PRECOND_LEARNED)
    sw.set_problem(L, 7, irr, L, 0, 0, mx)
    t0 = time.time()
    sw.dmrg_init(D, seed + 1)
    t_init = time.time() - t0
    log(f"sweep ({name}): MPS and operators of every boundary built in {t_init:.1f}s")
    t0 = time.time()
    e_left = sw.dmrg_sweep(False, D, rtol, 0.0, False)
    t_left = time.time() - t0
    log(f"sweep ({name}): left half (sites {L - 2}..1) in {t_left:.1f}s, energy {e_left:.10f}")
    t1 = time.time()
    e_right = sw.dmrg_sweep(True, D, rtol, 0.0, True)
    t_right = time.time() - t1
    phases, nmv = sw.dmrg_stats()
    hf, hms = sw.dmrg_heff_stats()
    low = sw.mixed_stats()[0]
    log(f"sweep ({name}): right half (sites 0..{L - 3}) in {t_right:.1f}s, energy {e_right:.10f}; {nmv} Heff "
        f"applications ({low} hi-only), {hf / max(hms, 1e-9) / 1e9:.2f} TFLOP/s over the sweep; phases {phases}")
    sw.close()
    return {
        "seconds": round(t_left + t_right, 3),
        "left_seconds": round(t_left, 3),
        "right_seconds": round(t_right, 3),
        "init_seconds": round(t_init, 3),
        "D": D,
        "davidson_rtol": rtol,
        "heff_applications": nmv,
        "heff_applications_hi_only": low,
        "heff_tflops_sweep_average": round(hf / max(hms, 1e-9) / 1e9, 4),
        "heff_device_seconds": round(hms / 1e3, 3),
        "energy_left_half": e_left,
        "energy_last_site": e_right,
        "phase_seconds": {k: round(v, 3) for k, v in phases.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--D", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-rtol", type=float, default=1e-5)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--check-rtol", type=float, default=1e-8)
    ap.add_argument("--no-learned", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import chemps2_amd
    from chemps2_amd.problems import mx_table, random_abinitio

    L, D, seed = args.L, args.D, args.seed
    site = L // 2 - 1
    irr, T, V = random_abinitio(L, seed, group=7)
    ctx = chemps2_amd.Context(0)
    ctx.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
    ctx.set_sectors_auto(D)
    ctx.fill_synthetic(site - 1, 1, seed + 1)
    ctx.fill_synthetic(site + 1, 0, seed + 2)
    t0 = time.time()
    Vlen = ctx.prepare(site)
    t_prep = time.time() - t0
    F_H, B_H = ctx.flops()
    st = ctx.plan_stats()
    log(f"prepare: {t_prep:.1f}s V={Vlen} F_H={F_H:.4e} stats={st}")

    torch.cuda.set_device(0)
    rng = np.random.default_rng(seed)
    xh = rng.normal(size=Vlen)
    x = torch.from_numpy(xh).to("cuda")
    y = torch.empty_like(x)
    ms, kms = timed_apply(ctx, x, y, args.steps, args.warmup)
    y_gpu = y.cpu().numpy()
    log(f"Heff x: {ms:.3f} ms per application, {F_H / ms / 1e9:.2f} TFLOP/s (F_H basis)")
    log(f"kernels (ms per application): GEMM tiles {kms[0]:.4f}, explicit block product {kms[1]:.4f}, "
        f"chained kernels {kms[2]:.4f}")
    nnz = st["explicit_doubles"]
    ncol_bytes = 4.0 * st.get("explicit_cols", 0)
    bspmv_bytes = 8.0 * nnz + 16.0 * Vlen + ncol_bytes
    achieved = bspmv_bytes / (kms[1] * 1e-3) / 1e9 if kms[1] > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": read_traffic(),
    }

    mixed = None
    if not args.no_mixed:
        # the same site prepared in the mixed form (same bytes for the explicit matrix)
        ctx.set_heff_form(chemps2_amd.FORM_MIXED)
        t0 = time.time()
        ctx.prepare(site)
        t_prep_m = time.time() - t0
        ms32, kms32 = timed_apply(ctx, x, y, args.steps, args.warmup, f32=True)
        d32 = float(np.abs(y.cpu().numpy() - y_gpu).max() / np.abs(y_gpu).max())
        msr, kmsr = timed_apply(ctx, x, y, args.steps, args.warmup)
        dr = float(np.abs(y.cpu().numpy() - y_gpu).max() / np.abs(y_gpu).max())
        b32 = 4.0 * nnz + 16.0 * Vlen + ncol_bytes
        log(f"mixed: prepare {t_prep_m:.1f}s; hi-only Heff x {ms32:.3f} ms ({F_H / ms32 / 1e9:.2f} TFLOP/s, explicit product "
            f"{kms32[1]:.4f} ms = {b32 / (kms32[1] * 1e-3) / 1e9:.0f} GB/s), max rel. difference to FP64 {d32:.2e}; "
            f"hi + lo {msr:.3f} ms, difference {dr:.2e}")
        mixed = {
            "heff_tflops_hi_only": round(F_H / (ms32 * 1e-3) / 1e12, 4),
            "ms_per_step_hi_only": round(ms32, 4),
            "explicit_product_gbs_hi_only": round(b32 / (kms32[1] * 1e-3) / 1e9, 1) if kms32[1] > 0 else None,
            "kernel_ms_hi_only": [round(v, 4) for v in kms32],
            "max_rel_diff_hi_only_vs_fp64": d32,
            "ms_per_step_refined": round(msr, 4),
            "max_rel_diff_refined_vs_fp64": dr,
        }

    mx = mx_table(T, V, L)
    ctx.close()

    cpu = None
    if not args.no_cpu:
        cpu, err = cpu_baseline(L, D, site, seed, irr, T, V, xh, y_gpu, args.cpu_seconds)
        if err > 1e-10:
            log(f"WARNING: GPU result differs from the oracle on the sample: {err:.2e}")

    sweep = None
    if not args.no_sweep:
        wpath = os.path.join(ROOT, "chemps2_amd", "data", "precond_mlp.npy")
        params = np.load(wpath, allow_pickle=False) if (not args.no_learned and os.path.exists(wpath)) else None
        forms = [("fp64", chemps2_amd.FORM_FP64, None)]
        if mixed is not None:
            forms.append(("mixed", chemps2_amd.FORM_MIXED, None))
            if params is not None:
                forms.append(("mixed+learned", chemps2_amd.FORM_MIXED, params))
        sweep = run_sweep(L, D, seed, irr, mx, args.sweep_rtol, chemps2_amd.FORM_FP64)
        if mixed is not None:
            mixed["same_site_check"] = run_check(L, D, seed, irr, mx, args.sweep_rtol, args.check_rtol, forms)
            sm = run_sweep(L, D, seed, irr, mx, args.sweep_rtol, chemps2_amd.FORM_MIXED)
            sm["delta_E_vs_fp64"] = abs(sm["energy_last_site"] - sweep["energy_last_site"])
            sm["speedup_vs_fp64"] = round(sweep["seconds"] / sm["seconds"], 3)
            mixed["sweep"] = sm
            log(f"mixed sweep: {sm['seconds']:.1f}s vs FP64 {sweep['seconds']:.1f}s, |dE| = {sm['delta_E_vs_fp64']:.2e}")
            if params is not None:
                sl = run_sweep(L, D, seed, irr, mx, args.sweep_rtol, chemps2_amd.FORM_MIXED, precond=params)
                sl["delta_E_vs_fp64"] = abs(sl["energy_last_site"] - sweep["energy_last_site"])
                sl["speedup_vs_fp64"] = round(sweep["seconds"] / sl["seconds"], 3)
                sl["network"] = "chemps2_amd/data/precond_mlp.npy (trained on L=24 D=1000 seeds 0-7)"
                mixed["sweep_learned"] = sl
                log(f"mixed+learned sweep: {sl['seconds']:.1f}s, {sl['heff_applications']} products "
                    f"(diagonal: {sm['heff_applications']}), |dE| vs FP64 = {sl['delta_E_vs_fp64']:.2e}")

    out = {
        "metric": "Heff·x TFLOP/s + Davidson sweep time at L=32 D=2000; |ΔE| vs FP64 ≤1e-8 Ha",
        "value": round(F_H / (ms * 1e-3) / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "Heff x, random ab-initio D2h Hamiltonian, L=32 N=32 2S=0 D=2000, sites 15-16",
            "L": L, "D": D, "site": site, "V": Vlen, "F_H": F_H, "B_H": B_H,
            "explicit_doubles": nnz, "prepare_s": round(t_prep, 2),
            "kernel_ms": [round(v, 4) for v in kms],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "sweep": sweep,
        "mixed": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
