"""GPU parity at the benchmark configurations, against the oracle:

* L = 32, N = 32, random D2h, D = 2000 (the benchmark's Hamiltonian): the whole Heff x
  and diagonal at sites 15 (explicit route) and 8 (chained route) in the FP64 and mixed
  forms, with the reference's pseudo-random operator shapes (chm_tensor_fill_synthetic);
* L = 24, N = 24, random D2h, D = 1000: a real state (GPU initialisation and left half
  sweep), whose MPS the oracle loads and builds its own operators from; product,
  diagonal and the Davidson energy at site 11 in the FP64 and mixed forms;
* the 4 x 8 Hubbard cylinder (U = 8, N = 28, C1): site 15 at D = 1000 in full and at
  D = 3000 on a sample of output blocks (blocks of hundreds: the MFMA GEMM tiles);
* the learned preconditioner with the committed network on an unseen L = 32 seed:
  the energy of a two-site step equals the FP64 form's.
"""
import os

import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import hubbard_cylinder, mx_table, random_abinitio
from oracle.wrap import OracleDMRG

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def synthetic_pair(L, group, irr, N, TwoS, T, V, D, site, seeds=(1, 2)):
    o = OracleDMRG(L, group, irr, N, TwoS, 0, T, V)
    n = o.synthetic(D, site, seeds[0], seeds[1])
    c = chemps2_amd.Context(0)
    c.set_problem(L, group, irr, N, TwoS, 0, mx_table(T, V, N))
    c.set_sectors_auto(D)
    if site > 0:
        c.fill_synthetic(site - 1, 1, seeds[0])
    if site < L - 2:
        c.fill_synthetic(site + 1, 0, seeds[1])
    return o, c, n


@pytest.mark.parametrize("site", [15, 8])
def test_config_L32_D2000_heff_fp64_and_mixed(site):
    irr, T, V = random_abinitio(32, 0, group=7)
    print(f"case L=32 D=2000 site {site}", flush=True)
    o, c, n = synthetic_pair(32, 7, irr, 32, 0, T, V, 2000, site)
    try:
        assert c.prepare(site) == n
        x = np.random.default_rng(site).normal(size=n)
        y0, f0 = o.apply(x)
        F_H, _ = c.flops()
        assert F_H == f0
        e64 = rel(c.apply(x), y0)
        ed = rel(c.diag(), o.diag())
        c.set_heff_form(chemps2_amd.FORM_MIXED)
        c.prepare(site)
        eref = rel(c.apply(x), y0)
        ehi = rel(c.apply(x, f32=True), y0)
        print(f"  V={n} F_H={F_H:.3e}: fp64 {e64:.2e}, diag {ed:.2e}, mixed refined {eref:.2e}, hi only {ehi:.2e}",
              flush=True)
        assert e64 <= 1e-12 and ed <= 1e-12 and eref <= 1e-12
        assert ehi <= 1e-6  # FP32 storage of the explicit part: about 2^-24 of its elements
    finally:
        c.close()


def test_config_L24_D1000_real_state_heff_and_davidson():
    L, D, site = 24, 1000, 11
    irr, T, V = random_abinitio(L, 0, group=7)
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
        print("case L=24 D=1000: init and left half to site 12", flush=True)
        c.dmrg_init(D, 5)
        c.dmrg_sweep_range(False, L - 2, site + 1, D, 1e-5, 0.0, False)
        mps = [c.mps_download(s) for s in range(L)]
        o = OracleDMRG(L, 7, irr, L, 0, 0, T, V)
        n = o.load_mps(D, site, mps)
        c.join(site)
        guess = c.s_get()
        assert c.prepare(site) == n
        x = np.random.default_rng(3).normal(size=n)
        y0, f0 = o.apply(x)
        e64 = rel(c.apply(x), y0)
        ed = rel(c.diag(), o.diag())
        assert c.flops()[0] == f0
        eo, _, nmo = o.davidson(1e-8)
        eg, nmg = c.davidson(guess.copy(), 1e-8)
        c.set_heff_form(chemps2_amd.FORM_MIXED)
        c.prepare(site)
        em, nmm = c.davidson(guess.copy(), 1e-8)
        print(f"  V={n}: product {e64:.2e}, diag {ed:.2e}; Davidson oracle {eo!r} ({nmo}), gpu {eg!r} ({nmg}), "
              f"mixed {em!r} ({nmm})", flush=True)
        assert e64 <= 1e-12 and ed <= 1e-12
        assert abs(eg - eo) <= 1e-9 and abs(em - eo) <= 1e-9
    finally:
        c.close()


def test_config_cylinder_D1000_heff():
    T, V, _ = hubbard_cylinder(8, 4, U=8.0)
    irr = np.zeros(32, dtype=np.int32)
    print("case cylinder D=1000 site 15", flush=True)
    o, c, n = synthetic_pair(32, 0, irr, 28, 0, T, V, 1000, 15)
    try:
        assert c.prepare(15) == n
        st = c.plan_stats()
        x = np.random.default_rng(7).normal(size=n)
        y0, f0 = o.apply(x)
        e64 = rel(c.apply(x), y0)
        ed = rel(c.diag(), o.diag())
        print(f"  V={n} F_H={f0:.3e} tiles={st['tiles1']}+{st['tiles2']}: product {e64:.2e}, diag {ed:.2e}", flush=True)
        assert c.flops()[0] == f0
        assert e64 <= 1e-12 and ed <= 1e-12
    finally:
        c.close()
    # the same product with every chained term on the MFMA GEMM tiles (CHM_CHAIN_KERNEL=0,
    # read when a context is created): the tile path at this configuration
    os.environ["CHM_CHAIN_KERNEL"] = "0"
    try:
        c = chemps2_amd.Context(0)
    finally:
        del os.environ["CHM_CHAIN_KERNEL"]
    try:
        c.set_problem(32, 0, irr, 28, 0, 0, mx_table(T, V, 28))
        c.set_sectors_auto(1000)
        c.fill_synthetic(14, 1, 1)
        c.fill_synthetic(16, 0, 2)
        assert c.prepare(15) == n
        st = c.plan_stats()
        et = rel(c.apply(x), y0)
        print(f"  chained terms on the GEMM tiles: tiles={st['tiles1']}+{st['tiles2']}, product {et:.2e}", flush=True)
        assert st["tiles1"] + st["tiles2"] > 0
        assert et <= 1e-12
    finally:
        c.close()


def test_config_cylinder_D3000_heff_sampled():
    T, V, _ = hubbard_cylinder(8, 4, U=8.0)
    irr = np.zeros(32, dtype=np.int32)
    print("case cylinder D=3000 site 15 (sampled output blocks)", flush=True)
    o, c, n = synthetic_pair(32, 0, irr, 28, 0, T, V, 3000, 15)
    try:
        assert c.prepare(15) == n
        x = np.random.default_rng(8).normal(size=n)
        y1 = c.apply(x)
        lab, k2i = o.kappa()
        nk = len(lab)
        # output blocks spread over the whole vector, as many as about 30 s of oracle time
        worst, secs, done = 0.0, 0.0, 0
        for k in np.linspace(0, nk - 1, 48).astype(int):
            y, _, s = o.apply_range(x, int(k), int(k) + 1)
            a, b = k2i[k], k2i[k + 1]
            if b > a:
                worst = max(worst, float(np.abs(y[a:b] - y1[a:b]).max() / max(np.abs(y[a:b]).max(), 1e-300)))
            secs += s
            done += 1
            if secs > 30.0:
                break
        print(f"  V={n}: {done} output blocks, worst {worst:.2e}, oracle {secs:.1f} s", flush=True)
        assert done >= 8 and worst <= 1e-12
    finally:
        c.close()


def test_config_learned_preconditioner_unseen_L32():
    wpath = os.path.join(ROOT, "chemps2_amd", "data", "precond_mlp.npy")
    params = np.load(wpath, allow_pickle=False)
    L, D, site = 32, 2000, 15
    irr, T, V = random_abinitio(L, 1000, group=7)  # not among the training seeds
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
        print("case learned preconditioner, L=32 seed 1000 D=2000", flush=True)
        c.dmrg_init(D, 3)
        c.dmrg_sweep_range(False, L - 2, site + 1, D, 1e-5, 0.0, False)
        c.join(site)
        guess = c.s_get()
        out = {}
        for name, form, prm in (("fp64", chemps2_amd.FORM_FP64, None),
                                ("mixed+learned", chemps2_amd.FORM_MIXED, params)):
            c.set_heff_form(form)
            c.precond_mode(chemps2_amd.PRECOND_DIAGONAL)
            if prm is not None:
                c.mlp_set(prm)
                c.precond_mode(chemps2_amd.PRECOND_LEARNED)
            c.prepare(site)
            out[name] = c.davidson(guess.copy(), 1e-8)
        print(f"  {out}", flush=True)
        assert abs(out["fp64"][0] - out["mixed+learned"][0]) <= 1e-8
    finally:
        c.close()
