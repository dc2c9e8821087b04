"""GPU tests of the sweep layer: Join and Split against the oracle, and the whole DMRG
on the device against FCI and the reference's known answers
(tests/test4.cpp.in:103, tests/test5.cpp.in:83)."""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import hubbard, molpro_to_xor, mx_table, permute, random_abinitio, read_fcidump, reorder_d2h
from fci import fci_energy
from helpers import symm_to_prog

pytestmark = pytest.mark.gpu
GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = chemps2_amd.Context(0)
    yield c
    c.close()


def load_state(ctx, o, D_fci):
    """Problem, FCI dimensions (through the library's own bookkeeper) and then the
    oracle's current dimensions and MPS."""
    L = o.L
    ctx.set_problem(L, int(o.group), o.irreps(), o.N, o.TwoS, o.Irrep, o.mx())
    ctx.set_sectors_auto(D_fci)
    for b in range(L + 1):
        ctx.set_sectors(b, o.sectors(b))
    for s in range(L):
        ctx.mps_upload(s, o.mps(s))


@pytest.mark.parametrize("index", [0, 3, 6])
def test_join_matches_oracle(ctx, oracle_lib, index):
    print(f"case join index {index}", flush=True)
    irr, T, V = random_abinitio(8, 4, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(16, 4)
    o.sweep_left_to(index, 16, rtol=1e-4)
    o.site(index)
    load_state(ctx, o, 16)
    n = ctx.join(index)
    assert n == o.V
    ref = symm_to_prog(o, o.site_x())
    assert np.abs(ctx.s_get() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("index,mr,D", [(2, True, 6), (4, False, 6), (3, True, 1000)])
def test_split_matches_oracle(ctx, oracle_lib, index, mr, D):
    print(f"case split index {index} mr {mr} D {D}", flush=True)
    # same two-site data split by both: equal discarded weight and dimensions, and the
    # re-joined tensor (independent of SVD sign choices) agrees
    irr, T, V = random_abinitio(8, 4, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(16, 4)
    o.sweep_left_to(index, 16, rtol=1e-4)
    o.site(index)
    load_state(ctx, o, 16)
    ctx.join(index)
    s_prog = symm_to_prog(o, o.davidson(rtol=1e-8)[1])
    ctx.s_set(s_prog)
    dw_gpu = ctx.split(index, D, mr, True)
    dw_orc = o.split_test(index, D, mr, True, s_prog)
    assert abs(dw_gpu - dw_orc) <= 1e-10 * max(1.0, dw_orc)
    assert np.array_equal(ctx.get_sectors(index + 1), o.sectors(index + 1))
    ctx.join(index)
    o.site(index)
    ref = symm_to_prog(o, o.site_x())
    assert np.abs(ctx.s_get() - ref).max() <= 1e-10 * np.abs(ref).max()


def run_dmrg(L, group, irr, N, TwoS, Irrep, T, V, scheme, D0, econst=0.0, seed=3):
    c = chemps2_amd.Context(0)
    c.set_problem(L, group, irr, N, TwoS, Irrep, mx_table(T, V, N), econst)
    c.dmrg_init(D0, seed)
    e = c.dmrg_solve(scheme)
    stats = c.dmrg_stats()
    c.close()
    return e, stats


def test_split_jacobi_fallback(monkeypatch):
    # Split's Jacobi kernel limited to one sweep: its sectors report non-convergence and
    # are decomposed again by rocSOLVER's dgesvd; the DMRG still reaches the energy of the
    # default run (and FCI)
    irr, T, V = random_abinitio(8, 5, group=7)
    for target in range(8):
        try:
            ef = fci_energy(8, irr, 8, 0, target, T, V)
            break
        except ValueError:
            continue
    print(f"case Split fallback L=8 N=8 irrep {target}", flush=True)
    scheme = [(60, 1e-11, 6, 0.0, 1e-10)]
    e0, _ = run_dmrg(8, 7, irr, 8, 0, target, T, V, scheme, 60)
    monkeypatch.setenv("CHM_SVD_MAX_SWEEPS", "1")
    e1, _ = run_dmrg(8, 7, irr, 8, 0, target, T, V, scheme, 60)
    print(f"  default {e0!r} fallback {e1!r} fci {ef!r}", flush=True)
    assert abs(e1 - e0) < 1e-9 and abs(e1 - ef) < 1e-8


@pytest.mark.parametrize("L,N,TwoS,seed", [(6, 6, 0, 1), (7, 7, 1, 2), (8, 6, 2, 3), (6, 7, 3, 4)])
def test_dmrg_equals_fci(L, N, TwoS, seed):
    irr, T, V = random_abinitio(L, seed, group=7)
    for target in range(8):
        try:
            ef = fci_energy(L, irr, N, TwoS, target, T, V)
            break
        except ValueError:
            continue
    print(f"case L={L} N={N} 2S={TwoS} seed={seed}", flush=True)
    e, stats = run_dmrg(L, 7, irr, N, TwoS, target, T, V, [(200, 1e-11, 10, 0.0, 1e-10)], 200)
    print(f"  gpu {e!r} fci {ef!r} stats {stats}", flush=True)
    assert abs(e - ef) < 1e-9


def test_dmrg_all_irreps_c2v():
    irr, T, V = random_abinitio(5, 9, group=5)
    for target in range(4):
        ef = fci_energy(5, irr, 4, 0, target, T, V)
        print(f"case C2v target {target}", flush=True)
        e, _ = run_dmrg(5, 5, irr, 4, 0, target, T, V, [(100, 1e-11, 10, 0.0, 1e-10)], 100)
        assert abs(e - ef) < 1e-9


def test_reference_test4_hubbard_on_gpu():
    # tests/test4.cpp.in:36-103: L=10 Hubbard, N=9, 2S=5, instructions (30, 1e-10, 3, 0.1)
    # and (1000, 1e-10, 10, 0.0); DMRG within 1e-8 of FCI
    T, V = hubbard(10, t=1.0, U=2.0)
    print("case test4 Hubbard L=10", flush=True)
    e, stats = run_dmrg(10, 0, [0] * 10, 9, 5, 0, T, V, [(30, 1e-10, 3, 0.1, 1e-5), (1000, 1e-10, 10, 0.0, 1e-5)], 30)
    ef = fci_energy(10, [0] * 10, 9, 5, 0, T, V)
    print(f"test4: gpu {e!r} fci {ef!r}", flush=True)
    assert abs(e - ef) < 1e-8


def test_reference_test5_n2_on_gpu():
    # tests/test5.cpp.in:39-83: N2 STO-3G, D2h reordered, D=1000, E0 = -107.648250974014
    d = read_fcidump(GOLDEN + "/N2.STO3G.FCIDUMP")
    irr = [molpro_to_xor(7, m) for m in d["molpro_irreps"]]
    print("case test5 N2 D=1000", flush=True)
    f2 = reorder_d2h(irr)
    irr2, T2, V2 = permute(irr, d["T"], d["V"], f2)
    e, stats = run_dmrg(d["L"], 7, irr2, d["N"], 0, 0, T2, V2, [(1000, 1e-12, 100, 0.0, 1e-5)], 1000, d["econst"])
    print(f"test5: gpu {e!r}", flush=True)
    assert abs(e + 107.648250974014) < 1e-8


def test_reference_test3_ch4_on_gpu():
    # tests/test3.cpp.in:40-115: CH4 STO-3G in C2v, DMRG within 1e-8 of FCI
    d = read_fcidump(GOLDEN + "/CH4.STO3G.FCIDUMP")
    irr = [molpro_to_xor(5, m) for m in d["molpro_irreps"]]
    L, N = d["L"], d["N"]
    print("case test3 CH4", flush=True)
    e, _ = run_dmrg(L, 5, irr, N, 0, 0, d["T"], d["V"], [(30, 1e-10, 3, 0.1, 1e-5), (1000, 1e-10, 10, 0.0, 1e-5)], 30,
                    d["econst"])
    ef = fci_energy(L, irr, N, 0, 0, d["T"], d["V"], d["econst"])
    assert abs(e - ef) < 1e-8


@pytest.mark.parametrize("mr,D", [(True, 10), (False, 1000)])
def test_split_svd_paths_agree(oracle_lib, monkeypatch, mr, D):
    # batched Jacobi kernel (default), rocSOLVER's dgesvd (CHM_SVD_JACOBI_MAXN=0) and its
    # gesvdj (CHM_SVD_GESVDJ=1) give the same truncation, discarded weight and re-joined
    # two-site object
    print(f"case svd paths mr {mr} D {D}", flush=True)
    irr, T, V = random_abinitio(8, 5, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(20, 4)
    o.sweep_left_to(3, 20, rtol=1e-4)
    o.site(3)
    s_prog = symm_to_prog(o, o.davidson(rtol=1e-8)[1])
    out = []
    for maxn, gesvdj in (("64", "0"), ("0", "0"), ("0", "1")):
        monkeypatch.setenv("CHM_SVD_JACOBI_MAXN", maxn)
        monkeypatch.setenv("CHM_SVD_GESVDJ", gesvdj)
        c = chemps2_amd.Context(0)
        load_state(c, o, 20)
        c.join(3)
        c.s_set(s_prog)
        dw = c.split(3, D, mr, True)
        sec = c.get_sectors(4)
        c.join(3)
        out.append((dw, sec, c.s_get()))
        c.close()
    dw0, sec0, s0 = out[0]
    for dw1, sec1, s1 in out[1:]:  # rocSOLVER's dgesvd, then its gesvdj
        assert np.array_equal(sec0, sec1)
        assert abs(dw0 - dw1) <= 1e-12 * max(1.0, dw1)
        assert np.abs(s0 - s1).max() <= 1e-11 * np.abs(s1).max()


def test_hubbard_l12_against_fci():
    # 1D Hubbard, L = 12, N = 12, 2S = 0, C1, t = 1 (open chain), U = 4; two instructions
    # at D = 200 with energy convergence 1e-8 (first with noise 0.05 and Davidson
    # tolerance 1e-5, second without noise at 1e-8); FCI over all 853,776 S_z = 0
    # determinants (tests/fci.py hubbard_fci_energy)
    from fci import hubbard_fci_energy

    print("case Hubbard L=12 N=12 U=4 D=200", flush=True)
    T, V = hubbard(12, t=1.0, U=4.0)
    ef = hubbard_fci_energy(12, 12, t=1.0, U=4.0)
    e, stats = run_dmrg(12, 0, [0] * 12, 12, 0, 0, T, V,
                        [(200, 1e-8, 10, 0.05, 1e-5), (200, 1e-8, 10, 0.0, 1e-8)], 200)
    print(f"  gpu {e!r} fci {ef!r} |dE| {abs(e - ef):.2e} stats {stats}", flush=True)
    assert abs(e - ef) < 1e-8


def test_sweep_range_continues_the_sweep():
    # a left half run as two ranges (sites 6..4, then 3..1) equals the single half sweep
    print("case sweep ranges L=8", flush=True)
    irr, T, V = random_abinitio(8, 21, group=7)
    mx = mx_table(T, V, 8)
    energies = []
    for split in (False, True):
        c = chemps2_amd.Context(0)
        c.set_problem(8, 7, irr, 8, 0, 0, mx)
        c.dmrg_init(64, 5)
        if split:
            c.dmrg_sweep_range(False, 6, 4, 64, 1e-8, 0.0, False)
            e_left = c.dmrg_sweep_range(False, 3, 1, 64, 1e-8, 0.0, False)
        else:
            e_left = c.dmrg_sweep(False, 64, 1e-8, 0.0, False)
        e_right = c.dmrg_sweep(True, 64, 1e-8, 0.0, True)
        with pytest.raises(chemps2_amd.ChmError):
            c.dmrg_sweep_range(False, 3, 5, 64, 1e-8, 0.0, False)  # out of order
        c.close()
        energies.append((e_left, e_right))
    assert abs(energies[0][0] - energies[1][0]) < 1e-12 and abs(energies[0][1] - energies[1][1]) < 1e-12
