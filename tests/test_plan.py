"""CPU checks of the chemps2_amd work-list generator against the oracle: the generated
terms, combinations and diagonal terms, executed with numpy, reproduce the oracle's
Heff x and diag(Heff), and the generator's F_H equals the oracle's GEMM count."""
import numpy as np
import pytest

from chemps2_amd.problems import hubbard, random_abinitio
from plancheck import Plan


@pytest.mark.parametrize("index", [0, 1, 3, 5, 6])
def test_plan_matches_oracle_d2h(oracle_lib, index):
    irr, T, V = random_abinitio(8, 4, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(12, 4)
    o.sweep_left_to(index, 12, rtol=1e-4)
    o.site(index)
    p = Plan(o, index)
    x = np.random.default_rng(index).normal(size=o.V)
    y0, f0 = o.apply(x)
    assert np.abs(p.apply(x) - y0).max() <= 1e-12 * np.abs(y0).max()
    assert p.flops == f0
    assert np.abs(p.diag() - o.diag()).max() <= 1e-12 * np.abs(o.diag()).max()


@pytest.mark.parametrize("TwoS,N", [(1, 7), (2, 8)])
def test_plan_matches_oracle_spin(oracle_lib, TwoS, N):
    irr, T, V = random_abinitio(7, 8, group=4)
    o = oracle_lib.OracleDMRG(7, 4, irr, N, TwoS, 0, T, V)
    o.init(20, 1)
    o.sweep_left_to(2, 20, rtol=1e-4)
    o.site(2)
    p = Plan(o, 2)
    x = np.random.default_rng(1).normal(size=o.V)
    y0, f0 = o.apply(x)
    assert np.abs(p.apply(x) - y0).max() <= 1e-12 * np.abs(y0).max()
    assert p.flops == f0


def test_plan_hubbard_right_half(oracle_lib):
    # index >= L/2 exercises the right-side pair sums of 2a and the diagonal
    T, V = hubbard(10, U=3.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, V)
    o.init(16, 2)
    o.sweep_left_to(6, 16, rtol=1e-4)
    o.site(6)
    p = Plan(o, 6)
    x = np.random.default_rng(2).normal(size=o.V)
    y0, _ = o.apply(x)
    assert np.abs(p.apply(x) - y0).max() <= 1e-12 * np.abs(y0).max()
    assert np.abs(p.diag() - o.diag()).max() <= 1e-12 * np.abs(o.diag()).max()


def test_mx_table_matches_oracle(oracle_lib):
    from chemps2_amd.problems import mx_table

    irr, T, V = random_abinitio(6, 2, group=7)
    o = oracle_lib.OracleDMRG(6, 7, irr, 6, 0, 0, T, V)
    o.bookkeeper(10)
    assert np.abs(o.mx() - mx_table(T, V, 6)).max() < 1e-14


@pytest.mark.parametrize("L,N,TwoS,D,group", [(32, 32, 0, 2000, 7), (12, 12, 0, 1000, 0), (10, 9, 3, 50, 4)])
def test_auto_sectors_match_oracle_bookkeeper(oracle_lib, L, N, TwoS, D, group):
    import plancheck

    irr, T, V = random_abinitio(L, 0, group=group)
    o = oracle_lib.OracleDMRG(L, group, irr, N, TwoS, 0, T, V)
    o.bookkeeper(D)
    from chemps2_amd.problems import mx_table

    l = plancheck.lib()
    h = l.pc_create(L, group, irr.astype(np.int32), N, TwoS, 0, mx_table(T, V, N))
    l.pc_sectors_auto(h, D)
    for b in range(L + 1):
        n = l.pc_nsectors(h, b, None)
        out = np.zeros(4 * n, dtype=np.int32)
        l.pc_nsectors(h, b, out.ctypes.data)
        assert np.array_equal(out.reshape(n, 4), o.sectors(b)), b
    l.pc_destroy(h)


def test_synthetic_plan_matches_oracle(oracle_lib):
    # the oracle's synthetic operators through the product's work list
    irr, T, V = random_abinitio(12, 5, group=7)
    o = oracle_lib.OracleDMRG(12, 7, irr, 12, 0, 0, T, V)
    o.synthetic(60, 5, 11, 12)
    p = Plan(o, 5)
    x = np.random.default_rng(0).normal(size=o.V)
    y0, f0 = o.apply(x)
    assert np.abs(p.apply(x) - y0).max() <= 1e-12 * np.abs(y0).max()
    y1, f1, sec = o.apply_range(x, 0, 10)
    k2i = p.k2i
    assert np.array_equal(y1[: k2i[10]], y0[: k2i[10]]) and not y1[k2i[10]:].any()


@pytest.mark.parametrize("index,mr", [(0, True), (2, True), (4, True), (6, False), (4, False), (2, False)])
def test_update_plan_matches_oracle(oracle_lib, index, mr):
    # operator updates (DMRGoperators.cpp:243-907) through the library's work list
    from plancheck import UpdatePlan, oracle_tensors

    irr, T, V = random_abinitio(8, 4, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(14, 4)
    if not mr:
        o.sweep_left_to(max(index - 1, 0), 14, rtol=1e-4)
    new = UpdatePlan(o, index, mr).tensors()
    ref = {(kd, k, m): d for kd, k, m, meta, d in oracle_tensors(o, index)}
    assert set(new) == set(ref)
    for key, b in ref.items():
        if b.size:
            assert np.abs(new[key] - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), key


def test_reorder_d2h_matches_oracle(oracle_lib):
    from chemps2_amd.problems import permute, reorder_d2h

    irr, T, V = random_abinitio(9, 6, group=7)
    o = oracle_lib.OracleDMRG(9, 7, irr, 8, 0, 0, T, V, reorder=True)
    o.bookkeeper(20)
    irr2, T2, V2 = permute(irr, T, V, reorder_d2h(list(irr)))
    from chemps2_amd.problems import mx_table

    assert np.array_equal(o.irreps(), irr2)
    assert np.abs(o.mx() - mx_table(T2, V2, 8)).max() < 1e-14
