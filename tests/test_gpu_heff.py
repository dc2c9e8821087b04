"""GPU parity of the chemps2_amd Heff path against the oracle, through the C ABI.

Bars: Heff x and diag(Heff) within 1e-12 relative (FP64, same association order, only
summation order differs); Davidson eigenvalue within 1e-9 of the oracle's Davidson at
rtol 1e-10, and the GPU eigenvector's residual under the oracle's Heff below 1e-7.
"""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import hubbard, random_abinitio
from helpers import load_site, symm_to_prog

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = chemps2_amd.Context(0)
    yield c
    c.close()


def d2h_case(oracle_lib, index, D=12, L=8, N=8, seed=4):
    irr, T, V = random_abinitio(L, seed, group=7)
    o = oracle_lib.OracleDMRG(L, 7, irr, N, 0, 0, T, V)
    o.init(D, seed)
    o.sweep_left_to(index, D, rtol=1e-4)
    o.site(index)
    return o


def check_apply_diag(ctx, o, index, seed=0):
    print(f"case L={o.L} site={index} V={o.V}", flush=True)
    V = load_site(ctx, o, index)
    assert V == o.V
    x = np.random.default_rng(seed).normal(size=V)
    y0, f0 = o.apply(x)
    y1 = ctx.apply(x)
    assert np.abs(y1 - y0).max() <= 1e-12 * np.abs(y0).max()
    d0 = o.diag()
    assert np.abs(ctx.diag() - d0).max() <= 1e-12 * np.abs(d0).max()
    F, B = ctx.flops()
    assert F == f0
    assert B > 16 * V


@pytest.mark.parametrize("index", [0, 1, 3, 5, 6])
def test_apply_and_diag_d2h(ctx, oracle_lib, index):
    check_apply_diag(ctx, d2h_case(oracle_lib, index), index)


def test_apply_large_blocks_c1(ctx, oracle_lib):
    # C1 Hubbard L=10 at D=1000: left dimensions up to 75, wider than one 64 x 64 tile
    T, V = hubbard(10, U=4.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, V)
    o.init(1000, 3)
    o.sweep_left_to(5, 1000, rtol=1e-3)
    o.site(5)
    assert o.sectors(5)[:, 3].max() > 64
    check_apply_diag(ctx, o, 5, seed=1)


def test_apply_spin_and_c2v(ctx, oracle_lib):
    irr, T, V = random_abinitio(7, 8, group=5)
    o = oracle_lib.OracleDMRG(7, 5, irr, 7, 3, 0, T, V)
    o.init(30, 1)
    o.sweep_left_to(2, 30, rtol=1e-4)
    o.site(2)
    check_apply_diag(ctx, o, 2, seed=3)


@pytest.mark.parametrize("index", [0, 3, 6])
def test_davidson(ctx, oracle_lib, index):
    o = d2h_case(oracle_lib, index, D=16)
    load_site(ctx, o, index)
    e0, x0, nm0 = o.davidson(rtol=1e-10)
    guess = symm_to_prog(o, o.site_x())
    e1, nm1 = ctx.davidson(guess, 1e-10)
    print(f"site {index}: oracle {e0!r} ({nm0} matvecs) gpu {e1!r} ({nm1} matvecs)", flush=True)
    assert abs(e1 - e0) <= 1e-9
    # residual of the GPU eigenvector (program convention -> symmetric) under the oracle
    lab, k2i = o.kappa()
    xs = guess.copy()
    for k in range(len(lab)):
        xs[k2i[k]:k2i[k + 1]] *= np.sqrt(lab[k, 7] + 1.0)
    hx, _ = o.apply(xs)
    assert np.linalg.norm(hx - e1 * xs) <= 1e-7
    assert abs(np.linalg.norm(xs) - 1.0) <= 1e-10


@pytest.mark.parametrize("explicit_max,chain_kernel,tiles_mfma,w_cap", [
    ("0", "1", "1", None), ("0", "0", "1", None), ("0", "0", "0", None), ("0", "0", "1", "20000"),
    ("64", "1", "1", None), ("100000000", "1", "1", None)])
def test_explicit_threshold_paths(oracle_lib, explicit_max, chain_kernel, tiles_mfma, w_cap, monkeypatch):
    # all chained (output-stationary kernel for blocks up to 64, GEMM tiles above; or
    # GEMM tiles only, on the matrix cores or with vector FMA, in one batch or in many
    # batches sharing a small intermediate buffer), a mix, and everything explicit
    monkeypatch.setenv("CHM_EXPLICIT_MAX", explicit_max)
    monkeypatch.setenv("CHM_CHAIN_KERNEL", chain_kernel)
    monkeypatch.setenv("CHM_TILES_MFMA", tiles_mfma)
    if w_cap:
        monkeypatch.setenv("CHM_TILE_W_CAP", w_cap)
    c = chemps2_amd.Context(0)
    T, V = hubbard(10, U=4.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, V)
    o.init(1000, 3)
    o.sweep_left_to(5, 1000, rtol=1e-3)
    o.site(5)
    check_apply_diag(c, o, 5, seed=2)
    st = c.plan_stats()
    print(explicit_max, st, flush=True)
    if explicit_max == "0":
        assert st["explicit_pairs"] == 0
    if explicit_max == "100000000":
        assert st["chain_terms"] == 0
    if chain_kernel == "0":
        assert st["chain_kernel_terms"] == 0
    c.close()


def test_gemm_tile_routes(oracle_lib, monkeypatch):
    # every chained term through the GEMM tiles (explicit limit 0, no chained kernels):
    # the default kernels (streamed phase-1 workgroups, wavefront tiles of at most
    # 32 x 32, grouped phase-2 tiles), every tile on 32 x 32 wavefronts, phase-2 output
    # blocks cut into segments of a few k chunks (partial sums and their reduction),
    # and the older phase-2 kernels, each against the oracle
    T, V = hubbard(10, U=4.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, V)
    o.init(1000, 3)
    o.sweep_left_to(5, 1000, rtol=1e-3)
    o.site(5)
    monkeypatch.setenv("CHM_EXPLICIT_MAX", "0")
    monkeypatch.setenv("CHM_CHAIN_KERNEL", "0")
    monkeypatch.setenv("CHM_FCHAIN", "0")
    tiles2 = {}
    for name, env in [("default", {}), ("wave", {"CHM_WAVE_TILES": "1"}),
                      ("segments", {"CHM_TILE_SEGMENTS": "100000000"}),
                      ("wave+segments", {"CHM_WAVE_TILES": "1", "CHM_TILE_SEGMENTS": "100000000"}),
                      ("segments, k_tiles_mfma", {"CHM_TILE_SEGMENTS": "100000000", "CHM_TILES_MFMA": "1"}),
                      ("segments, k_tiles", {"CHM_TILE_SEGMENTS": "100000000", "CHM_TILES_MFMA": "0"})]:
        for k in ("CHM_WAVE_TILES", "CHM_TILE_SEGMENTS", "CHM_TILE_SEG_MIN", "CHM_TILES_MFMA"):
            monkeypatch.delenv(k, raising=False)
        if "CHM_TILE_SEGMENTS" in env:
            env = dict(env, CHM_TILE_SEG_MIN="2")
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        c = chemps2_amd.Context(0)
        try:
            print(f"tile route {name}", flush=True)
            check_apply_diag(c, o, 5, seed=3)
            st = c.plan_stats()
            assert st["explicit_pairs"] == 0 and st["chain_kernel_terms"] == 0 and st["tiles1"] > 0
            tiles2[name] = st["tiles2"]
        finally:
            c.close()
    # segments multiply the phase-2 tiles; 32 x 32 tiles cut the larger blocks in four
    assert tiles2["segments"] > tiles2["default"]
    assert tiles2["wave"] >= tiles2["default"]


@pytest.mark.parametrize("site", [0, 3, 6])
@pytest.mark.parametrize("mfma", ["1", "0"])
def test_chain_kernel_d2h(oracle_lib, site, mfma, monkeypatch):
    # every term through the output-stationary chained kernels (small D2h blocks: all
    # five term modes, transposed and plain operator blocks), on the matrix cores
    # (k_chain_mfma) or in LDS tiles (k_chain)
    monkeypatch.setenv("CHM_EXPLICIT_MAX", "0")
    monkeypatch.setenv("CHM_CHAIN_KERNEL", "1")
    monkeypatch.setenv("CHM_CHAIN_MFMA", mfma)
    c = chemps2_amd.Context(0)
    irr, T, V = random_abinitio(8, 11, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(24, 2)
    o.sweep_left_to(site, 24, rtol=1e-4)
    o.site(site)
    check_apply_diag(c, o, site, seed=site + 5)
    st = c.plan_stats()
    assert st["explicit_pairs"] == 0 and st["chain_kernel_terms"] == st["terms"]
    c.close()


def test_synthetic_matches_oracle(ctx, oracle_lib):
    # the library's pseudo-random operators equal the oracle's synthetic ones
    from chemps2_amd.problems import mx_table

    irr, T, V = random_abinitio(12, 5, group=7)
    o = oracle_lib.OracleDMRG(12, 7, irr, 12, 0, 0, T, V)
    n = o.synthetic(60, 5, 11, 12)
    ctx.set_problem(12, 7, irr, 12, 0, 0, mx_table(T, V, 12))
    ctx.set_sectors_auto(60)
    ctx.fill_synthetic(4, 1, 11)
    ctx.fill_synthetic(6, 0, 12)
    assert ctx.prepare(5) == n
    x = np.random.default_rng(7).normal(size=n)
    y0, f0 = o.apply(x)
    assert np.abs(ctx.apply(x) - y0).max() <= 1e-12 * np.abs(y0).max()
    assert ctx.flops()[0] == f0


def test_synthetic_fill_and_auto_sectors(ctx):
    # library-derived dimensions and shapes: prepare succeeds and Heff is symmetric
    irr, T, V = random_abinitio(10, 3, group=7)
    from chemps2_amd.problems import mx_table
    ctx.set_problem(10, 7, irr, 10, 0, 0, mx_table(T, V, 10))
    ctx.set_sectors_auto(40)
    ctx.fill_synthetic(3, 1, 11)
    ctx.fill_synthetic(5, 0, 12)
    Vlen = ctx.prepare(4)
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=Vlen), rng.normal(size=Vlen)
    ha, hb = ctx.apply(a), ctx.apply(b)
    assert np.isfinite(ha).all() and np.abs(ha).max() > 0
