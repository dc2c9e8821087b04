"""GPU parity of the renormalised-operator updates (DMRGoperators.cpp:243-907) against
the oracle, through the C ABI: every tensor of the new slot within 1e-12 relative."""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import random_abinitio
from plancheck import oracle_tensors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = chemps2_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("index,mr", [(0, True), (3, True), (5, True), (6, False), (3, False), (1, False)])
def test_update_operators(ctx, oracle_lib, index, mr):
    L = 8
    irr, T, V = random_abinitio(L, 4, group=7)
    o = oracle_lib.OracleDMRG(L, 7, irr, 8, 0, 0, T, V)
    o.init(14, 4)
    if not mr:
        o.sweep_left_to(max(index - 1, 0), 14, rtol=1e-4)
    print(f"update index={index} moving_right={mr}", flush=True)
    ctx.set_problem(L, 7, o.irreps(), 8, 0, 0, o.mx())
    for b in range(L + 1):
        ctx.set_sectors(b, o.sectors(b))
    prev = index - 1 if mr else index + 1
    if 0 <= prev <= L - 2:
        for kind, k, m, meta, data in oracle_tensors(o, prev):
            ctx.upload(prev, mr, kind, k, m, int(meta[1]), int(meta[2]), int(meta[3]), data)
    tsite = index if mr else index + 1
    ctx.mps_upload(tsite, o.mps(tsite))
    ctx.update_operators(index, mr)
    n = 0
    for kind, k, m, meta, ref in oracle_tensors(o, index):
        got = ctx.tensor_download(index, kind, k, m, ref.size)
        if ref.size:
            assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), (kind, k, m)
        n += 1
    assert n > 50


@pytest.mark.parametrize("index,mr", [(3, True), (3, False), (0, True), (6, False)])
@pytest.mark.parametrize("ops_kernel,tiles_mfma", [("1", "2"), ("0", "1"), ("0", "0")])
def test_update_operators_other_paths(oracle_lib, monkeypatch, index, mr, ops_kernel, tiles_mfma):
    # small products through k_ops and the rest on the default tiles (CHM_OPS_KERNEL=1),
    # and every product through the older GEMM tiles, on the matrix cores or with vector
    # FMA, agree as well
    monkeypatch.setenv("CHM_OPS_KERNEL", ops_kernel)
    monkeypatch.setenv("CHM_TILES_MFMA", tiles_mfma)
    c = chemps2_amd.Context(0)
    try:
        test_update_operators(c, oracle_lib, index, mr)
    finally:
        c.close()
