"""The fused MFMA chained kernel (k_fchain: chained products of blocks up to 80 on the
matrix cores, one workgroup per work item, row strips per wavefront, the other
association order as the transposed block) against the oracle's Heff (HeffDiagrams2-5,
Heff.cpp:43-248): every chained block it may take is sent to it (CHM_FCHAIN_MIN=0,
CHM_FCHAIN_MAXD=80), at an oracle state of a D2h Hamiltonian (interior and edge sites)
and at the 4 x 8 Hubbard cylinder at D = 1000 (blocks of 65-80 with their shared fifth
strip, larger ones on the GEMM tiles). Products to 1e-12 relative; the FP32 variant of
the mixed form within its error bound, refined to 1e-12."""
import os

import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import hubbard_cylinder, mx_table, random_abinitio
from helpers import load_site
from oracle.wrap import OracleDMRG

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _ctx(env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return chemps2_amd.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("site", [1, 5, 9])
def test_fchain_oracle_state(site):
    irr, T, V = random_abinitio(12, 9, group=7)
    o = OracleDMRG(12, 7, irr, 12, 0, 0, T, V)
    o.init(150, 2)
    o.sweep_left_to(site, 150, rtol=1e-4)
    n = o.site(site)
    x = np.random.default_rng(site).normal(size=n)
    y0, f0 = o.apply(x)
    print(f"case fused chain kernel, L=12 D=150 site {site}, V={n}", flush=True)
    for env in ({"CHM_CHAIN_KERNEL": "1", "CHM_FCHAIN_MIN": "0", "CHM_FCHAIN_MAXD": "80", "CHM_EXPLICIT_MAX": "0"}, {"CHM_CHAIN_KERNEL": "1", "CHM_FCHAIN_MIN": "0", "CHM_FCHAIN_MAXD": "80"}):
        c = _ctx(env)
        try:
            assert load_site(c, o, site) == n
            st = c.plan_stats()
            e = rel(c.apply(x), y0)
            print(f"  {env}: chain-kernel terms {st['chain_kernel_terms']}, product {e:.2e}", flush=True)
            assert c.flops()[0] == f0
            assert e <= 1e-12
        finally:
            c.close()


def test_fchain_cylinder_D1000():
    T, V, _ = hubbard_cylinder(8, 4, U=8.0)
    irr = np.zeros(32, dtype=np.int32)
    o = OracleDMRG(32, 0, irr, 28, 0, 0, T, V)
    n = o.synthetic(1000, 15, 1, 2)
    x = np.random.default_rng(3).normal(size=n)
    y0, _ = o.apply(x)
    print(f"case fused chain kernel, cylinder D=1000 site 15, V={n}", flush=True)
    c = _ctx({"CHM_CHAIN_KERNEL": "1", "CHM_FCHAIN_MIN": "0", "CHM_FCHAIN_MAXD": "80"})
    try:
        c.set_problem(32, 0, irr, 28, 0, 0, mx_table(T, V, 28))
        c.set_sectors_auto(1000)
        c.fill_synthetic(14, 1, 1)
        c.fill_synthetic(16, 0, 2)
        assert c.prepare(15) == n
        e = rel(c.apply(x), y0)
        print(f"  product {e:.2e}, plan {c.plan_stats()}", flush=True)
        assert e <= 1e-12
    finally:
        c.close()


@pytest.mark.parametrize("site", [5, 9])
def test_fchain_fp32_mixed_form(site):
    """The mixed form's low-precision product (k_fchain32 on v_mfma_f32_16x16x4f32 for the
    fused items, hi floats for the explicit part) is within the FP32 error bound of the
    oracle, and the refined product (FP64 kernels, hi + lo) equals it to 1e-12."""
    irr, T, V = random_abinitio(12, 9, group=7)
    o = OracleDMRG(12, 7, irr, 12, 0, 0, T, V)
    o.init(150, 2)
    o.sweep_left_to(site, 150, rtol=1e-4)
    n = o.site(site)
    x = np.random.default_rng(site).normal(size=n)
    y0, _ = o.apply(x)
    print(f"case FP32 fused chain kernel, mixed form, L=12 D=150 site {site}, V={n}", flush=True)
    for env in ({"CHM_CHAIN_KERNEL": "1", "CHM_FCHAIN_MIN": "0", "CHM_FCHAIN_MAXD": "80", "CHM_EXPLICIT_MAX": "0"}, {"CHM_CHAIN_KERNEL": "1", "CHM_FCHAIN_MIN": "0", "CHM_FCHAIN_MAXD": "80"}):
        c = _ctx(env)
        try:
            c.set_heff_form(chemps2_amd.FORM_MIXED)
            assert load_site(c, o, site) == n
            e_low = rel(c.apply(x, f32=True), y0)
            e_ref = rel(c.apply(x), y0)
            print(f"  {env}: low precision {e_low:.2e}, refined {e_ref:.2e}", flush=True)
            assert 1e-12 < e_low <= 1e-5  # FP32 products were used, within their bound
            assert e_ref <= 1e-12
        finally:
            c.close()
