"""chm_heff_apply_noalloc (the Heff product with its partial sums in the library's static
device scratch, no device allocation) against chm_heff_apply on the same prepared site:
the same kernels in the same order, so the results are identical. Two calls in a row,
the second on a smaller site, through the C ABI."""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import mx_table, random_abinitio
from oracle.wrap import OracleDMRG

pytestmark = pytest.mark.gpu


def _site(L, D, site, seed):
    irr, T, V = random_abinitio(L, seed, group=7)
    o = OracleDMRG(L, 7, irr, L, 0, 0, T, V)
    n = o.synthetic(D, site, 1, 2)
    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, L, 0, 0, mx_table(T, V, L))
    c.set_sectors_auto(D)
    c.fill_synthetic(site - 1, 1, 1)
    c.fill_synthetic(site + 1, 0, 2)
    assert c.prepare(site) == n
    return c, n


def test_noalloc_product_equals_apply():
    import torch

    sizes = []
    for L, D, site, seed in ((16, 400, 7, 3), (10, 60, 4, 5)):  # the second input is smaller
        print(f"case no-alloc product L={L} D={D} site {site}", flush=True)
        c, n = _site(L, D, site, seed)
        try:
            x = np.random.default_rng(seed).normal(size=n)
            y_ref = c.apply(x)
            xd = torch.tensor(x, dtype=torch.float64, device="cuda:0")
            yd = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
            c.apply_noalloc(xd, yd)
            torch.cuda.synchronize()
            y = yd.cpu().numpy()
            print(f"  V={n}: max |diff| {np.abs(y - y_ref).max():.3e}", flush=True)
            assert np.array_equal(y, y_ref)
            sizes.append(n)
        finally:
            c.close()
    assert sizes[1] < sizes[0]
