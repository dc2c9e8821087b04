"""The mixed Heff form against the FP64 form, the oracle and FCI.

In the mixed form the explicit sector matrix is stored as float pairs (hi = fl32(v),
lo = fl32(v - hi)); the Davidson iterates with hi-only products and refines with hi + lo
products once the residual is below 1e-5. Bars: the refined product within 1e-12 of the
oracle (hi + lo is v to 2^-48); the hi-only product within 1e-6 and, as evidence that
the float path ran, not within 1e-12; Davidson and whole-DMRG energies within 1e-9 of
the oracle and of FCI, as in the FP64 tests."""
import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import hubbard, mx_table, random_abinitio
from fci import fci_energy
from helpers import load_site, symm_to_prog

pytestmark = pytest.mark.gpu


def hubbard_site(oracle_lib):
    T, V = hubbard(10, U=4.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, V)
    o.init(1000, 3)
    o.sweep_left_to(5, 1000, rtol=1e-3)
    o.site(5)
    return o


@pytest.mark.parametrize("explicit_max", ["100000000", "4096"])
def test_mixed_apply(oracle_lib, explicit_max, monkeypatch):
    print(f"case mixed apply, explicit max {explicit_max}", flush=True)
    monkeypatch.setenv("CHM_EXPLICIT_MAX", explicit_max)
    c = chemps2_amd.Context(0)
    c.set_heff_form(chemps2_amd.FORM_MIXED)
    o = hubbard_site(oracle_lib)
    V = load_site(c, o, 5)
    assert c.plan_stats()["explicit_pairs"] > 0
    x = np.random.default_rng(5).normal(size=V)
    y0, _ = o.apply(x)
    scale = np.abs(y0).max()
    assert np.abs(c.apply(x) - y0).max() <= 1e-12 * scale
    err = np.abs(c.apply(x, f32=True) - y0).max()
    print(f"  hi-only product: max error {err / scale:.3e} relative", flush=True)
    assert 1e-12 * scale < err <= 1e-6 * scale
    c.close()


def test_apply_f32_needs_mixed_form(oracle_lib):
    c = chemps2_amd.Context(0)
    o = hubbard_site(oracle_lib)
    V = load_site(c, o, 5)
    with pytest.raises(chemps2_amd.ChmError):
        c.apply(np.ones(V), f32=True)
    c.close()


@pytest.mark.parametrize("index", [1, 3, 5])
def test_mixed_davidson(oracle_lib, index):
    print(f"case mixed Davidson site {index}", flush=True)
    irr, T, V = random_abinitio(8, 4, group=7)
    o = oracle_lib.OracleDMRG(8, 7, irr, 8, 0, 0, T, V)
    o.init(16, 4)
    o.sweep_left_to(index, 16, rtol=1e-4)
    o.site(index)
    e0, _, _ = o.davidson(rtol=1e-10)
    c = chemps2_amd.Context(0)
    c.set_heff_form(chemps2_amd.FORM_MIXED)
    load_site(c, o, index)
    guess = symm_to_prog(o, o.site_x())
    e1, nm = c.davidson(guess, 1e-10)
    low, full = c.mixed_stats()
    print(f"  oracle {e0!r} mixed {e1!r}: {low} hi-only + {full} refined products", flush=True)
    assert abs(e1 - e0) <= 1e-9
    assert low > 0 and full > 0 and low + full == nm
    c.close()


@pytest.mark.parametrize("L,N,TwoS,seed", [(6, 6, 0, 1), (8, 6, 2, 3)])
def test_mixed_dmrg_equals_fci(L, N, TwoS, seed):
    irr, T, V = random_abinitio(L, seed, group=7)
    for target in range(8):
        try:
            ef = fci_energy(L, irr, N, TwoS, target, T, V)
            break
        except ValueError:
            continue
    print(f"case mixed DMRG L={L} N={N} 2S={TwoS} seed={seed}", flush=True)
    c = chemps2_amd.Context(0)
    c.set_heff_form(chemps2_amd.FORM_MIXED)
    c.set_problem(L, 7, irr, N, TwoS, target, mx_table(T, V, N))
    c.dmrg_init(200, 3)
    e = c.dmrg_solve([(200, 1e-11, 10, 0.0, 1e-10)])
    low, full = c.mixed_stats()
    c.close()
    print(f"  gpu {e!r} fci {ef!r}: {low} hi-only + {full} refined products", flush=True)
    assert abs(e - ef) < 1e-9
    assert low > 0
