"""Lower-state projector terms of excited-state DMRG on the GPU, through the C ABI.

Heff gains sum_q (v_q . x) v_q and its diagonal sum_q v_q^2, where v_q is the
reference's VeffTilde of lower state q (HeffDiagrams1.cpp:65-86, HeffDiagonal.cpp:621-642).
Bars: Heff x and the diagonal within 1e-12 relative of the plain product plus those
terms; removing the vectors gives the plain product bit for bit; and, fixed by the
mathematics, the Davidson on Heff + E_shift w0 w0^T (w0 the ground state of the dense
Heff built column by column with the oracle) converges to the dense Heff's second
eigenvalue within 1e-9, in the FP64 and the mixed form and with the learned
preconditioner.
"""
import numpy as np
import pytest
import torch

import chemps2_amd
from chemps2_amd.problems import hubbard
from helpers import load_site
from test_gpu_heff import d2h_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = chemps2_amd.Context(0)
    yield c
    c.close()


def c1_case(oracle_lib):
    # C1 Hubbard L=10 at D=1000, site 5: V in the tens of thousands (many dot partials)
    T, Vint = hubbard(10, U=4.0)
    o = oracle_lib.OracleDMRG(10, 0, [0] * 10, 10, 0, 0, T, Vint)
    o.init(1000, 3)
    o.sweep_left_to(5, 1000, rtol=1e-3)
    o.site(5)
    return o


@pytest.mark.parametrize("case", ["d2h", "c1"])
def test_lower_terms_apply_and_diag(ctx, oracle_lib, case):
    site = 3 if case == "d2h" else 5
    o = d2h_case(oracle_lib, 3) if case == "d2h" else c1_case(oracle_lib)
    V = load_site(ctx, o, site)
    print(f"case {case} L={o.L} site={site} V={V}", flush=True)
    rng = np.random.default_rng(5)
    x = rng.normal(size=V)
    y0, d0 = ctx.apply(x), ctx.diag()
    W = rng.normal(size=(2, V))
    ctx.set_lower(W)
    y1, d1 = ctx.apply(x), ctx.diag()
    ye = y0 + W.T @ (W @ x)
    de = d0 + (W * W).sum(axis=0)
    assert np.abs(y1 - ye).max() <= 1e-12 * np.abs(ye).max()
    assert np.abs(d1 - de).max() <= 1e-12 * np.abs(de).max()
    # device-resident vectors, input and output
    Wd = torch.from_numpy(W[:1].copy()).cuda()
    ctx.set_lower(Wd)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(V, dtype=torch.float64, device="cuda")
    ctx.apply(xd, yd)
    torch.cuda.synchronize()
    ye1 = y0 + W[0] * (W[0] @ x)
    assert np.abs(yd.cpu().numpy() - ye1).max() <= 1e-12 * np.abs(ye1).max()
    ctx.set_lower(None)
    assert np.array_equal(ctx.apply(x), y0)
    assert np.array_equal(ctx.diag(), d0)
    # a new prepare clears them
    ctx.set_lower(W)
    ctx.prepare(site)
    assert np.array_equal(ctx.apply(x), y0)
    with pytest.raises(ValueError):
        ctx.set_lower(np.zeros((1, V + 1)))


@pytest.mark.parametrize("form,learned", [("fp64", False), ("mixed", False), ("mixed", True)])
def test_lower_state_davidson_gives_second_eigenvalue(oracle_lib, form, learned):
    import os

    o = d2h_case(oracle_lib, 3)
    c = chemps2_amd.Context(0)
    try:
        if form == "mixed":
            c.set_heff_form(chemps2_amd.FORM_MIXED)
        if learned:
            path = os.path.join(os.path.dirname(chemps2_amd.__file__), "data", "precond_mlp.npy")
            c.mlp_set(np.load(path, allow_pickle=False))
            c.precond_mode(chemps2_amd.PRECOND_LEARNED)
        V = load_site(c, o, 3)
        H = np.empty((V, V))
        e = np.zeros(V)
        for i in range(V):
            e[i] = 1.0
            H[:, i] = o.apply(e)[0]
            e[i] = 0.0
        ev, vec = np.linalg.eigh(0.5 * (H + H.T))
        print(f"{form} learned={learned} V={V}: E0 {ev[0]!r} E1 {ev[1]!r}", flush=True)
        assert ev[1] - ev[0] > 1e-4
        shift = 2.0 * (ev[1] - ev[0]) + 1.0
        c.set_lower(np.sqrt(shift) * vec[:, :1].T)
        e1, nm = c.davidson(np.random.default_rng(2).normal(size=V), 1e-10)
        print(f"Davidson with the lower state: {e1!r} ({nm} products)", flush=True)
        assert abs(e1 - ev[1]) <= 1e-9
        c.set_lower(None)
        e0, _ = c.davidson(np.random.default_rng(3).normal(size=V), 1e-10)
        assert abs(e0 - ev[0]) <= 1e-9
    finally:
        c.close()
