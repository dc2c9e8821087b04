"""The learned preconditioner on the device: the FP32 network kernels against the
float64 numpy network (tests/mlp_ref.py, itself pinned by finite differences), training,
the identity of a zero network with the diagonal preconditioner, and harvest -> train ->
solve on an unseen Hamiltonian with the energy unchanged."""
import numpy as np
import pytest

import chemps2_amd
import mlp_ref
from chemps2_amd.problems import mx_table, random_abinitio
from fci import fci_energy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = chemps2_amd.Context(0)
    yield c
    c.close()


def random_params(seed, scale=0.15):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=chemps2_amd.mlp_nparams()).astype(np.float32)


def test_forward_matches_numpy(ctx):
    print("case network forward", flush=True)
    params = random_params(1)
    ctx.mlp_set(params)
    X = np.random.default_rng(2).normal(size=(1000, chemps2_amd.MLP_IN)).astype(np.float32)
    got = ctx.mlp_forward(X)
    ref, _ = mlp_ref.forward(params.astype(np.float64), X.astype(np.float64))
    assert np.abs(got - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [700, 5000])
def test_gradient_matches_numpy(ctx, n):
    # n = 5000 runs the weight gradients with K split over workgroups
    print(f"case network gradient n={n}", flush=True)
    params = random_params(3)
    ctx.mlp_set(params)
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, chemps2_amd.MLP_IN)).astype(np.float32)
    p, e = rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    loss, g = ctx.mlp_loss_grad(X, p, e, w)
    lref, gref = mlp_ref.loss_grad(params.astype(np.float64), X, p, e, w)
    print(f"  loss {loss!r} ref {lref!r}; gradient error {np.linalg.norm(g - gref) / np.linalg.norm(gref):.2e}", flush=True)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-3 * np.linalg.norm(gref)


def test_training_reduces_loss(ctx):
    print("case network training", flush=True)
    rng = np.random.default_rng(7)
    n = 65536
    X = rng.normal(size=(n, chemps2_amd.MLP_IN)).astype(np.float32)
    X[:, 5:] = 0.0
    p = rng.normal(size=n).astype(np.float32)
    e = (p * (1.0 + 0.5 * np.tanh(X[:, 0]) - 0.2 * X[:, 1] ** 2 / (1 + X[:, 1] ** 2))).astype(np.float32)
    w = np.ones(n, dtype=np.float32)
    ctx.mlp_set(chemps2_amd.mlp_init(0))
    l0, l1 = ctx.mlp_train(X, p, e, w, batch=2048, epochs=6, lr=2e-3, seed=3)
    print(f"  loss first epoch {l0:.4e}, last {l1:.4e}", flush=True)
    assert l1 < 0.1 * l0


def d2h_problem(seed, L=8, N=8):
    irr, T, V = random_abinitio(L, seed, group=7)
    return irr, T, V, mx_table(T, V, N)


def run(irr, mx, L, N, params=None, harvest=0, D=200):
    c = chemps2_amd.Context(0)
    c.set_problem(L, 7, irr, N, 0, 0, mx)
    if params is not None:
        c.mlp_set(params)
        c.precond_mode(chemps2_amd.PRECOND_LEARNED)
    if harvest:
        c.harvest(harvest)
    c.dmrg_init(D, 3)
    e = c.dmrg_solve([(D, 1e-11, 8, 0.0, 1e-9)])
    _, nmv = c.dmrg_stats()
    data = c.harvest_take() if harvest else None
    c.close()
    return e, nmv, data


def test_zero_network_equals_diagonal():
    print("case zero network", flush=True)
    irr, T, V, mx = d2h_problem(5)
    e0, n0, _ = run(irr, mx, 8, 8)
    e1, n1, _ = run(irr, mx, 8, 8, params=chemps2_amd.mlp_init(0))
    assert e1 == e0 and n1 == n0


def test_harvest_train_solve_unseen():
    print("case harvest, train, solve an unseen Hamiltonian", flush=True)
    parts = []
    for seed in (11, 12, 13):
        irr, T, V, mx = d2h_problem(seed)
        _, _, data = run(irr, mx, 8, 8, harvest=256)
        assert all(np.all(np.isfinite(a)) for a in data) and len(data[1]) > 0
        parts.append(data)
    X, p, e, w = (np.concatenate([d[k] for d in parts]) for k in range(4))
    c = chemps2_amd.Context(0)
    c.mlp_set(chemps2_amd.mlp_init(0))
    l0, l1 = c.mlp_train(X, p, e, w, batch=1024, epochs=8, lr=1e-3, seed=1)
    params = c.mlp_get()
    c.close()
    irr, T, V, mx = d2h_problem(99)
    ef = fci_energy(8, irr, 8, 0, 0, T, V)
    ed, nd, _ = run(irr, mx, 8, 8)
    el, nl, _ = run(irr, mx, 8, 8, params=params)
    print(f"  {len(p)} samples, loss {l0:.4e} -> {l1:.4e}; products diagonal {nd}, learned {nl}; "
          f"E diagonal {ed!r} learned {el!r} FCI {ef!r}", flush=True)
    assert l1 < l0
    assert abs(ed - ef) < 1e-9 and abs(el - ef) < 1e-9


def test_committed_network_unseen():
    # the parameters shipped in chemps2_amd/data (trained by chemps2_amd.precond_train on
    # L = 24, D = 1000 sweeps) on a Hamiltonian none of its samples came from
    import os

    print("case committed network", flush=True)
    path = os.path.join(os.path.dirname(chemps2_amd.__file__), "data", "precond_mlp.npy")
    params = np.load(path, allow_pickle=False)
    irr, T, V, mx = d2h_problem(77)
    ef = fci_energy(8, irr, 8, 0, 0, T, V)
    ed, nd, _ = run(irr, mx, 8, 8)
    el, nl, _ = run(irr, mx, 8, 8, params=params)
    print(f"  products diagonal {nd}, learned {nl}; |E - FCI| {abs(ed - ef):.1e} / {abs(el - ef):.1e}", flush=True)
    assert abs(ed - ef) < 1e-9 and abs(el - ef) < 1e-9
