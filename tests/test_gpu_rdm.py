"""GPU tests of the 2-RDM (chm_dmrg_2rdm: DMRG::calc_rdms_and_correlations, 2-RDM part,
DMRGtechnics.cpp:40-110; TwoDM, TwoDM.cpp:40-640).

1. The device 2-RDM of the oracle's converged MPS equals the oracle's 2-RDM of the same
   MPS (oracle/twodm.cpp, itself pinned by the reference library's 2-RDMs), to 1e-11.
2. The whole path on the device (DMRG, then the 2-RDM sweep) against the 2-RDMs the
   reference library wrote (tests/golden/ref_rdm.npz) for the same problems and
   schedules, with the double trace N (N - 1) and the energy 0.5 sum A h + Econst equal
   to the DMRG energy.
"""
import importlib.util
import os

import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import mx_table

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_rdm.npz"))


def _cases():
    spec = importlib.util.spec_from_file_location("mrv", os.path.join(HERE, "golden", "make_ref_vectors.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.RDM_CASES


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_2rdm_same_state_as_oracle(name, oracle_lib):
    mk, sched = CASES[name]
    p = mk()
    L, N = p["L"], p["N"]
    print(f"case {name}: oracle DMRG, then the 2-RDM of its MPS on the device", flush=True)
    o = oracle_lib.OracleDMRG(L, p["group"], p["irreps"], N, p["TwoS"], p["Irrep"], p["T"], p["V"], p["econst"])
    o.init(sched[0][0], 1)
    o.solve(sched)
    mps = [o.mps(s) for s in range(L)]
    secs = [o.sectors(b) for b in range(L + 1)]
    A0, B0, e0 = o.two_rdm()
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, p["group"], o.irreps(), N, p["TwoS"], p["Irrep"], o.mx(), p["econst"])
        c.set_sectors_auto(sched[0][0])
        for b in range(L + 1):
            c.set_sectors(b, secs[b])
        for s in range(L):
            c.mps_upload(s, mps[s])
        c.dmrg_presolve()
        A1, B1, e1 = c.two_rdm()
    finally:
        c.close()
    scale = max(np.abs(A0).max(), 1.0)
    print(f"{name}: |dA| {np.abs(A1 - A0).max():.2e} |dB| {np.abs(B1 - B0).max():.2e} energy {e1!r} vs {e0!r}", flush=True)
    assert np.abs(A1 - A0).max() < 1e-11 * scale
    assert np.abs(B1 - B0).max() < 1e-11 * scale
    assert abs(e1 - e0) < 1e-10


@pytest.mark.parametrize("name", sorted(CASES))
def test_2rdm_device_dmrg_matches_reference(name):
    mk, sched = CASES[name]
    p = mk()
    L, N = p["L"], p["N"]
    print(f"case {name}: DMRG and 2-RDM on the device", flush=True)
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, p["group"], p["irreps"], N, p["TwoS"], p["Irrep"], mx_table(p["T"], p["V"], N), p["econst"])
        c.dmrg_init(sched[0][0], 5)
        E = c.dmrg_solve(sched)
        A, B, e = c.two_rdm()
    finally:
        c.close()
    Ar = GOLD[name + "_A"].reshape(L, L, L, L, order="F")
    Br = GOLD[name + "_B"].reshape(L, L, L, L, order="F")
    print(f"{name}: E {E!r} ref {GOLD[name + '_energy'][0]!r}; |dA| {np.abs(A - Ar).max():.2e} |dB| {np.abs(B - Br).max():.2e}",
          flush=True)
    assert abs(E - GOLD[name + "_energy"][0]) < 1e-9
    # independent runs converged to econv 1e-12: the 2-RDMs agree to the state's accuracy
    assert np.abs(A - Ar).max() < 1e-6
    assert np.abs(B - Br).max() < 1e-6
    assert abs(np.einsum("ijij", A) - N * (N - 1)) < 1e-8
    assert abs(e - E) < 1e-8
