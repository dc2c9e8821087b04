"""The oracle's 2-RDM (oracle/twodm.cpp: TwoDM::FillSite and the left sweep of
DMRG::calc_rdms_and_correlations, DMRGtechnics.cpp:40-110) against the 2-RDMs the
reference library itself wrote (tests/golden/ref_rdm.npz, tests/golden/make_ref_vectors.py)
and against what the mathematics fixes: the double trace sum_ij A[i,j,i,j] = N (N - 1),
the energy 0.5 sum A h + Econst equals the DMRG energy, the index symmetries of
TwoDM::set_2rdm_A_DMRG (TwoDM.cpp:76-96), and the 1-RDM trace N."""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_rdm.npz"))


def _cases():
    spec = importlib.util.spec_from_file_location("mrv", os.path.join(HERE, "golden", "make_ref_vectors.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.RDM_CASES


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_2rdm_matches_reference(name, oracle_lib):
    mk, sched = CASES[name]
    p = mk()
    o = oracle_lib.OracleDMRG(p["L"], p["group"], p["irreps"], p["N"], p["TwoS"], p["Irrep"], p["T"], p["V"], p["econst"])
    o.init(sched[0][0], 1)
    E = o.solve(sched)
    A, B, e_rdm = o.two_rdm()
    L, N = p["L"], p["N"]
    Ar = GOLD[name + "_A"].reshape(L, L, L, L, order="F")
    Br = GOLD[name + "_B"].reshape(L, L, L, L, order="F")
    e_ref = GOLD[name + "_energy"]
    # same converged state (econv 1e-12): energies to 1e-9, the RDMs to 1e-7
    assert abs(E - e_ref[0]) < 1e-9
    assert np.abs(A - Ar).max() < 1e-7
    assert np.abs(B - Br).max() < 1e-7
    assert abs(np.einsum("ijij", A) - N * (N - 1)) < 1e-8
    assert abs(e_rdm - E) < 1e-8
    assert abs(e_ref[1] - e_ref[0]) < 1e-8
    # TwoDM::set_2rdm_A_DMRG symmetries
    for R in (A, B):
        assert np.abs(R - R.transpose(1, 0, 3, 2)).max() < 1e-12
        assert np.abs(R - R.transpose(2, 3, 0, 1)).max() < 1e-12
    # 1-RDM (TwoDM::get1RDM_DMRG) trace
    one = np.einsum("ikjk->ij", A) / (N - 1)
    assert abs(np.trace(one) - N) < 1e-8
