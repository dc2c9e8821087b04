"""CPU checks against golden vectors written by the reference CheMPS2 library itself
(built from its sources by oracle/ref.mk; tests/golden/make_ref_vectors.py).

The virtual dimensions of SyBookkeeper(Prob, D) (SyBookkeeper.cpp:29-269) fix every
sector list, every V and every work list, so they are compared bit for bit: the
library's own bookkeeper (chemps2_amd/csrc/bookkeeper.cpp, host build) and the oracle's
(oracle/core.cpp), current and FCI dimensions, at the benchmark configurations.
"""
import os

import numpy as np
import pytest

import plancheck
from chemps2_amd.problems import mx_table

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_dims.npz"))


def _cases():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mrv", os.path.join(HERE, "golden", "make_ref_vectors.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.DIMS_CASES


CASES = _cases()


def _rows(name, b):
    g = GOLD[name]
    return g[g[:, 0] == b][:, 1:]


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_bookkeeper_matches_reference(name):
    mk, D = CASES[name]
    p = mk()
    l = plancheck.lib()
    h = l.pc_create(p["L"], p["group"], np.asarray(p["irreps"], dtype=np.int32), p["N"], p["TwoS"], p["Irrep"],
                    mx_table(p["T"], p["V"], p["N"]))
    l.pc_sectors_auto(h, D)
    try:
        for b in range(p["L"] + 1):
            ref = _rows(name, b)  # N, 2S, I, current, fci
            n = l.pc_nsectors(h, b, None)
            cur = np.zeros(4 * n, dtype=np.int32)
            l.pc_nsectors(h, b, cur.ctypes.data)
            want = ref[ref[:, 3] > 0][:, [0, 1, 2, 3]]
            assert np.array_equal(cur.reshape(n, 4), want), (name, b)
            nf = l.pc_fci(h, b, None)
            fci = np.zeros(4 * nf, dtype=np.int32)
            l.pc_fci(h, b, fci.ctypes.data)
            assert np.array_equal(fci.reshape(nf, 4), ref[:, [0, 1, 2, 4]]), (name, b)
    finally:
        l.pc_destroy(h)


@pytest.mark.parametrize("name", ["hubbard_L12_N12_D200", "abinitio_L24_seed0_D1000",
                                  "abinitio_L10_seed3_N9_S3_I2_C2v_D50", "n2_sto3g_D30"])
def test_oracle_bookkeeper_matches_reference(oracle_lib, name):
    mk, D = CASES[name]
    p = mk()
    o = oracle_lib.OracleDMRG(p["L"], p["group"], np.asarray(p["irreps"]), p["N"], p["TwoS"], p["Irrep"], p["T"],
                              p["V"])
    o.bookkeeper(D)
    for b in range(p["L"] + 1):
        ref = _rows(name, b)
        assert np.array_equal(o.sectors(b), ref[ref[:, 3] > 0][:, [0, 1, 2, 3]]), (name, b)


def test_golden_dims_cover_the_benchmark_workload():
    # the middle boundaries of the L = 32, D = 2000 workload: V at sites 15-16 comes out
    # of these integers (186,316 with the SyBookkeeper dimensions)
    g = GOLD["abinitio_L32_seed0_D2000"]
    assert g.shape[1] == 6 and (g[:, 4] >= 1).all()
    tot = [g[(g[:, 0] == b)][:, 4].sum() for b in range(33)]
    assert tot[0] == 1 and tot[32] == 1 and max(tot) >= 1900
