"""Checkpoint and restart on the device: DMRG::saveMPS after Solve, then in a new
context SyBookkeeper(D) + loadDIM + loadMPS + PreSolve + Solve
(DMRG.cpp:100-116, DMRGmpsio.cpp:27-162)."""
import os

import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import mx_table, random_abinitio
from fci import fci_energy

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(chemps2_amd.IO_LIB_PATH), reason="checkpoint library not built (no HDF5)")]


def test_save_load_resume(tmp_path):
    print("case checkpoint L=8 N=8 D2h D=200", flush=True)
    L, N, TS, Irrep, D = 8, 8, 0, 0, 200
    irr, T, V = random_abinitio(L, 11, group=7)
    mx = mx_table(T, V, N)
    path = str(tmp_path / "CheMPS2_MPS0.h5")

    a = chemps2_amd.Context(0)
    a.set_problem(L, 7, irr, N, TS, Irrep, mx)
    a.dmrg_init(8, 5)
    e1 = a.dmrg_solve([(D, 1e-11, 10, 0.0, 1e-9)])
    a.mps_save(path, converged=True)
    secs = [a.get_sectors(b) for b in range(L + 1)]
    sites = [a.mps_download(s) for s in range(L)]
    a.close()
    assert abs(e1 - fci_energy(L, irr, N, TS, Irrep, T, V)) < 1e-8

    b = chemps2_amd.Context(0)
    b.set_problem(L, 7, irr, N, TS, Irrep, mx)
    b.set_sectors_auto(D)
    b.mps_load_dims(path)
    assert all(np.array_equal(b.get_sectors(k), secs[k]) for k in range(L + 1))
    assert b.mps_load(path) is True
    assert all(np.array_equal(b.mps_download(s), sites[s]) for s in range(L))
    b.dmrg_presolve()
    e2 = b.dmrg_solve([(D, 1e-11, 2, 0.0, 1e-9)])
    b.close()
    assert abs(e2 - e1) < 1e-9
