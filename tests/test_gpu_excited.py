"""GPU tests of excited-state DMRG: lower states kept on the device, overlap operators
TensorO, the VeffTilde projector in Heff and the newExcitation driver
(DMRG.cpp:464-505, DMRGtechnics.cpp:540-620, TensorO.cpp).

Pinned by the reference's own known answers (tests/test5.cpp.in:83-85: N2 STO-3G, D2h
reordered, D = 1000, level shift 20) and by energies the reference library itself
computed (tests/golden/ref_energies.json, written by tests/golden/make_ref_vectors.py
from oracle/_ref).
"""
import json
import os

import numpy as np
import pytest

import chemps2_amd
from chemps2_amd.problems import molpro_to_xor, mx_table, permute, random_abinitio, read_fcidump, reorder_d2h

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def run_states(L, group, irr, N, TwoS, Irrep, T, V, scheme, nstates, eshift, econst=0.0, seed=3):
    c = chemps2_amd.Context(0)
    try:
        c.set_problem(L, group, irr, N, TwoS, Irrep, mx_table(T, V, N), econst)
        c.dmrg_init(scheme[0][0], seed)
        out = [c.dmrg_solve(scheme)]
        for s in range(1, nstates):
            c.new_excitation(eshift, scheme[0][0], seed + 17 * s)
            assert c.nstates() == s + 1
            out.append(c.dmrg_solve(scheme))
        return out
    finally:
        c.close()


def test_reference_test5_excited_states_on_gpu():
    # tests/test5.cpp.in:39-85: ground state, then activateExcitations + newExcitation(20.0)
    # twice, each Solve with D = 1000, Econv 1e-12, 100 sweeps
    d = read_fcidump(os.path.join(GOLDEN, "N2.STO3G.FCIDUMP"))
    irr = [molpro_to_xor(7, m) for m in d["molpro_irreps"]]
    irr2, T2, V2 = permute(irr, d["T"], d["V"], reorder_d2h(irr))
    print("case test5 N2 three states", flush=True)
    e = run_states(d["L"], 7, irr2, d["N"], 0, 0, T2, V2, [(1000, 1e-12, 100, 0.0, 1e-5)], 3, 20.0, d["econst"])
    print(f"test5: gpu {e!r}", flush=True)
    assert abs(e[0] + 107.648250974014) < 1e-8
    assert abs(e[1] + 106.944757308768) < 1e-8
    assert abs(e[2] + 106.92314213886) < 1e-8


def test_excited_states_match_reference_library():
    with open(os.path.join(GOLDEN, "ref_energies.json")) as f:
        ref = json.load(f)["abinitio_L10_seed11_N10_3states"]
    irr, T, V = random_abinitio(10, 11, group=7)
    print("case synthetic L=10 seed 11 three states", flush=True)
    sched = [tuple(s) for s in ref["schedule"]]
    e = run_states(10, 7, irr, 10, 0, 0, T, V, sched, ref["nstates"], ref["eshift"])
    print(f"gpu {e!r} reference {ref['energies']!r}", flush=True)
    assert np.allclose(e, ref["energies"], atol=1e-8, rtol=0)


def test_new_excitation_needs_a_state():
    c = chemps2_amd.Context(0)
    try:
        irr, T, V = random_abinitio(6, 1, group=7)
        c.set_problem(6, 7, irr, 6, 0, 0, mx_table(T, V, 6))
        with pytest.raises(chemps2_amd.ChmError):
            c.new_excitation(20.0, 50)
        c.dmrg_init(50, 1)
        with pytest.raises(chemps2_amd.ChmError):
            c.new_excitation(-1.0, 50)
        assert c.nstates() == 1
    finally:
        c.close()
