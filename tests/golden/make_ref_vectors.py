"""Writes golden vectors from the reference CheMPS2 library, built from its own sources by
``oracle/ref.mk`` (run ``make -f oracle/ref.mk -j8`` first; needs /root/reference).

    python tests/golden/make_ref_vectors.py dims      -> tests/golden/ref_dims.npz
    python tests/golden/make_ref_vectors.py energies  -> tests/golden/ref_energies.json
    python tests/golden/make_ref_vectors.py rdm       -> tests/golden/ref_rdm.npz

dims: SyBookkeeper(Prob, D) (SyBookkeeper.cpp:29-269) for the benchmark configurations
and a few small odd cases: rows (b, N, 2S, I, current dimension, FCI dimension) of every
sector with a non-zero FCI dimension.
energies: the reference DMRG's converged energies (DMRG::Solve, DMRG.cpp:268-417) for
problems that no closed form or FCI of this repository covers at their size.
rdm: the reference's 2-RDM (DMRG::calc2DMandCorrelations, DMRGtechnics.cpp:40-261;
TwoDM::getTwoDMA_DMRG / getTwoDMB_DMRG) of small problems.
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from chemps2_amd.problems import (hubbard, hubbard_cylinder, molpro_to_xor, permute, random_abinitio,  # noqa: E402
                                  read_fcidump, reorder_d2h)
from oracle import refdrv  # noqa: E402


def n2_sto3g():
    d = read_fcidump(os.path.join(HERE, "N2.STO3G.FCIDUMP"))
    irr = [molpro_to_xor(7, m) for m in d["molpro_irreps"]]
    irr2, T2, V2 = permute(irr, d["T"], d["V"], reorder_d2h(irr))
    return dict(L=d["L"], group=7, irreps=list(irr2), N=d["N"], TwoS=0, Irrep=0, T=T2, V=V2, econst=d["econst"])


def abinitio(L, seed, N, TwoS=0, Irrep=0, group=7):
    irr, T, V = random_abinitio(L, seed, group=group)
    return dict(L=L, group=group, irreps=list(irr), N=N, TwoS=TwoS, Irrep=Irrep, T=T, V=V, econst=0.0)


def hub(L, N, U, TwoS=0):
    T, V = hubbard(L, U=U)
    return dict(L=L, group=0, irreps=[0] * L, N=N, TwoS=TwoS, Irrep=0, T=T, V=V, econst=0.0)


def cylinder(N=28):
    T, V, _ = hubbard_cylinder(8, 4, U=8.0)
    return dict(L=32, group=0, irreps=[0] * 32, N=N, TwoS=0, Irrep=0, T=T, V=V, econst=0.0)


# name -> (problem, D)
DIMS_CASES = {
    "hubbard_L12_N12_D200": (lambda: hub(12, 12, 4.0), 200),
    "abinitio_L24_seed0_D1000": (lambda: abinitio(24, 0, 24), 1000),
    "abinitio_L32_seed0_D2000": (lambda: abinitio(32, 0, 32), 2000),
    "cylinder_4x8_N28_D3000": (lambda: cylinder(28), 3000),
    "abinitio_L10_seed3_N9_S3_I2_C2v_D50": (lambda: abinitio(10, 3, 9, TwoS=3, Irrep=2, group=5), 50),
    "n2_sto3g_D1000": (n2_sto3g, 1000),
    "n2_sto3g_D30": (n2_sto3g, 30),
}


def problem_file(tmp, name, p):
    path = os.path.join(tmp, name + ".bin")
    refdrv.write_problem(path, p["L"], p["group"], p["irreps"], p["N"], p["TwoS"], p["Irrep"], p["T"], p["V"],
                         p["econst"], p.get("mx"))
    return path


def make_dims():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (mk, D) in DIMS_CASES.items():
            p = mk()
            rows = []
            for b, lst in refdrv.dims(problem_file(tmp, name, p), D).items():
                rows += [(b,) + r for r in lst]
            out[name] = np.array(rows, dtype=np.int32)
            print(name, out[name].shape, flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_dims.npz"), **out)


# name -> (problem, schedule, nstates, eshift)
ENERGY_CASES = {
    # config-2 generator at a size the reference converges in minutes on 8 cores
    "abinitio_L16_seed5_N16": (lambda: abinitio(16, 5, 16), [(500, 1e-10, 20, 0.0, 1e-8)], 1, 0.0),
    # odd electron count, spin 3/2, a non-trivial target irrep, C2v
    "abinitio_L12_seed7_N11_S3_I1_C2v": (lambda: abinitio(12, 7, 11, TwoS=3, Irrep=1, group=5),
                                         [(400, 1e-10, 20, 0.0, 1e-8)], 1, 0.0),
    # doped 1D Hubbard at U = 8
    "hubbard_L16_N14_U8": (lambda: hub(16, 14, 8.0), [(500, 1e-10, 20, 0.0, 1e-8)], 1, 0.0),
    # three lowest states of a synthetic D2h problem (excited-state driver, test5 style)
    "abinitio_L10_seed11_N10_3states": (lambda: abinitio(10, 11, 10), [(300, 1e-10, 20, 0.0, 1e-8)], 3, 20.0),
}


def make_energies():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (mk, sched, nst, esh) in ENERGY_CASES.items():
            p = mk()
            r = refdrv.dmrg(problem_file(tmp, name, p), sched, nstates=nst, eshift=esh)
            out[name] = {"energies": r["energies"], "schedule": sched, "nstates": nst, "eshift": esh,
                         "seconds": r["seconds"]}
            print(name, r["energies"], r["seconds"], flush=True)
    with open(os.path.join(HERE, "ref_energies.json"), "w") as f:
        json.dump(out, f, indent=1)


RDM_CASES = {
    "n2_sto3g": (n2_sto3g, [(1000, 1e-12, 100, 0.0, 1e-8)]),
    "hubbard_L8_N6_S2_U4": (lambda: hub(8, 6, 4.0, TwoS=2), [(500, 1e-12, 40, 0.0, 1e-8)]),
    "abinitio_L8_seed4_N7_S1_C2v": (lambda: abinitio(8, 4, 7, TwoS=1, group=5), [(500, 1e-12, 40, 0.0, 1e-8)]),
}


def make_rdm():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (mk, sched) in RDM_CASES.items():
            p = mk()
            r = refdrv.dmrg(problem_file(tmp, name, p), sched, rdm=True)
            A, B = r["rdm"]
            out[name + "_A"] = A
            out[name + "_B"] = B
            out[name + "_energy"] = np.array([r["energies"][-1], r["rdm_energy"]])
            print(name, r["energies"], r["rdm_energy"], flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_rdm.npz"), **out)


if __name__ == "__main__":
    what = sys.argv[1:] or ["dims", "energies", "rdm"]
    if "dims" in what:
        make_dims()
    if "energies" in what:
        make_energies()
    if "rdm" in what:
        make_rdm()
