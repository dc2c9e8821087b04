"""Runs the reference CheMPS2 library, built from its own sources by ``oracle/ref.mk``
into ``oracle/_ref/``, through ``oracle/ref_driver.cpp``.

Test infrastructure only: the scripts under ``tests/golden/`` use it to write golden
vectors, and the CPU-baseline leg of the benchmark times it. The product never loads it.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
DRIVER = os.path.join(REF_DIR, "ref_driver")
REFERENCE = "/root/reference/CheMPS2"


def available():
    return os.path.exists(DRIVER)


def build(jobs=8):
    """Compile the reference and the driver (needs the reference sources)."""
    if not os.path.isdir(REFERENCE):
        raise RuntimeError("reference sources not present")
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "ref.mk"), f"-j{jobs}"], check=True)


def write_problem(path, L, group, irreps, N, TwoS, Irrep, T, V, econst=0.0, mx=None):
    """Problem file of ref_driver.cpp: header, irreps, econst, T, V (physicist <ij|kl>) and
    optionally the raw Problem::setMxElement table mx[a + L(b + L(c + L d))]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<7i", 0x50484D43, L, group, N, TwoS, Irrep, 1 if mx is not None else 0))
        f.write(np.asarray(irreps, dtype="<i4").tobytes())
        f.write(struct.pack("<d", econst))
        f.write(np.ascontiguousarray(T, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(V, dtype="<f8").tobytes())
        if mx is not None:
            f.write(np.ascontiguousarray(mx, dtype="<f8").tobytes())


def _run(args, threads=None, cwd=None, timeout=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    # the reference parallelises with OpenMP over blocks; MKL runs inside those threads
    env.setdefault("MKL_THREADING_LAYER", "GNU")
    r = subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, env=env, cwd=cwd,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"ref_driver failed ({r.returncode}): {r.stderr[-2000:]}")
    return r.stdout


def dims(prob_path, D):
    """SyBookkeeper(Prob, D): {b: [(N, 2S, I, current, fci), ...]}."""
    out = {}
    for line in _run(["dims", prob_path, D]).splitlines():
        b, N, S, I, cur, fci = (int(x) for x in line.split())
        out.setdefault(b, []).append((N, S, I, cur, fci))
    return out


def dmrg(prob_path, schedule, nstates=1, eshift=0.0, rdm=False, threads=None, timeout=None):
    """Reference DMRG: schedule = [(D, econv, max_sweeps, noise, rtol), ...]. Returns
    {"energies": [...], "seconds": [...], "init_seconds": t, "rdm": (A, B) or None,
    "rdm_energy": E or None}."""
    args = ["dmrg", prob_path, nstates, repr(float(eshift)), len(schedule)]
    for D, econv, ms, noise, rtol in schedule:
        args += [D, repr(float(econv)), ms, repr(float(noise)), repr(float(rtol))]
    with tempfile.TemporaryDirectory() as tmp:
        rdm_path = os.path.join(tmp, "rdm.bin")
        if rdm:
            args += ["rdm", rdm_path]
        out = _run(args, threads=threads, cwd=tmp, timeout=timeout)
        res = {"energies": [], "seconds": [], "init_seconds": None, "rdm": None, "rdm_energy": None}
        for line in out.splitlines():
            p = line.split()
            if len(p) == 6 and p[0] == "state" and p[2] == "energy":
                res["energies"].append(float(p[3]))
                res["seconds"].append(float(p[5]))
            elif len(p) == 3 and p[0] == "init" and p[1] == "seconds":
                res["init_seconds"] = float(p[2])
            elif len(p) >= 3 and p[0] == "rdm" and p[1] == "energy":
                res["rdm_energy"] = float(p[2])
        if rdm:
            a = np.fromfile(rdm_path, dtype="<f8")
            n = a.size // 2
            res["rdm"] = (a[:n].copy(), a[n:].copy())
        res["stdout"] = out
    return res


def steps(prob_path, D, first, last, rtol=1e-5, budget_s=1e30, threads=None, chkpt_file=None, timeout=None):
    """The reference's first left-sweep steps (ref_driver steps), until they have taken
    budget_s seconds: {"init_seconds": t, "steps": [(index, energy, seconds, products),
    ...]}; chkpt_file: an MPS checkpoint in the reference's HDF5 layout to start from."""
    import shutil

    with tempfile.TemporaryDirectory() as tmp:
        args = ["steps", prob_path, D, first, last, repr(float(rtol)), repr(float(budget_s))]
        if chkpt_file:
            shutil.copy(chkpt_file, os.path.join(tmp, "CheMPS2_MPS0.h5"))
            args.append("chkpt")
        out = _run(args, threads=threads, cwd=tmp, timeout=timeout)
    res = {"init_seconds": None, "steps": []}
    nits = []
    for line in out.splitlines():
        p = line.split()
        if len(p) == 3 and p[0] == "init" and p[1] == "seconds":
            res["init_seconds"] = float(p[2])
        elif len(p) == 6 and p[0] == "step" and p[2] == "energy":
            res["steps"].append([int(p[1]), float(p[3]), float(p[5])])
        elif "nIt(DAVIDSON)" in line:
            nits.append(int(line.split("=")[-1]))
    # the last nIt lines belong to the steps (one Davidson per step)
    for q, st in enumerate(res["steps"]):
        k = len(nits) - len(res["steps"]) + q
        st.append(nits[k] if 0 <= k < len(nits) else None)
    return res
