"""MPS checkpoints in the reference's HDF5 layout (DMRGmpsio.cpp:27-162), host side.

The file layout is checked with h5dump against names and types built here from the
reference's saveMPS, and the sector list against brute-force reachability: a sector
(b, N, 2S) is listed iff b orbitals can hold N electrons at spin S and the other L - b
orbitals can complete it to the target (N, 2S) of the whole chain."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import chemps2_amd

pytestmark = pytest.mark.skipif(not os.path.exists(chemps2_amd.IO_LIB_PATH), reason="checkpoint library not built (no HDF5)")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def spins(k, n):
    """Twice the spins n electrons can couple to on k spatial orbitals."""
    if n < 0 or n > 2 * k:
        return []
    return list(range(n % 2, min(n, 2 * k - n) + 1, 2))


def reachable(L, N, TS, nirrep):
    out = []
    for b in range(L + 1):
        for n in range(2 * b + 1):
            for s in spins(b, n):
                if any(abs(s - sr) <= TS <= s + sr for sr in spins(L - b, N - n)):
                    out += [(b, n, s, i) for i in range(nirrep)]
    return out


def write_random(path, L, N, TS, nirrep, converged, seed):
    rng = np.random.default_rng(seed)
    secs = reachable(L, N, TS, nirrep)
    dims = rng.integers(0, 5, size=len(secs)).astype(np.int32)
    mps = [rng.standard_normal(int(rng.integers(1, 40))) for _ in range(L)]
    chemps2_amd.io_write(str(path), L, N, TS, nirrep, converged, dims, mps)
    return secs, dims, mps


@pytest.mark.parametrize("L,N,TS,nirrep", [(2, 2, 0, 1), (5, 4, 2, 2), (6, 7, 1, 4), (8, 8, 0, 8)])
def test_round_trip(tmp_path, L, N, TS, nirrep):
    path = tmp_path / "CheMPS2_MPS0.h5"
    secs, dims, mps = write_random(path, L, N, TS, nirrep, True, L * 100 + N)
    assert np.array_equal(chemps2_amd.io_read_dims(str(path), L, N, TS, nirrep), dims)
    for s in range(L):
        vals, conv = chemps2_amd.io_read_site(str(path), s, len(mps[s]))
        assert conv == 1
        assert np.array_equal(vals, mps[s])


def test_sector_list_is_reachability():
    # chm_io_read_dims with no output array returns the number of (b, N, 2S, I) entries
    # of SyBookkeeper's ranges (SyBookkeeper.cpp:84-92); the names themselves are
    # compared in test_layout_matches_save_mps
    for L in range(2, 9):
        for N in range(2 * L + 1):
            for TS in spins(L, N):
                cnt = ctypes.c_int(0)
                assert chemps2_amd.load_io().chm_io_read_dims(b"unused", L, N, TS, 2, None, ctypes.byref(cnt)) == 0
                assert cnt.value == len(reachable(L, N, TS, 2)), (L, N, TS)


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_layout_matches_save_mps(tmp_path):
    # DMRGmpsio.cpp:30-92: /Convergence/Converged_yn (I32LE), /VirtDim_b_N_2S_I/Value
    # (I32LE) and /MPS_site/Values (IEEE F64LE, TensorT length)
    L, N, TS, nirrep = 5, 4, 2, 2
    path = tmp_path / "CheMPS2_MPS0.h5"
    secs, dims, mps = write_random(path, L, N, TS, nirrep, False, 7)
    listing = subprocess.run([H5DUMP, "-n", str(path)], capture_output=True, text=True, check=True).stdout
    objs = {tuple(line.split()) for line in listing.splitlines() if line.split()[:1] in (["group"], ["dataset"])}
    want = {("group", "/"), ("group", "/Convergence"), ("dataset", "/Convergence/Converged_yn")}
    for b, n, s, i in secs:
        want |= {("group", f"/VirtDim_{b}_{n}_{s}_{i}"), ("dataset", f"/VirtDim_{b}_{n}_{s}_{i}/Value")}
    for s in range(L):
        want |= {("group", f"/MPS_{s}"), ("dataset", f"/MPS_{s}/Values")}
    assert objs == want
    head = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True).stdout
    assert head.count("H5T_STD_I32LE") == 1 + len(secs)
    assert head.count("H5T_IEEE_F64LE") == L
    for s in range(L):
        assert f"SIMPLE {{ ( {len(mps[s])} ) / ( {len(mps[s])} ) }}" in head
    conv = subprocess.run([H5DUMP, "-d", "/Convergence/Converged_yn", str(path)], capture_output=True, text=True,
                          check=True).stdout
    assert "(0): 0" in conv


def test_errors(tmp_path):
    path = tmp_path / "m.h5"
    _, _, mps = write_random(path, 4, 4, 0, 1, False, 1)
    with pytest.raises(chemps2_amd.ChmError):
        chemps2_amd.io_read_site(str(path), 0, len(mps[0]) + 1)  # wrong TensorT size
    with pytest.raises(chemps2_amd.ChmError):
        chemps2_amd.io_read_site(str(path), 4, 3)  # no such site
    with pytest.raises(chemps2_amd.ChmError):
        chemps2_amd.io_read_dims(str(tmp_path / "absent.h5"), 4, 4, 0, 1)
    with pytest.raises(chemps2_amd.ChmError):
        chemps2_amd.io_read_dims(str(path), 4, 4, 0, 2)  # other symmetry group: sectors absent
