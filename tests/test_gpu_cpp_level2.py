"""The C++ drop-in at level 2: tests/integration/level2_check.cpp, compiled
against the reference's own headers and library (oracle/_ref, built from the reference's
sources by oracle/ref.mk) and against libchemps2_amd.so, builds a CheMPS2::Problem from the
N2 STO-3G FCIDUMP with SetupReorderD2h, solves it through chm_dmrg_init / chm_dmrg_solve as
the replaced DMRG constructor and DMRG::Solve would, and solves it again with the
reference's own DMRG on the host: the energies agree to 1e-8 (tests/test5.cpp.in:83)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "level2_check")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/level2_check is built only where the reference's sources are")
def test_cpp_level2_n2_sto3g():
    fcidump = os.path.join(ROOT, "tests", "golden", "N2.STO3G.FCIDUMP")
    print("case C++ level-2 check, N2 STO-3G D=1000", flush=True)
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "16"), MKL_THREADING_LAYER="GNU")
    r = subprocess.run([BIN, fcidump, "7", "14", "0", "0", "1", "1000", "1e-10", "20"], capture_output=True, text=True,
                       timeout=240, env=env)
    line = [x for x in r.stdout.splitlines() if x.startswith("level2_check:")]
    print(line[-1] if line else r.stdout[-2000:] + r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    e_gpu = float(line[-1].split("chemps2_amd")[1].split(",")[0])
    assert abs(e_gpu + 107.648250974014) < 1e-8
