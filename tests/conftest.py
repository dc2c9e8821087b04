import os
import sys

import pytest

# torch brings its own HIP runtime libraries: they are loaded before the library's (and
# torch's device context opened first, see _torch_hip_first), as in bench.py, whatever
# test modules a run collects
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU case")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import wrap

    wrap.build()
    return wrap


_torch_gpu_ready = False


@pytest.fixture(autouse=True)
def _torch_hip_first(request):
    """The first GPU test of the session opens torch's device context before the library
    opens its own."""
    global _torch_gpu_ready
    if not _torch_gpu_ready and request.node.get_closest_marker("gpu") is not None:
        import torch

        torch.cuda.init()
        torch.empty(1, device="cuda:0")
        _torch_gpu_ready = True
    yield
