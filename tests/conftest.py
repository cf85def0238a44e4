import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "arc-demo_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run on the MI355X box)")
    config.addinivalue_line("markers", "slow: longer CPU test")
    # When GPU tests may run, torch's HIP runtime is loaded before any test
    # loads libarcsim, so that the library binds to it (the order the
    # PyTorch interop requires) and the tests that use torch tensors can run
    # after the others in one process.
    if config.getoption("markexpr", "").replace(" ", "") != "notgpu":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
