import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(ROOT))

_NATIVE = [ROOT / "spice2_amd" / "lib" / "libspice_hip.so",
           ROOT / "spice2_amd" / "lib" / "libspice_workloads.so",
           ROOT / "oracle" / "liboracle.so",
           ROOT / "tests" / "native" / "bin" / "api_host_tests"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long check at the N=500k scale")
    if not all(p.exists() for p in _NATIVE):  # build once if nothing is built yet
        import __graft_entry__
        __graft_entry__.build()
