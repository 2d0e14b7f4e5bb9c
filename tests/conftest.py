import ctypes as C
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

ORACLE_SO = ROOT / "oracle" / "liboracle.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _load_oracle():
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    from hsolver_amd import abi
    lib = C.CDLL(str(ORACLE_SO))
    abi.declare(lib, "or")
    return lib


@pytest.fixture(scope="session")
def oracle_lib():
    return _load_oracle()


@pytest.fixture(scope="session")
def hs_lib():
    # torch first: its bundled HIP runtime then serves the library too (one runtime per
    # process; loading the library first maps a second runtime that torch.cuda cannot use)
    import torch  # noqa: F401
    from hsolver_amd import core
    return core.library()
