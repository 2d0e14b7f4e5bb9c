"""CPU checks of the host driver's command line: a bad iteration count, an unknown flag
or a flag without its value ends with the usage text and exit code 1 before any mesh is
read or device touched."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "hsolver_amd" / "hsolver_main"
EX13 = str(ROOT / "tests" / "golden" / "example13.in")


@pytest.mark.parametrize("args", [[EX13, "0"], [EX13, "-3"], [EX13, "12abc"], [EX13, "--bogus"],
                                  [EX13, "5", "out.vtk", "extra"], [EX13, "--order", "zigzag"],
                                  [EX13, "--order"], []])
def test_usage_errors(args, tmp_path):
    if not EXE.exists():
        pytest.skip("hsolver_main not built")
    r = subprocess.run([str(EXE)] + args, capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "usage:" in r.stderr
    assert not (tmp_path / "out.vtk").exists() and not (tmp_path / "result.vtk").exists()
