"""The C++ host driver (main.cpp shape) on the reference's own input example13.in
(50^3 stretched grid, one 10^3 solid block), reference physics, against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import configs, mesh

from helpers import oracle_ctx, rel_l2

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def read_vtk(path, n):
    lines = Path(path).read_text().splitlines()
    def block(tag, width):
        i = next(k for k, l in enumerate(lines) if l.startswith(tag))
        start = i + (2 if tag.startswith("SCALARS") else 1)
        return np.array([[float(v) for v in l.split()] for l in lines[start:start + n]]).reshape(n, width)
    return block("SCALARS pressure", 1)[:, 0], block("VECTORS velocity", 3), block("SCALARS temperature", 1)[:, 0]


def test_driver_example13_matches_oracle(oracle_lib, tmp_path):
    iters = 3
    out = tmp_path / "r.vtk"
    r = subprocess.run([str(ROOT / "hsolver_amd" / "hsolver_main"), str(ROOT / "tests" / "golden" / "example13.in"),
                        str(iters), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "VTK file written" in r.stdout
    m = mesh.read_in(str(ROOT / "tests" / "golden" / "example13.in"))
    nx, ny, nz = 50, 50, 50
    text = out.read_text()
    assert "DIMENSIONS 50 50 50" in text
    p, vel, t = read_vtk(out, nx * ny * nz)
    case = configs.Case("ex13", m["xv"], m["yv"], m["zv"], m["cell_type"], configs.mode_r_params())
    o = oracle_ctx(oracle_lib, case)
    o.init()
    for _ in range(iters):
        o.step()
    f = o.fields()
    assert rel_l2(p, f["p"].ravel()) < 1e-9
    assert rel_l2(t, f["t"].ravel()) < 1e-12
    assert rel_l2(vel[:, 0], f["u"][:, :, :nx].ravel()) < 1e-9
    assert rel_l2(vel[:, 1], f["v"][:, :ny, :].ravel()) < 1e-9
    assert rel_l2(vel[:, 2], f["w"][:nz, :, :].ravel()) < 1e-9


def test_driver_reads_processor_files(tmp_path):
    # main.cpp:16-21: grid from the .in, cell types from processor0/ (written here by the
    # one-part decomposer): the same run as from the .in alone, byte for byte
    exe = str(ROOT / "hsolver_amd" / "hsolver_main")
    ex13 = str(ROOT / "tests" / "golden" / "example13.in")
    r = subprocess.run([exe, ex13, "--decompose", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    for out, extra in ((a, []), (b, ["--proc", str(tmp_path)])):
        r = subprocess.run([exe, ex13, "2", str(out)] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
    assert "cell types from" in r.stdout
    assert a.read_bytes() == b.read_bytes()
