"""The C++ host driver (main.cpp shape) and the C++ shim on the reference's own input
example13.in (50^3 stretched grid, one 10^3 solid block), reference physics, against
the oracle; the output records of hs_get_output against an independent numpy packing
of the reference's rules (solver.h:160-213)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import abi, configs, mesh

from helpers import gpu_ctx, oracle_ctx, pack_reference, rel_l2, small_ref_case

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
EX13 = str(ROOT / "tests" / "golden" / "example13.in")
EXE = str(ROOT / "hsolver_amd" / "hsolver_main")


def read_vtk(path, n):
    lines = Path(path).read_text().splitlines()

    def block(tag, width, start_off=None, conv=float):
        i = next(k for k, l in enumerate(lines) if l.startswith(tag))
        start = i + (start_off if start_off is not None else (2 if tag.startswith("SCALARS") else 1))
        return np.array([[conv(v) for v in l.split()] for l in lines[start:start + n]]).reshape(n, width)

    return {"loc": block("POINTS", 3), "p": block("SCALARS pressure", 1)[:, 0], "vel": block("VECTORS velocity", 3),
            "t": block("SCALARS temperature", 1)[:, 0], "id": block("SCALARS id", 1, conv=int)[:, 0]}


def ex13_oracle(oracle_lib, iters):
    m = mesh.read_in(EX13)
    case = configs.Case("ex13", m["xv"], m["yv"], m["zv"], m["cell_type"], configs.mode_r_params())
    o = oracle_ctx(oracle_lib, case)
    o.init()
    for _ in range(iters):
        o.step()
    return case, o


def test_driver_example13_matches_oracle(oracle_lib, tmp_path):
    """result.vtk in the reference's records and format (12 decimals, boundary t = 293,
    fluid row ids, zeroed solid cells) against the oracle's fields packed by the
    reference's rules; --raw against the oracle's raw fields."""
    iters = 3
    out, raw = tmp_path / "r.vtk", tmp_path / "raw.vtk"
    for path, extra in ((out, []), (raw, ["--raw"])):
        r = subprocess.run([EXE, EX13, str(iters), str(path)] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
        assert "VTK file written" in r.stdout
    nx = ny = nz = 50
    n = nx * ny * nz
    text = out.read_text()
    assert "DIMENSIONS 50 50 50" in text and f"POINTS {n} double" in text
    case, o = ex13_oracle(oracle_lib, iters)
    want = pack_reference(case.grid, o.state())
    got = read_vtk(out, n)
    assert np.array_equal(got["id"], want["id"])
    assert np.abs(got["loc"] - want["loc"]).max() < 1e-12
    # 12 printed decimals: 5e-13 of rounding on top of the GPU-vs-oracle difference
    for key in ("p", "t"):
        assert np.abs(got[key] - want[key]).max() < 1e-9 * max(1.0, np.abs(want[key]).max()), key
    assert np.abs(got["vel"] - want["vel"]).max() < 1e-9
    assert np.all(got["t"][want["id"] < 0] == want["t"][want["id"] < 0]) or \
        np.abs(got["t"] - want["t"]).max() < 1e-9
    solid = want["id"] < 0
    assert solid.sum() == 1000 and np.all(got["p"][solid] == 0) and np.all(got["vel"][solid] == 0)
    # raw fields
    f = o.fields()
    rg = read_vtk(raw, n)
    assert rel_l2(rg["p"], f["p"].ravel()) < 1e-9
    assert rel_l2(rg["t"], f["t"].ravel()) < 1e-12
    assert rel_l2(rg["vel"][:, 0], f["u"][:, :, :nx].ravel()) < 1e-9
    assert rel_l2(rg["vel"][:, 1], f["v"][:, :ny, :].ravel()) < 1e-9
    assert rel_l2(rg["vel"][:, 2], f["w"][:nz, :, :].ravel()) < 1e-9


def test_driver_reads_processor_files(tmp_path):
    # main.cpp:16-21: grid from the .in, cell types from processor0/ (written here by the
    # one-part decomposer): the same run as from the .in alone, byte for byte
    r = subprocess.run([EXE, EX13, "--decompose", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    for out, extra in ((a, []), (b, ["--proc", str(tmp_path)])):
        r = subprocess.run([EXE, EX13, "2", str(out)] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
    assert "cell types from" in r.stdout
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("order", ["lex", "rb"])
def test_shim_reference_loop_matches_abi(order):
    """FluidSolver / MobSolver driven by a host loop in the order of Solver::solve
    (solver.h:44-155), and hsolver::Solver::solve itself, leave every device array bit
    for bit as hs_run does; the shim's result / idArray / idToData hold hs_get_output's
    records (hsolver_shim_check exits 0 only then)."""
    exe = str(ROOT / "hsolver_amd" / "hsolver_shim_check")
    r = subprocess.run([exe, EX13, "3", "--order", order], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "bit-identical" in r.stdout


def test_output_records_mode_r(hs_lib, oracle_lib):
    """hs_get_output against the numpy packing of the same context's raw arrays (exact),
    before any p' assembly (ids 0) and after two SIMPLE iterations, and against the
    oracle's arrays packed the same way."""
    case = small_ref_case()
    g, o = gpu_ctx(hs_lib, case), oracle_ctx(oracle_lib, case)
    g.init(); o.init()
    got = g.output()
    want = pack_reference(case.grid, g.state(), assembled=False)
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    for _ in range(2):
        g.step(); o.step()
    got = g.output()
    want = pack_reference(case.grid, g.state())
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    ref = pack_reference(case.grid, o.state())
    assert np.array_equal(got["id"], ref["id"])
    for key in ("vel", "p", "t"):
        assert rel_l2(got[key], ref[key]) < 1e-9, key
    # the boundary rule: every fluid cell on the domain boundary reads 293
    nx, ny, nz = g.nx, g.ny, g.nz
    edge = np.zeros((nz, ny, nx), bool)
    edge[0], edge[-1], edge[:, 0], edge[:, -1], edge[:, :, 0], edge[:, :, -1] = (True,) * 6
    fl = case.grid.cell_type.reshape(nz, ny, nx) == 0
    assert np.all(got["t"][(edge & fl).ravel()] == 293.0)
    assert np.all(got["id"][~fl.ravel()] == -1)


def test_output_records_mode_c(hs_lib):
    """Mode C: the conjugate temperature in every cell (no boundary override), fluid row
    ids, zero velocity and p on solid cells."""
    case = configs.c2(scale=8)
    case.params.order = abi.HS_ORDER_RB
    g = gpu_ctx(hs_lib, case)
    g.init()
    g.step()
    got = g.output()
    want = pack_reference(case.grid, g.state(), mode_r=False)
    for key in want:
        assert np.array_equal(got[key], want[key]), key


def cylindrical_in(path):
    """A periodic cylindrical-form input (nyr = 90 over 360 -> 90 rows 4 wide, mesh.cpp:
    1644-1700) with one solid block away from the y ends.  (A block split over rows 1 and
    90 is one region to the decomposer, which wraps y, but two to the solver, which keeps
    walls at the y ends as the reference's does; the reference-physics exchange takes one
    solid region, so such a split is refused at hs_create.)"""
    xs = ",".join(str(i) for i in range(17)) + ","
    zs = ",".join(str(i) for i in range(11)) + ","
    path.write_text(f"xgrid = {xs}\n\nyl(1)=0,yc(1)=0,nyl(1)=0,nyr(1)=90,dynm(1)=0,yl(2)=360,\n\n"
                    f"zgrid = {zs}\n\nmobs =\n 6,8,31,33,4,6,\n\nwalls =\n\nholes=\n\nends\n")
    return str(path)


def test_driver_cylindrical_periodic_matches_oracle(oracle_lib, tmp_path):
    src = cylindrical_in(tmp_path / "cyl.in")
    m = mesh.read_in(src)
    assert m["periodic_y"] and m["cell_type"].shape == (10, 90, 16) and m["solid_regions"] == 1
    assert m["cell_type"].sum() == 27 and np.all(np.diff(m["yv"]) == 4)
    iters = 3
    raw = tmp_path / "raw.vtk"
    r = subprocess.run([EXE, src, str(iters), str(raw), "--raw"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "DIMENSIONS 16 90 10" in raw.read_text()
    case = configs.Case("cyl", m["xv"], m["yv"], m["zv"], m["cell_type"], configs.mode_r_params())
    o = oracle_ctx(oracle_lib, case)
    o.init()
    for _ in range(iters):
        o.step()
    f = o.fields()
    assert np.isfinite(f["t"]).all() and np.abs(f["t"]).max() < 400
    n = 16 * 90 * 10
    rg = read_vtk(raw, n)
    assert rel_l2(rg["p"], f["p"].ravel()) < 1e-9
    assert rel_l2(rg["t"], f["t"].ravel()) < 1e-12
    assert rel_l2(rg["vel"][:, 0], f["u"][:, :, :16].ravel()) < 1e-9
    assert rel_l2(rg["vel"][:, 1], f["v"][:, :90, :].ravel()) < 1e-9
    assert rel_l2(rg["vel"][:, 2], f["w"][:10, :, :].ravel()) < 1e-9
    # the decomposer's files for the periodic typing drive the same run
    r = subprocess.run([EXE, src, "--decompose", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    for out, extra in ((a, []), (b, ["--proc", str(tmp_path)])):
        r = subprocess.run([EXE, src, "2", str(out)] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr + r.stdout
    assert a.read_bytes() == b.read_bytes()
