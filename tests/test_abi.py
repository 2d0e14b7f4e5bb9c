"""CPU checks of the C ABI: the library loads without a GPU, exports every entry point
the public headers declare, its struct layouts match the Python bindings, its
defaults match, and argument errors come back as codes (no device is touched)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import abi, configs

ROOT = Path(__file__).resolve().parents[1]


def declared(header: Path):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(hs_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(hs_lib):
    names = declared(ROOT / "include" / "hsolver.h") + declared(ROOT / "include" / "hsolver_mesh.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(hs_lib, n)]
    assert not missing, missing


def test_struct_sizes_match_bindings(hs_lib):
    s = (C.c_int32 * 3)()
    assert hs_lib.hs_abi_sizes(s) == 0
    assert list(s) == [C.sizeof(abi.Grid), C.sizeof(abi.Params), C.sizeof(abi.Stats)]


@pytest.mark.parametrize("mode", [abi.HS_MODE_REF, abi.HS_MODE_CONFIG])
def test_params_default_matches(hs_lib, mode):
    p = abi.Params()
    hs_lib.hs_params_default(mode, C.byref(p))
    q = abi.params_default(mode)
    for name, _ in abi.Params._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if name == "reserved":
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_reference_constants_in_mode_r_defaults():
    # fluid_solver.h:21-37, mobs_solver.h:15-29, fluid_solver.h:1871-1875, solver.h:52
    p = abi.params_default(abi.HS_MODE_REF)
    assert (p.u_lid, p.nu, p.alpha_f, p.rho) == (0.1, 0.001, 0.001, 1.0)
    assert (p.relax_u, p.relax_p, p.relax_t, p.relax_ts) == (0.3, 0.3, 0.3, 0.3)
    assert (p.mom_tol, p.t_tol, p.ts_tol) == (2e-3, 2e-3, 1e-3)
    assert (p.mom_sweeps_max, p.t_sweeps_max, p.ts_sweeps_max) == (100, 100, 100)
    assert (p.p_tol, p.p_maxit, p.freeze_p_matrix) == (1e-5, 1000, 1)
    assert (p.t_fluid_init, p.t_solid_init) == (293.0, 333.0)


def test_strerror_and_argument_errors(hs_lib):
    assert hs_lib.hs_strerror(0) == b"ok"
    assert b"converge" in hs_lib.hs_strerror(abi.HS_E_NOCONV)
    h = C.c_void_p()
    p = abi.params_default(abi.HS_MODE_REF)
    g = abi.GridArrays(np.linspace(0, 1, 3), np.linspace(0, 1, 5), np.linspace(0, 1, 5), np.zeros((4, 4, 2), np.uint8))
    assert hs_lib.hs_create(C.byref(g.struct), C.byref(p), C.byref(h)) == abi.HS_E_ARG   # nx < 3
    assert hs_lib.hs_create(None, C.byref(p), C.byref(h)) == abi.HS_E_ARG
    bad = abi.params_default(abi.HS_MODE_REF)
    bad.mode = 7
    g2 = abi.GridArrays(np.linspace(0, 1, 5), np.linspace(0, 1, 5), np.linspace(0, 1, 5), np.zeros((4, 4, 4), np.uint8))
    assert hs_lib.hs_create(C.byref(g2.struct), C.byref(bad), C.byref(h)) == abi.HS_E_ARG


def test_mode_r_rejects_two_solid_regions(hs_lib):
    ct = configs.box_cells(8, 8, 8, [(2, 3, 2, 3, 2, 3), (6, 7, 6, 7, 6, 7)])
    g = abi.GridArrays(np.linspace(0, 1, 9), np.linspace(0, 1, 9), np.linspace(0, 1, 9), ct)
    p = abi.params_default(abi.HS_MODE_REF)
    h = C.c_void_p()
    assert hs_lib.hs_create(C.byref(g.struct), C.byref(p), C.byref(h)) == abi.HS_E_ARG


@pytest.mark.parametrize("shape", [(24, 12, 1), (40, 1, 1), (2, 8, 8)])
def test_fewer_than_three_cells_refused(hs_lib, shape):
    """hs_create refuses a grid with fewer than three cells along an axis before it touches
    the device (HS_E_ARG), in both physics modes."""
    import numpy as np
    from hsolver_amd import configs
    nx, ny, nz = shape
    grid = abi.GridArrays(configs.uniform_axis(nx, 1.0), configs.uniform_axis(ny, 1.0), configs.uniform_axis(nz, 1.0),
                          np.zeros(nx * ny * nz, dtype=np.uint8))
    for mode in (abi.HS_MODE_REF, abi.HS_MODE_CONFIG):
        p = abi.params_default(mode)
        h = C.c_void_p()
        assert hs_lib.hs_create(C.byref(grid.struct), C.byref(p), C.byref(h)) == abi.HS_E_ARG
        assert not h.value
