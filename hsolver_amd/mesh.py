"""`.in` input files through the library's host mesh module (include/hsolver_mesh.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, configs


def _decl(lib):
    lib.hs_mesh_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    lib.hs_mesh_parse.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    I = C.POINTER(C.c_int32)
    lib.hs_mesh_info.argtypes = [C.c_void_p, I, I, I, I, I, I, I]
    lib.hs_mesh_grid.argtypes = [C.c_void_p, C.POINTER(abi.Grid)]
    lib.hs_mesh_box.argtypes = [C.c_void_p, C.c_int32, I]
    lib.hs_mesh_free.argtypes = [C.c_void_p]
    lib.hs_mesh_free.restype = None


def read_in(path: str | None = None, text: str | None = None, lib: C.CDLL | None = None) -> dict:
    """Parse a `.in` file (or its text).  Returns vertex arrays, cell types (z,y,x),
    the mobs boxes and the region counts."""
    if lib is None:
        from . import core
        lib = core.library()
    _decl(lib)
    h = C.c_void_p()
    rc = lib.hs_mesh_read(path.encode(), C.byref(h)) if path is not None else lib.hs_mesh_parse(text.encode(), C.byref(h))
    if rc != 0:
        raise OSError(f"hs_mesh_read failed ({rc}) for {path}")
    try:
        v = [C.c_int32() for _ in range(7)]
        lib.hs_mesh_info(h, *[C.byref(x) for x in v])
        nx, ny, nz, nm, nw, nf, ns = (x.value for x in v)
        g = abi.Grid()
        lib.hs_mesh_grid(h, C.byref(g))
        xv = np.ctypeslib.as_array(g.xv, shape=(nx + 1,)).copy()
        yv = np.ctypeslib.as_array(g.yv, shape=(ny + 1,)).copy()
        zv = np.ctypeslib.as_array(g.zv, shape=(nz + 1,)).copy()
        ct = np.ctypeslib.as_array(g.cell_type, shape=(nx * ny * nz,)).copy().reshape(nz, ny, nx)
        boxes = []
        for b in range(nm):
            box = (C.c_int32 * 6)()
            lib.hs_mesh_box(h, b, box)
            boxes.append(tuple(box))
        return {"xv": xv, "yv": yv, "zv": zv, "cell_type": ct, "mobs": boxes, "nwalls": nw,
                "fluid_regions": nf, "solid_regions": ns}
    finally:
        lib.hs_mesh_free(h)


def case_from_in(path: str, params: abi.Params | None = None, lib=None) -> configs.Case:
    m = read_in(path, lib=lib)
    p = params if params is not None else configs.mode_r_params()
    return configs.Case(str(path), m["xv"], m["yv"], m["zv"], m["cell_type"], p,
                        workload=f"{path} ({len(m['xv'])-1}x{len(m['yv'])-1}x{len(m['zv'])-1})")
