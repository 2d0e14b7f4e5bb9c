"""`.in` input files and one rank's processor files (processor0/cells.txt, faces.txt,
boundary.txt) through the library's host mesh module (include/hsolver_mesh.h)."""
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
    lib.hs_mesh_y_period.argtypes = [C.c_void_p, I, I]
    lib.hs_mesh_free.argtypes = [C.c_void_p]
    lib.hs_mesh_free.restype = None
    lib.hs_mesh_write_proc.argtypes = [C.c_void_p, C.c_char_p]
    lib.hs_mesh_read_proc.argtypes = [C.c_void_p, C.c_char_p]
    L = C.POINTER(C.c_int64)
    lib.hs_mesh_proc_part.argtypes = [C.c_void_p, C.c_int32, I, I, L, L, I]
    lib.hs_mesh_proc_boundary.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32, L, L]


HS_E_ARG, HS_E_IO = abi.HS_E_ARG, abi.HS_E_IO


def _open(path, text, lib):
    if lib is None:
        from . import core
        lib = core.library()
    _decl(lib)
    h = C.c_void_p()
    rc = lib.hs_mesh_read(path.encode(), C.byref(h)) if path is not None else lib.hs_mesh_parse(text.encode(), C.byref(h))
    if rc != 0:
        raise OSError(f"hs_mesh_read failed ({rc}) for {path}")
    return lib, h


def _summary(lib, h) -> dict:
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
    span, periodic = C.c_int32(), C.c_int32()
    lib.hs_mesh_y_period(h, C.byref(span), C.byref(periodic))
    return {"xv": xv, "yv": yv, "zv": zv, "cell_type": ct, "mobs": boxes, "nwalls": nw,
            "fluid_regions": nf, "solid_regions": ns, "y_span": span.value, "periodic_y": bool(periodic.value)}


def _parts(lib, h) -> dict:
    out = {}
    for which, name in ((0, "fluid"), (1, "mobs")):
        region, group, nb = C.c_int32(), C.c_int32(), C.c_int32()
        nc, nfc = C.c_int64(), C.c_int64()
        rc = lib.hs_mesh_proc_part(h, which, C.byref(region), C.byref(group), C.byref(nc), C.byref(nfc), C.byref(nb))
        if rc != 0:
            raise ValueError(f"hs_mesh_proc_part failed ({rc}) for {name}")
        bounds = []
        for b in range(nb.value):
            buf = C.create_string_buffer(64)
            st, n = C.c_int64(), C.c_int64()
            lib.hs_mesh_proc_boundary(h, which, b, buf, 64, C.byref(st), C.byref(n))
            bounds.append((buf.value.decode(), st.value, n.value))
        out[name] = {"region": region.value, "group": group.value, "ncells": nc.value, "nfaces": nfc.value,
                     "bounds": bounds}
    return out


def read_in(path: str | None = None, text: str | None = None, lib: C.CDLL | None = None) -> dict:
    """Parse a `.in` file (or its text).  Returns vertex arrays, cell types (z,y,x),
    the mobs boxes, the region counts and the cylindrical y form's span and periodicity
    (`y_span` 0 and `periodic_y` False for a ygrid file)."""
    lib, h = _open(path, text, lib)
    try:
        return _summary(lib, h)
    finally:
        lib.hs_mesh_free(h)


def write_proc(directory: str, path: str | None = None, text: str | None = None, lib: C.CDLL | None = None) -> dict:
    """The reference's decomposer at one part: write `directory`/processor0/ for a `.in`
    file (or its text).  Returns the per-class summary (cells, faces, boundary groups).
    ValueError when the typing has more than one fluid or solid region."""
    lib, h = _open(path, text, lib)
    try:
        rc = lib.hs_mesh_write_proc(h, str(directory).encode())
        if rc == HS_E_ARG:
            raise ValueError("one part holds one fluid and one solid region at most")
        if rc != 0:
            raise OSError(f"hs_mesh_write_proc failed ({rc}) in {directory}")
        return _parts(lib, h)
    finally:
        lib.hs_mesh_free(h)


def read_proc(directory: str, path: str | None = None, text: str | None = None, lib: C.CDLL | None = None) -> dict:
    """The solver-side reader (main.cpp:18-21): grid from the `.in`, cell types from
    `directory`/processor0/.  Returns read_in's fields with the files' typing, plus
    "parts" (the per-class summary as read).  OSError when the files are missing,
    malformed or do not cover the grid."""
    lib, h = _open(path, text, lib)
    try:
        rc = lib.hs_mesh_read_proc(h, str(directory).encode())
        if rc != 0:
            raise OSError(f"hs_mesh_read_proc failed ({rc}) for {directory}/processor0")
        out = _summary(lib, h)
        out["parts"] = _parts(lib, h)
        return out
    finally:
        lib.hs_mesh_free(h)


def case_from_in(path: str, params: abi.Params | None = None, lib=None) -> configs.Case:
    m = read_in(path, lib=lib)
    p = params if params is not None else configs.mode_r_params()
    return configs.Case(str(path), m["xv"], m["yv"], m["zv"], m["cell_type"], p,
                        workload=f"{path} ({len(m['xv'])-1}x{len(m['yv'])-1}x{len(m['zv'])-1})")
