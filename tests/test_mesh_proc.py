"""One rank's processor files: the decomposer at one part (decomposeMesh mesh.cpp:735-1581
with nParts = 1, writeProcFile :2277-2374) and the solver-side reader (readCells /
readFaces / readBoundaries :2507-2642), on the CPU.

Pinned by: files worked out by hand from those functions for two- and three-cell grids,
closed-form face counts for solid boxes inside and on the edge of a box domain, the
geometry of every face record of the reference's example13.in, a byte-exact
write -> read -> write round trip, and refusal of malformed, partial and multi-region
inputs."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import mesh

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def grid_text(nx, ny, nz, mobs=(), walls=()):
    pts = lambda n: ",".join(str(v) for v in range(n + 1)) + ","
    lines = [f"xgrid = {pts(nx)}", f"ygrid = {pts(ny)}", f"zgrid = {pts(nz)}", "mobs ="]
    lines += [" " + ",".join(str(v) for v in b) + "," for b in mobs]
    lines.append("walls =")
    lines += [" " + ",".join(str(v) for v in w) + "," for w in walls]
    lines.append("holes =")
    return "\n".join(lines) + "\n"


def files(d):
    return {f: (Path(d) / "processor0" / f"{f}.txt").read_text() for f in ("cells", "faces", "boundary")}


def parse(d):
    """The three files as {type: (cells, faces[n,3], [(name, start, n)])}, parsed here."""
    out = {}
    cl = files(d)["cells"].splitlines()
    for a, line in enumerate(cl):
        if line.startswith("meta"):
            _, t, n, _, _ = line.split()
            out[t] = [np.array(cl[a + 1].split(), dtype=np.int64), None, []]
            assert len(out[t][0]) == int(n)
    fl = files(d)["faces"].splitlines()
    for a, line in enumerate(fl):
        if line.startswith("meta"):
            _, t, n, _, _ = line.split()
            rows = [r.split() for r in fl[a + 1:a + 1 + int(n)]]
            out[t][1] = np.array(rows, dtype=np.int64).reshape(-1, 3)
    bl = files(d)["boundary"].splitlines()
    for a, line in enumerate(bl):
        if line.startswith("meta"):
            _, t, n, _, _ = line.split()
            out[t][2] = [(r.split()[0], int(r.split()[1]), int(r.split()[2])) for r in bl[a + 1:a + 1 + int(n)]]
    return out


def cell_faces(c, nx, ny, nz):
    """Global face ids of 1-based cells c (mesh.h:209-219, 245-253): x faces, then y, then z."""
    q = np.asarray(c) - 1
    i, j, k = q % nx, (q // nx) % ny, q // (nx * ny)
    nxf, nyf = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    fx = lambda i, j, k: i + (nx + 1) * (j + ny * k)
    fy = lambda i, j, k: nxf + i + nx * (j + (ny + 1) * k)
    fz = lambda i, j, k: nxf + nyf + i + nx * (j + ny * k)
    return np.stack([fx(i, j, k), fx(i + 1, j, k), fy(i, j, k), fy(i, j + 1, k), fz(i, j, k), fz(i, j, k + 1)], -1)


def test_two_cells_by_hand(hs_lib, tmp_path):
    # 2x1x1, cell 2 solid.  Face ids: x 0-2, y 3-6, z 7-10.  Each cell: its face to the
    # other class in "ib_0_0", its five outer faces in "wall"; groups in name order.
    mesh.write_proc(tmp_path, text=grid_text(2, 1, 1, mobs=[(2, 2, 1, 1, 1, 1)]), lib=hs_lib)
    f = files(tmp_path)
    assert f["cells"] == "meta fluid 1 0 0\n1 \nmeta mobs 1 0 0\n2 \n"
    assert f["faces"] == ("meta fluid 6 0 0\n1 0 0\n0 0 0\n3 0 0\n5 0 0\n7 0 0\n9 0 0\n\n"
                          "meta mobs 6 0 0\n1 0 0\n2 0 0\n4 0 0\n6 0 0\n8 0 0\n10 0 0\n\n")
    assert f["boundary"] == ("meta fluid 2 0 0\nib_0_0 0 1\nwall 1 5\n\n"
                             "meta mobs 2 0 0\nib_0_0 0 1\nwall 1 5\n\n")


def test_three_cell_line_by_hand(hs_lib, tmp_path):
    # 3x1x1 fluid: internal faces 1 (cells 0|1) and 2 (cells 1|2) owned by the earlier
    # cell; 14 outer faces in id order, each owned by its only cell, neighbour 0
    mesh.write_proc(tmp_path, text=grid_text(3, 1, 1), lib=hs_lib)
    f = files(tmp_path)
    assert f["cells"] == "meta fluid 3 0 0\n1 2 3 \n"
    owners = {0: 0, 3: 2, 4: 0, 5: 1, 6: 2, 7: 0, 8: 1, 9: 2, 10: 0, 11: 1, 12: 2, 13: 0, 14: 1, 15: 2}
    body = "".join(f"{g} {o} 0\n" for g, o in owners.items())
    assert f["faces"] == "meta fluid 16 0 0\n1 0 1\n2 1 2\n" + body + "\n"
    assert f["boundary"] == "meta fluid 1 0 0\nwall 2 14\n\n"


def test_interior_wall_face_keeps_later_owner(hs_lib, tmp_path):
    # 3x2x1 with a wall on x face i = 2 at j = 1 only: the region stays connected through
    # j = 2.  Both cells list the face in "wall"; the group keeps the later one (cell 2,
    # position 1) and neighbour 0, and the face is not internal.
    parts = mesh.write_proc(tmp_path, text=grid_text(3, 2, 1, walls=[(2, 2, 1, 1, 1, 1, 1, 0)]), lib=hs_lib)
    p = parse(tmp_path)["fluid"]
    name, start, n = p[2][0]
    assert name == "wall"
    wall = p[1][start:start + n]
    assert 1 in wall[:, 0] and 1 not in p[1][:start, 0]
    assert tuple(wall[wall[:, 0] == 1][0]) == (1, 1, 0)
    # 3x2x1 has 4 + 3 + 0 = 7 internal faces; the wall takes one
    assert start == 6 and n == 2 * (2 + 3 + 6) + 1
    assert parts["fluid"]["bounds"] == [("wall", 6, 23)]


def box_counts(nx, ny, nz, box):
    """Face counts of a one-part decomposition with one solid box, in closed form."""
    i0, i1, j0, j1, k0, k1 = box
    a, b, c = i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1
    total_int = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    total_out = 2 * (ny * nz + nx * nz + nx * ny)
    m_int = (a - 1) * b * c + a * (b - 1) * c + a * b * (c - 1)
    m_wall = (b * c * ((i0 == 1) + (i1 == nx)) + a * c * ((j0 == 1) + (j1 == ny)) + a * b * ((k0 == 1) + (k1 == nz)))
    ib = 2 * (a * b + b * c + c * a) - m_wall
    f_int = total_int - m_int - ib
    fluid = [("ib_0_0", f_int, ib), ("wall", f_int + ib, total_out - m_wall)]
    mobs = [("ib_0_0", m_int, ib)] + ([("wall", m_int + ib, m_wall)] if m_wall else [])
    return nx * ny * nz - a * b * c, a * b * c, fluid, mobs


@pytest.mark.parametrize("dims,box", [((6, 5, 4), (2, 3, 2, 4, 2, 3)),     # inside
                                      ((6, 5, 4), (1, 2, 2, 3, 3, 4)),     # on x-min and z-max
                                      ((6, 5, 4), (2, 3, 1, 5, 1, 2)),     # across y, on z-min
                                      ((7, 3, 5), (7, 7, 1, 1, 5, 5))])    # one corner cell
def test_solid_box_closed_form(hs_lib, tmp_path, dims, box):
    parts = mesh.write_proc(tmp_path, text=grid_text(*dims, mobs=[box]), lib=hs_lib)
    nf, nm, fb, mb = box_counts(*dims, box)
    assert parts["fluid"]["ncells"] == nf and parts["mobs"]["ncells"] == nm
    assert parts["fluid"]["bounds"] == fb and parts["mobs"]["bounds"] == mb
    assert parts["fluid"]["nfaces"] == fb[-1][1] + fb[-1][2]


def test_all_solid_and_all_fluid(hs_lib, tmp_path):
    (tmp_path / "s").mkdir()
    parts = mesh.write_proc(tmp_path / "s", text=grid_text(3, 4, 5, mobs=[(1, 3, 1, 4, 1, 5)]), lib=hs_lib)
    n_int = 2 * 4 * 5 + 3 * 3 * 5 + 3 * 4 * 4
    assert parts["fluid"]["ncells"] == 0 and parts["fluid"]["bounds"] == []
    assert parts["mobs"]["bounds"] == [("wall", n_int, 2 * (20 + 15 + 12))]
    assert files(tmp_path / "s")["cells"].startswith("meta mobs 60 0 0\n")
    parts = mesh.write_proc(tmp_path, text=grid_text(3, 4, 5), lib=hs_lib)
    assert parts["mobs"]["ncells"] == 0
    assert "mobs" not in files(tmp_path)["faces"]


def test_example13_face_geometry(hs_lib, tmp_path):
    parts = mesh.write_proc(tmp_path, path=str(GOLDEN / "example13.in"), lib=hs_lib)
    nf, nm, fb, mb = box_counts(50, 50, 50, (21, 30, 21, 30, 21, 30))
    assert (parts["fluid"]["bounds"], parts["mobs"]["bounds"]) == (fb, mb)
    m = mesh.read_in(str(GOLDEN / "example13.in"), lib=hs_lib)
    ct = m["cell_type"].ravel()
    got = parse(tmp_path)
    for t, cls in (("fluid", 0), ("mobs", 1)):
        cells, faces, bounds = got[t]
        assert np.array_equal(cells, np.flatnonzero(ct == cls) + 1)          # global order
        cf = cell_faces(cells, 50, 50, 50)
        n_int = bounds[0][1]
        own = cf[faces[:, 1]]
        assert (own == faces[:, 0:1]).any(1).all()                            # a face of its owner
        nb = cf[faces[:n_int, 2]]
        assert (nb == faces[:n_int, 0:1]).any(1).all()                        # and of its neighbour
        assert (faces[:n_int, 1] < faces[:n_int, 2]).all()
        assert (faces[n_int:, 2] == 0).all()
        assert [b[0] for b in bounds] == sorted(b[0] for b in bounds)
        for name, s, n in bounds:
            g = faces[s:s + n, 0]
            assert (np.diff(g) > 0).all()
            # each interface face is a face of a cell of the other class
            if name == "ib_0_0":
                other = cell_faces(np.flatnonzero(ct != cls) + 1, 50, 50, 50).ravel()
                assert np.isin(g, other).all()


def _roundtrip(lib, text, src, dst):
    mesh._decl(lib)
    h = C.c_void_p()
    assert lib.hs_mesh_parse(text.encode(), C.byref(h)) == 0
    try:
        rc = lib.hs_mesh_read_proc(h, str(src).encode())
        if rc == 0:
            assert lib.hs_mesh_write_proc(h, str(dst).encode()) == 0
        return rc
    finally:
        lib.hs_mesh_free(h)


def test_write_read_write_is_byte_exact(hs_lib, tmp_path):
    text = (GOLDEN / "example13.in").read_text()
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir(), b.mkdir()
    mesh.write_proc(a, text=text, lib=hs_lib)
    assert _roundtrip(hs_lib, text, a, b) == 0
    assert files(a) == files(b)
    r = mesh.read_proc(a, text=text, lib=hs_lib)
    assert np.array_equal(r["cell_type"], mesh.read_in(text=text, lib=hs_lib)["cell_type"])
    assert (r["fluid_regions"], r["solid_regions"]) == (1, 1)
    assert r["parts"] == mesh.write_proc(b, text=text, lib=hs_lib)


def test_cell_types_come_from_the_files(hs_lib, tmp_path):
    # the solver's typing is the processor files', not the .in's mobs boxes (main.cpp:18-21)
    mesh.write_proc(tmp_path, text=grid_text(6, 5, 4, mobs=[(2, 3, 2, 4, 2, 3)]), lib=hs_lib)
    r = mesh.read_proc(tmp_path, text=grid_text(6, 5, 4), lib=hs_lib)
    assert r["cell_type"].sum() == 12 and r["cell_type"][1:3, 1:4, 1:3].all()
    assert r["solid_regions"] == 1


def _corrupt(d, name, old, new, count=1):
    p = Path(d) / "processor0" / f"{name}.txt"
    t = p.read_text()
    assert old in t
    p.write_text(t.replace(old, new, count))


@pytest.mark.parametrize("case", ["missing", "cell_dropped", "cell_twice", "face_not_of_owner", "face_id_range",
                                  "group_past_end", "bad_meta", "region_mismatch", "half_of_two_parts"])
def test_reader_refuses(hs_lib, tmp_path, case):
    text = grid_text(6, 5, 4, mobs=[(2, 3, 2, 4, 2, 3)])
    src = tmp_path / "src"
    src.mkdir()
    if case != "missing":
        mesh.write_proc(src, text=text, lib=hs_lib)
    if case == "cell_dropped":
        _corrupt(src, "cells", "meta mobs 12 0 0\n", "meta mobs 11 0 0\n")
        p = src / "processor0" / "cells.txt"
        lines = p.read_text().split("\n")
        lines[3] = " ".join(lines[3].split()[1:]) + " "
        p.write_text("\n".join(lines))
    elif case == "cell_twice":
        p = src / "processor0" / "cells.txt"
        lines = p.read_text().split("\n")
        first_solid = lines[3].split()[0]
        lines[1] = lines[1].replace("1 ", first_solid + " ", 1)
        p.write_text("\n".join(lines))
    elif case == "face_not_of_owner":
        p = src / "processor0" / "faces.txt"
        lines = p.read_text().split("\n")
        g, o, n = lines[1].split()
        lines[1] = f"{g} {int(o) + 7} {n}"
        p.write_text("\n".join(lines))
    elif case == "face_id_range":
        p = src / "processor0" / "faces.txt"
        lines = p.read_text().split("\n")
        lines[1] = "999999 " + " ".join(lines[1].split()[1:])
        p.write_text("\n".join(lines))
    elif case == "group_past_end":
        _corrupt(src, "boundary", "wall ", "wall 1", 1)
    elif case == "bad_meta":
        _corrupt(src, "faces", "meta fluid", "meta fluid x")
    elif case == "region_mismatch":
        _corrupt(src, "boundary", "meta mobs 1 0 0", "meta mobs 1 1 0")
    elif case == "half_of_two_parts":
        # the files of one rank of a two-part run hold part of the grid
        p = src / "processor0" / "cells.txt"
        lines = p.read_text().split("\n")
        fluid = lines[1].split()
        lines[0] = f"meta fluid {len(fluid) // 2} 0 0"
        lines[1] = " ".join(fluid[:len(fluid) // 2]) + " "
        p.write_text("\n".join(lines[:2]) + "\n")
        p = src / "processor0" / "faces.txt"
        p.write_text("meta fluid 0 0 0\n\n")
        p = src / "processor0" / "boundary.txt"
        p.write_text("meta fluid 0 0 0\n\n")
    with pytest.raises(OSError):
        mesh.read_proc(src, text=text, lib=hs_lib)


@pytest.mark.parametrize("text", [grid_text(6, 5, 4, mobs=[(2, 2, 2, 2, 2, 2), (4, 4, 3, 3, 3, 3)]),
                                  grid_text(4, 2, 2, walls=[(3, 3, 1, 2, 1, 2, 1, 0)])])
def test_one_part_refuses_several_regions(hs_lib, tmp_path, text):
    with pytest.raises(ValueError):
        mesh.write_proc(tmp_path, text=text, lib=hs_lib)
    assert not (tmp_path / "processor0").exists()


def test_driver_decompose_then_read(tmp_path):
    exe = ROOT / "hsolver_amd" / "hsolver_main"
    r = subprocess.run([str(exe), str(GOLDEN / "example13.in"), "--decompose", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert files(tmp_path)["boundary"] == ("meta fluid 2 0 0\nib_0_0 364200 600\nwall 364800 15000\n\n"
                                           "meta mobs 1 0 0\nib_0_0 2700 600\n\n")


def test_grid_view_sees_types_read_from_files(hs_lib, tmp_path):
    from hsolver_amd import abi
    mesh.write_proc(tmp_path, text=grid_text(6, 5, 4, mobs=[(2, 3, 2, 4, 2, 3)]), lib=hs_lib)
    h = C.c_void_p()
    assert hs_lib.hs_mesh_parse(grid_text(6, 5, 4).encode(), C.byref(h)) == 0
    try:
        g = abi.Grid()
        hs_lib.hs_mesh_grid(h, C.byref(g))             # view taken before the read
        assert hs_lib.hs_mesh_read_proc(h, str(tmp_path).encode()) == 0
        ct = np.ctypeslib.as_array(g.cell_type, shape=(120,))
        assert ct.sum() == 12
    finally:
        hs_lib.hs_mesh_free(h)
