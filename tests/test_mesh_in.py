"""Host mesh module: the reference's `.in` grammar (mesh.cpp:1616-1893, 2384-2488) and
single-partition typing (mesh.cpp:84-114, 298-454) on its own input files."""
from pathlib import Path

import numpy as np
import pytest

from hsolver_amd import mesh

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_example13(hs_lib):
    m = mesh.read_in(str(GOLDEN / "example13.in"), lib=hs_lib)
    assert (len(m["xv"]), len(m["yv"]), len(m["zv"])) == (51, 51, 51)
    assert m["mobs"] == [(21, 30, 21, 30, 21, 30)]           # the only uncommented mobs line
    assert m["cell_type"].sum() == 1000
    assert m["cell_type"][20:30, 20:30, 20:30].all()
    assert (m["fluid_regions"], m["solid_regions"], m["nwalls"]) == (1, 1, 0)
    assert m["xv"][1] == 0.003 and m["xv"][-1] == 1.0
    assert np.all(np.diff(m["xv"]) > 0)


def test_example21(hs_lib):
    m = mesh.read_in(str(GOLDEN / "example21.in"), lib=hs_lib)
    assert m["cell_type"].shape == (120, 120, 120)
    assert m["mobs"] == [(46, 75, 46, 75, 46, 75)]
    assert m["cell_type"].sum() == 27000
    assert m["fluid_regions"] == 1 and m["solid_regions"] == 1


def test_grammar_comments_and_trailing_separator(hs_lib):
    text = ("xgrid = 0.0, 0.5,\n 1.0, 1.5 ; comment 9.9,\n"
            "ygrid = 0, 1, 2, 3,\nzgrid = 0,1,2,3,\n"
            "mobs =\n ;1,1,1,1,1,1,\n 2,2,2,3,1,3,\n"
            "walls = 2,2,1,3,1,3,1,0,\n holes= \nends")
    m = mesh.read_in(text=text, lib=hs_lib)
    # "1.5" has no trailing separator: the reference's `>> value >> sep` drops it (Q24)
    assert list(m["xv"]) == [0.0, 0.5, 1.0]
    assert list(m["yv"]) == [0, 1, 2, 3] and list(m["zv"]) == [0, 1, 2, 3]
    assert m["mobs"] == [(2, 2, 2, 3, 1, 3)]
    assert m["nwalls"] == 1
    ct = m["cell_type"]
    assert ct.shape == (3, 3, 2) and ct[:, 1:3, 1].all() and ct.sum() == 6


def test_wall_splits_fluid_region(hs_lib):
    text = ("xgrid = 0,1,2,3,4,\nygrid = 0,1,2,\nzgrid = 0,1,2,\nmobs =\n"
            "walls = 3,3,1,2,1,2,1,0,\nholes =\n")
    m = mesh.read_in(text=text, lib=hs_lib)
    assert m["fluid_regions"] == 2 and m["solid_regions"] == 0


def test_missing_file(hs_lib):
    with pytest.raises(OSError):
        mesh.read_in(str(GOLDEN / "does_not_exist.in"), lib=hs_lib)


def _cyl(nyr, yl1, yl2, nx=2, nz=2, mobs="", lead=""):
    xs = ",".join(str(i) for i in range(nx + 1)) + ","
    zs = ",".join(str(i) for i in range(nz + 1)) + ","
    return (f"xgrid = {xs}\n{lead}yl(1)={yl1},yc(1)=0,nyl(1)=0,nyr(1)={nyr},dynm(1)=0,yl(2)={yl2},\n"
            f"zgrid = {zs}\nmobs =\n{mobs}\nwalls =\nholes =\nends\n")


def test_cylindrical_y_form(hs_lib):
    """mesh.cpp:1644-1700: nyr points (b - a) / n * i from 0, in integer arithmetic;
    b - a = 360 gives n periodic cells (setDims mesh.cpp:66-69)."""
    m = mesh.read_in(text=_cyl(72, 0, 360), lib=hs_lib)
    assert m["y_span"] == 360 and m["periodic_y"]
    assert list(m["xv"]) == [0, 1, 2] and list(m["zv"]) == [0, 1, 2]
    # 72 cells: the 72 points 0, 5, ..., 355 and one more so row 72 is 5 wide
    assert list(m["yv"]) == [5.0 * i for i in range(73)]
    assert m["cell_type"].shape == (2, 72, 2)
    # the step truncates (360 / 7 = 51) and the points start at 0 whatever yl(1) is
    m = mesh.read_in(text=_cyl(7, 30, 390), lib=hs_lib)
    assert m["periodic_y"] and list(m["yv"]) == [51.0 * i for i in range(8)]
    # a span other than 360 is an ordinary walled y direction of n - 1 cells
    m = mesh.read_in(text=_cyl(4, 0, 180), lib=hs_lib)
    assert m["y_span"] == 180 and not m["periodic_y"]
    assert list(m["yv"]) == [0, 45, 90, 135]
    # the ygrid form reports no span
    m = mesh.read_in(str(GOLDEN / "example13.in"), lib=hs_lib)
    assert m["y_span"] == 0 and not m["periodic_y"]


def test_cylindrical_values_read_as_stoi(hs_lib):
    """Each value is the text from the key to the next comma read by std::stoi, and the
    keys are found as substrings anywhere in the file (mesh.cpp:1682-1694)."""
    # "nyl(1)=" holds "yl(1)=", so a nyl(1) written first supplies yl(1)
    m = mesh.read_in(text=_cyl(4, 0, 360, lead="nyl(1)=40,"), lib=hs_lib)
    assert m["y_span"] == 320 and not m["periodic_y"] and list(m["yv"]) == [0, 80, 160, 240]
    # leading blanks and a fractional tail are what stoi skips and stops at
    m = mesh.read_in(text=_cyl(" 4", "0.9", "360.5"), lib=hs_lib)
    assert m["y_span"] == 360 and len(m["yv"]) == 5
    for bad in (_cyl(0, 0, 360), _cyl("x", 0, 360), _cyl(4, 0, "end")):
        with pytest.raises(OSError):
            mesh.read_in(text=bad, lib=hs_lib)
    with pytest.raises(OSError):  # yl(2) missing
        mesh.read_in(text="xgrid = 0,1,\nyl(1)=0,nyr(1)=4,\nzgrid = 0,1,\nmobs =\n", lib=hs_lib)


def test_periodic_regions_wrap(hs_lib):
    """getConnectedRegions mesh.cpp:322-327: with periodic y, row 1 neighbours row n, so
    solid boxes at j = 1 and j = n are one region; with a 180 span they are two."""
    boxes = "1,1,1,1,1,2,\n1,1,4,4,1,2,\n"
    m = mesh.read_in(text=_cyl(4, 0, 360, mobs=boxes), lib=hs_lib)
    assert m["cell_type"].sum() == 4
    assert (m["fluid_regions"], m["solid_regions"]) == (1, 1)
    m = mesh.read_in(text=_cyl(5, 0, 180, mobs=boxes), lib=hs_lib)
    assert (m["fluid_regions"], m["solid_regions"]) == (1, 2)
    # three rows with row 2 solid: fluid rows 1 and 3 meet across the wrap; walled, they
    # are two regions
    cut = "1,2,2,2,1,2,\n"
    m = mesh.read_in(text=_cyl(3, 0, 360, mobs=cut), lib=hs_lib)
    assert (m["fluid_regions"], m["solid_regions"]) == (1, 1)
    m = mesh.read_in(text=_cyl(4, 0, 180, mobs=cut), lib=hs_lib)
    assert m["cell_type"].shape == (2, 3, 2) and (m["fluid_regions"], m["solid_regions"]) == (2, 1)


def _faces(path):
    lines = (path / "processor0" / "faces.txt").read_text().split("\n")
    assert lines[0].split()[:3] == ["meta", "fluid", str(len([x for x in lines[1:] if x.strip()]))]
    return [tuple(int(v) for v in x.split()) for x in lines[1:] if x.strip()]


def test_periodic_processor_faces(hs_lib, tmp_path):
    """The decomposer's face list for a periodic 1 x 2 x 2 all-fluid grid, worked out by hand
    from decomposeMesh (mesh.cpp:1241-1389) and the face numbering with NyP = Ny
    (mesh.h:209-219): x faces 0-7, y faces 8-11, z faces 12-17.  Row 2's north face
    fy(1,3,k) is the south face of row 1 one plane up (fy(1,1,k+1)), and for the top
    plane the bottom z face 12.  So the y-face pairs run 1-2, 2-3 (cells 0-based 0-1,
    1-2, 2-3), face 8 (row 1's south face on the first plane, seen once, whose linear
    neighbour index is 0) is dropped, and face 12 is a wall listed from both cells,
    keeping the later owner."""
    out = mesh.write_proc(str(tmp_path), text=_cyl(2, 0, 360, nx=1, nz=2), lib=hs_lib)
    internal = [(9, 0, 1), (14, 0, 2), (10, 1, 2), (15, 1, 3), (11, 2, 3)]
    wall = [(0, 0), (1, 0), (2, 1), (3, 1), (4, 2), (5, 2), (6, 3), (7, 3), (12, 3), (13, 1), (16, 2), (17, 3)]
    assert _faces(tmp_path) == internal + [(g, o, 0) for g, o in wall]
    assert out["fluid"]["bounds"] == [("wall", 5, 12)]
    # the files read back
    back = mesh.read_proc(str(tmp_path), text=_cyl(2, 0, 360, nx=1, nz=2), lib=hs_lib)
    assert back["parts"]["fluid"]["nfaces"] == 17 and back["periodic_y"]
