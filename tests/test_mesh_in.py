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
