"""Grid loaders of the C ABI (sg_read_grid: Mesh::plot3d_loader and
simple_loader, src/utils/mesh.cpp:133-169) on the host; no GPU needed."""
import os

import numpy as np
import pytest

from structured_amd import grids, lib

REF_GRID = "/root/reference/examples/laminar/channel/channel_3x101.unf2"


def test_p3d_round_trip(tmp_path):
    xv, yv, _ = grids.flat_plate(13, 9, 1e-3)
    p = str(tmp_path / "g.p3d")
    grids.write_p3d(p, xv, yv)
    x2, y2 = lib.read_grid(p, "p3d", 13, 9)
    assert np.array_equal(x2, xv) and np.array_equal(y2, yv)


def test_simple_round_trip(tmp_path):
    xv, yv = grids.bump_channel(11, 7, 1e-2, 0.05, 0.2)
    p = str(tmp_path / "g.txt")
    with open(p, "w") as f:
        for j in range(7):
            for i in range(11):
                f.write("%.20e %.20e\n" % (xv[i, j], yv[i, j]))
    x2, y2 = lib.read_grid(p, "simple", 11, 7)
    assert np.array_equal(x2, xv) and np.array_equal(y2, yv)


def test_grid_errors(tmp_path):
    xv, yv, _ = grids.flat_plate(13, 9, 1e-3)
    p = str(tmp_path / "g.p3d")
    grids.write_p3d(p, xv, yv)
    with pytest.raises(lib.SgError, match="header"):
        lib.read_grid(p, "p3d", 13, 10)
    with open(p) as f:
        head = f.read()[:400]
    with open(p, "w") as f:
        f.write(head)
    with pytest.raises(lib.SgError, match="truncated"):
        lib.read_grid(p, "p3d", 13, 9)
    with pytest.raises(lib.SgError, match="format"):
        lib.read_grid(p, "plot3d", 13, 9)


@pytest.mark.skipif(not os.path.exists(REF_GRID), reason="reference example grid not present")
def test_reference_example_grid():
    """The reference's own channel grid file reads the same through the C ABI
    and the Python loader."""
    x1, y1 = lib.read_grid(REF_GRID, "p3d", 3, 101)
    x2, y2 = grids.read_p3d(REF_GRID, 3, 101)
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    assert y1[0, 0] == 0.0 and y1[0, -1] > 0.0
