"""File writers of the C ABI (src/utils/io.cpp layouts) against the
reference's own files: the restart round trip, the Tecplot file the reference
wrote for a state (golden .tec), and the surface file against the
reference's surface values (golden fixtures)."""
import os

import numpy as np
import pytest

import golden_util as gu
from structured_amd import lib

pytestmark = pytest.mark.gpu


def test_restart_round_trip_and_size_check(tmp_path):
    case, z, _ = gu.load("laminar_plate_25x17")
    eq = lib.Equation(case)
    p = str(tmp_path / "case.out")
    eq.write_restart(p, z["q"])
    assert os.path.getsize(p) == z["q"].size * 8
    # write_restart's layout: raw float64 q[i][j][k]
    assert np.array_equal(np.fromfile(p), z["q"].reshape(-1))
    assert np.array_equal(eq.read_restart(p), z["q"])
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(lib.SgError, match="expected"):
        eq.read_restart(p)


def test_tecplot_matches_reference_file(tmp_path):
    name = "explicit_fe_ramp_laminar_plate_25x17"
    case, _, _, q = gu.load_explicit(name)
    eq = lib.Equation(case)
    p = str(tmp_path / "case.tec")
    eq.write_tecplot(p, q)
    mine = open(p).read().splitlines()
    ref = open(os.path.join(gu.HERE, name + ".tec")).read().splitlines()
    assert mine[:3] == ref[:3] and len(mine) == len(ref)
    same = sum(a == b for a, b in zip(mine, ref))
    print("identical lines", same, "of", len(ref), flush=True)
    a = np.array([[float(t) for t in ln.split()] for ln in mine[3:]])
    b = np.array([[float(t) for t in ln.split()] for ln in ref[3:]])
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=0)
    assert same >= len(ref) - 3


@pytest.mark.parametrize("name", [n for n in gu.names() if not n.startswith("inviscid")])
def test_surface_file_matches_reference(tmp_path, name):
    case, z, _ = gu.load(name)
    eq = lib.Equation(case)
    p = str(tmp_path / "case.surface")
    f = eq.write_surface(p, z["q"])
    s = np.loadtxt(p)
    ref = z["surface"]
    assert s.shape == ref.shape
    np.testing.assert_allclose(s[:, 0], ref[:, 0], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(s[:, 1], ref[:, 1], rtol=2e-5, atol=1e-5 * np.abs(ref[:, 1]).max())
    np.testing.assert_allclose(s[:, 2], ref[:, 2], rtol=2e-5, atol=1e-5 * np.abs(ref[:, 2]).max())
    _, _, _, f2 = eq.wall(z["q"])
    assert np.array_equal(f, f2)


def test_npz_fields(tmp_path):
    case, z, _ = gu.load("laminar_plate_25x17")
    eq = lib.Equation(case)
    p = str(tmp_path / "case.npz")
    eq.write_npz(p, z["q"])
    d = np.load(p)
    assert d["q"].shape == (case.nic, case.njc, case.nvar) and np.array_equal(d["q"], z["q"])
    assert d["xc"].shape == (case.nic, case.njc)
    np.testing.assert_allclose(d["rho"] * d["u"], z["q"][..., 1], rtol=1e-14)
