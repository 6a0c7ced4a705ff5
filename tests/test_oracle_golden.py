"""The CPU oracle against the reference's own outputs (golden fixtures made
by compiling the reference's sources, tests/golden/make_golden.py), and
against the live compiled reference where oracle/_ref exists."""
import numpy as np
import pytest

import golden_util as gu
import oracle_ffi as of

NAMES = gu.names()


@pytest.mark.parametrize("name", NAMES)
def test_residual_matches_reference(name):
    case, z, _ = gu.load(name)
    r, l2 = of.residual(case, z["q"])
    scale = np.abs(z["R"]).max()
    assert np.abs(r - z["R"]).max() <= 1e-13 * scale
    # Solver::step norm, src/solver/solver.cpp:125-134 (unnormalised)
    np.testing.assert_allclose(l2, np.sqrt((z["R"] ** 2).sum(axis=(0, 1))), rtol=1e-13)


@pytest.mark.parametrize("name", NAMES)
def test_dt_matches_reference(name):
    case, z, meta = gu.load(name)
    np.testing.assert_allclose(of.dt(case, z["q"], meta["cfl"]), z["dt"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("inviscid")])
def test_surface_matches_reference(name):
    case, z, _ = gu.load(name)
    x, cp, cf, _ = of.surface(case, z["q"])
    s = z["surface"]
    # the reference writes the .surface file with 6 significant digits
    np.testing.assert_allclose(x, s[:, 0], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(cp, s[:, 1], rtol=2e-5, atol=1e-5 * np.abs(s[:, 1]).max())
    np.testing.assert_allclose(cf, s[:, 2], rtol=2e-5, atol=1e-5 * np.abs(s[:, 2]).max())


@pytest.mark.parametrize("name", NAMES)
def test_jv_dual_matches_reference_fd(name):
    """Dual-number J.v of the oracle against a central difference of the
    compiled reference's residual (O(h^2) truncation + roundoff/h)."""
    case, z, _ = gu.load(name)
    _, jv = of.jv(case, z["q"], z["v"])
    scale = np.abs(z["jv_fd"]).max()
    assert np.abs(jv - z["jv_fd"]).max() <= 1e-6 * scale


@pytest.mark.skipif(not of.have_ref(), reason="compiled reference (oracle/_ref) not built")
def test_live_reference_seeds():
    """Fresh seeds against the live compiled reference."""
    for name in ("laminar_plate_25x17", "wake_25x11", "bump_iso_33x17", "laminar_plate_25x17_ausm",
                 "inviscid_bump_ausm_33x17"):
        case, _, _ = gu.load(name)
        ref = of.RefRun(case)
        for seed in (3, 4):
            q = case.perturbed_state(seed=seed, amp=0.02)
            r, _ = of.residual(case, q)
            rr = ref.residual(q)
            assert np.abs(r - rr).max() <= 1e-13 * np.abs(rr).max()


@pytest.mark.parametrize("name", gu.explicit_names())
def test_explicit_solve_matches_reference(name):
    """The oracle's composition of forward Euler / RK4-Jameson steps, the
    stop rule and the CFL ramp against the reference's own Solver::solve
    (compiled without ADOL-C) from the same state."""
    case, q0, solver, q_ref = gu.load_explicit(name)
    q, hist = of.explicit_solve(case, q0, **solver)
    assert len(hist) == solver["iteration_max"] + 2
    assert np.abs(q - q_ref).max() <= 1e-12 * np.abs(q_ref - q0).max()
    # a dropped stage, a wrong stage weight or an off-by-one step count moves q
    # by far more than that tolerance
    for wrong in (dict(solver, iteration_max=solver["iteration_max"] - 1),
                  dict(solver, scheme="forward_euler" if solver["scheme"] != "forward_euler" else "rk4_jameson")):
        qw, _ = of.explicit_solve(case, q0, **wrong)
        assert np.abs(qw - q_ref).max() > 1e-6 * np.abs(q_ref - q0).max()


def test_naca0012_example_matches_reference():
    """The reference's own NACA 0012 example (C-grid, wake cut, wall, far
    field; its grid file and config values): residual, norms, dt and the
    surface file."""
    case, q, z = gu.load_naca()
    n = int(z["sub"])
    r, l2 = of.residual(case, q)
    assert np.abs(r[::n, ::n] - z["R_sub"]).max() <= 1e-13 * np.abs(z["R_sub"]).max()
    np.testing.assert_allclose(l2, z["l2"], rtol=1e-13)
    np.testing.assert_allclose(of.dt(case, q, 3.0)[::n, ::n], z["dt_sub"], rtol=1e-14, atol=0)
    x, cp, cf, _ = of.surface(case, q)
    s = z["surface"]
    np.testing.assert_allclose(x, s[:, 0], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(cp, s[:, 1], rtol=2e-5, atol=1e-5 * np.abs(s[:, 1]).max())
    np.testing.assert_allclose(cf, s[:, 2], rtol=2e-5, atol=1e-5 * np.abs(s[:, 2]).max())
