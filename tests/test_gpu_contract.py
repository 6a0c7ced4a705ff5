"""The C ABI's contract on the GPU: the Newton loop's stop rule and ramps as
the reference's Solver::solve, call-order errors of the linear-solver
interface, failure injection (non-finite states, linear solves that cannot
converge), the size of the caller's norm buffer, the initial state, the wall
force coefficients against the reference's own printout and the wall
objective against its definition."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import golden_util as gu
from structured_amd import cases, lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "ref_scalars.json")) as _f:
    REF_SCALARS = json.load(_f)


def reference_ramps(cfl0, ur0, n, cfl_ramp, cfl_it, cfl_e, ur_ramp, ur_it, ur_e):
    """CFL and UNDER_RELAXATION in force at counter 0..n-1 of Solver::solve
    (src/solver/solver.cpp:197-230): after the step of counter c the counter
    becomes c + 1 and, if it exceeds the ramp iteration, CFL <- min(CFL^e,
    1e12) and UR <- min(UR^e, 10)."""
    cfl, ur, out_c, out_u = cfl0, ur0, [], []
    for c in range(n):
        out_c.append(cfl)
        out_u.append(ur)
        counter = c + 1
        if cfl_ramp and counter > cfl_it:
            cfl = min(cfl ** cfl_e, 1e12)
        if ur_ramp and counter > ur_it:
            ur = min(ur ** ur_e, 10.0)
    return np.array(out_c), np.array(out_u)


@pytest.mark.parametrize("iteration_max", [0, 3, 6])
def test_newton_stop_rule_and_ramps_follow_solver_solve(iteration_max):
    """tolerance 0 (the reference's disabled test): iteration_max + 1 linear
    solves (counters 0..iteration_max), iteration_max + 2 residual
    evaluations, and the CFL / under-relaxation of each counter as
    Solver::solve ramps them."""
    case = cases.laminar_plate(33, 21)
    eq = lib.Equation(case)
    opts = dict(cfl=2.0, cfl_ramp=1, cfl_ramp_iteration=1, cfl_ramp_exponent=1.3, cfl_max=1e12,
                under_relaxation=0.5, under_relaxation_ramp=1, under_relaxation_ramp_iteration=2,
                under_relaxation_ramp_exponent=0.8, iteration_max=iteration_max, tolerance=0.0, gmres_restart=20,
                gmres_maxit=40, gmres_rtol=1e-2, precond=4)
    q, rep, hist = eq.solve(case.freestream_state(), **opts)
    assert rep["iterations"] == iteration_max + 1
    assert rep["converged"] == 0 and rep["code"] == -5
    assert len(hist["cfl"]) == iteration_max + 2
    cfl, ur = reference_ramps(2.0, 0.5, iteration_max + 2, True, 1, 1.3, True, 2, 0.8)
    np.testing.assert_array_equal(hist["cfl"], cfl)
    np.testing.assert_array_equal(hist["ur"], ur)
    eq.close()


def test_ls_calls_out_of_order_are_state_errors():
    """Changing the preconditioner after set_lhs (preallocate with another
    precond, set_pc with another chunk width) or overwriting the
    linearisation state (an adjoint product) invalidates the factors:
    solve and precond_apply then return SG_ERR_STATE before any launch."""
    case = cases.sa_plate(41, 25, dy1=1e-4)
    q = case.perturbed_state(seed=2, amp=0.002)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 10.0)
    r, _ = eq.calc_residual(q)

    def expect_state(fn):
        with pytest.raises(lib.SgError) as e:
            fn()
        assert e.value.code == -7, e.value

    eq.ls_preallocate(restart=20, precond=4, maxit=20, rtol=1e-2)
    eq.ls_set_lhs(q, dt)
    eq.ls_set_rhs(r)
    eq.ls_precond_apply(r)  # in order: fine
    eq.ls_preallocate(restart=20, precond=6, maxit=20, rtol=1e-2)
    expect_state(lambda: eq.ls_precond_apply(r))
    expect_state(lambda: eq.ls_solve())
    eq.ls_set_pc(lhs_order=1, pc_dt_factor=1.0, pc_width=16)
    eq.ls_set_lhs(q, dt)
    eq.ls_precond_apply(r)
    eq.ls_set_pc(lhs_order=1, pc_dt_factor=1.0, pc_width=32)
    expect_state(lambda: eq.ls_precond_apply(r))
    eq.ls_set_lhs(q, dt)
    eq.ls_precond_apply(r)
    eq.jtv(q, r)  # overwrites the linearisation state and the Jacobian blocks
    expect_state(lambda: eq.ls_solve())
    eq.close()


def test_nonfinite_state_is_reported():
    """A NaN in one cell of q: sg_residual returns SG_ERR_NAN (the reference
    traps floating-point exceptions, src/main.cpp:9); sg_solve from that
    state stops with SG_ERR_NAN at its first residual; an explicit step
    reports it too."""
    case = cases.laminar_plate(33, 21)
    q = case.freestream_state()
    q[7, 5, 3] = np.nan
    eq = lib.Equation(case)
    r, l2 = eq.calc_residual(q)
    assert not np.all(np.isfinite(l2))
    rc = lib.load().sg_residual(eq.h, lib._p(q), lib._p(np.empty_like(q)), lib._p(np.empty(4)))
    assert rc == -3
    with pytest.raises(lib.SgError) as e:
        eq.solve(q, cfl=1.0, iteration_max=3, tolerance=1e-12, gmres_restart=10, gmres_maxit=10, precond=4)
    assert e.value.code == -3
    with pytest.raises(lib.SgError) as e:
        eq.solve_explicit(q, scheme="forward_euler", cfl=0.5, iteration_max=2)
    assert e.value.code == -3
    eq.close()


def test_gmres_budget_exhausted_is_noconv():
    """maxit = 1 with a tight tolerance: sg_ls_solve makes one Krylov
    iteration and returns SG_ERR_NOCONV with the residual it reached; the
    Newton loop with that budget finishes its iterations without converging
    and reports SG_ERR_NOCONV."""
    case = cases.laminar_plate(33, 21)
    q = case.perturbed_state(seed=5, amp=0.002)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    r, _ = eq.calc_residual(q)
    eq.ls_preallocate(restart=10, precond=4, maxit=1, rtol=1e-12)
    eq.ls_set_lhs(q, dt)
    eq.ls_set_rhs(r)
    rc = lib.load().sg_ls_solve(eq.h, lib._p(np.empty(eq.shape)), C.byref(C.c_int()), C.byref(C.c_double()))
    assert rc == -5
    _, it, rr = eq.ls_solve()
    assert it == 1 and 1e-12 < rr < 1.0
    q2, rep, _ = eq.solve(case.freestream_state(), cfl=5.0, iteration_max=2, tolerance=1e-12, gmres_restart=10,
                          gmres_maxit=1, gmres_rtol=1e-12, precond=4)
    assert rep["code"] == -5 and rep["converged"] == 0 and rep["linear_iterations"] == 3
    assert np.all(np.isfinite(q2))
    eq.close()


def test_async_norm_writes_exactly_nvar_entries():
    """sg_residual_dev_async writes nvar norms into the caller's buffer and
    nothing after them: rows of a packed [k, nvar] array with guard values
    keep their neighbours."""
    torch = pytest.importorskip("torch")
    case = cases.sa_plate(65, 33, dy1=1e-4)
    q = case.perturbed_state(seed=6)
    eq = lib.Equation(case)
    _, l2_host = eq.calc_residual(q)
    dev = torch.device("cuda:0")
    aos = torch.from_numpy(q.reshape(-1).copy()).to(dev)
    qd = torch.zeros(eq.nvar * eq.plane, dtype=torch.float64, device=dev)
    rd = torch.zeros_like(qd)
    eq.aos_to_soa_dev(aos.data_ptr(), qd.data_ptr())
    nv = eq.nvar
    packed = torch.full((3, nv), -7.0, dtype=torch.float64, device=dev)
    eq.synchronize()
    eq.residual_dev_async(qd.data_ptr(), rd.data_ptr(), packed[1].data_ptr())
    eq.synchronize()
    out = packed.cpu().numpy()
    assert np.all(out[0] == -7.0) and np.all(out[2] == -7.0)
    np.testing.assert_allclose(out[1], l2_host, rtol=1e-13)
    eq.close()


@pytest.mark.parametrize("name", sorted(REF_SCALARS))
def test_initialize_matches_reference(name):
    """sg_initialize against the state the reference's
    EulerEquation::initialize sets (src/model/eulerequation.cpp:262-277)."""
    case = gu.load_naca()[0] if name.startswith("naca") else gu.load(name)[0]
    eq = lib.Equation(case)
    q0 = eq.initialize()
    assert np.array_equal(q0[..., :4], np.broadcast_to(REF_SCALARS[name]["init"], q0[..., :4].shape))
    eq.close()


def test_initialize_sa_fifth_unknown():
    case = cases.sa_plate(33, 21, dy1=1e-4)
    eq = lib.Equation(case)
    q0 = eq.initialize()
    assert np.array_equal(q0, case.freestream_state())
    assert np.all(q0[..., 4] == case.fs["rho_inf"] * case.fs["nut_inf"])
    eq.close()


@pytest.mark.parametrize("name", sorted(REF_SCALARS))
def test_wall_forces_match_reference_printout(name):
    """sg_wall's six force coefficients against the reference's own printed
    Pressure / Viscous / Total Cl Cd (src/utils/io.cpp:249-251, 6 digits)."""
    if name.startswith("naca"):
        case, q, _ = gu.load_naca()
    else:
        case, z, _ = gu.load(name)
        q = z["q"]
    eq = lib.Equation(case)
    _, _, _, f = eq.wall(q)
    ref = np.array(REF_SCALARS[name]["forces"])
    tol = 5.0001e-6 * np.abs(ref) + 1e-15 * np.abs(f).max()
    assert np.all(np.abs(f - ref) <= tol), (f, ref)
    eq.close()


def test_objective_value_is_its_definition():
    """sg_objective_cf's F equals sum_i ds_i (Cf_i - Cf*_i)^2 evaluated here
    from sg_wall's Cf (pinned to the reference's surface file) and the wall
    face lengths |normal_eta[i][0]| of sg_get_metrics."""
    case = cases.sa_bump(49, 25, dy1=1e-4)
    q = case.perturbed_state(seed=9, amp=0.003)
    eq = lib.Equation(case)
    _, _, cf, _ = eq.wall(q)
    target = cf * (1.0 + 0.1 * np.sin(np.arange(cf.size)))
    F, _ = eq.objective_cf(q, target, grad=False)
    neta = np.empty((case.nic, case.nj, 2))
    lib.load().sg_get_metrics(eq.h, None, lib._p(neta), None, None, None)
    first = case.tail - 1
    ds = np.hypot(neta[first:first + cf.size, 0, 0], neta[first:first + cf.size, 0, 1])
    expect = float(np.sum(ds * (cf - target) ** 2))
    assert abs(F - expect) <= 1e-13 * expect
    # a plain sum without the face lengths is far off
    assert abs(F - np.sum((cf - target) ** 2)) > 0.5 * expect
    eq.close()
