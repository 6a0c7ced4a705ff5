"""Matrix-free Newton-Krylov on the GPU (linear-solver interface and the
pseudo-transient Newton loop).

Linear solve: the FGMRES solution of (I/dt - J) dq = R is checked with the
oracle's own dual-number J.v (||A dq - R|| / ||R|| <= rtol + slack).
Nonlinear solve: the laminar flat plate (97x65, M=0.2, Re=1e5) converges to
max_k ||R_k||_2 < 1e-10; at that state the oracle (the reference's residual,
bit-for-bit) gives a residual of the same size, the GPU wall Cf equals the
oracle's wall Cf to 1e-12 of its range, and Cf sqrt(Re_x) is near the
Blasius value 0.664 on the plate (sanity band, not parity)."""
import numpy as np
import pytest

import oracle_ffi as of
from structured_amd import cases, lib

pytestmark = pytest.mark.gpu


def test_linear_solve_residual():
    """FGMRES's residual estimate equals the true residual of I/dt - J built
    from the oracle's J.v; the j-line preconditioner beats block Jacobi,
    which beats no preconditioning, at a fixed iteration budget."""
    case = cases.laminar_plate(41, 25)
    q = case.perturbed_state(seed=4, amp=0.002)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 5.0)
    r, _ = eq.calc_residual(q)
    res = {}
    for precond in (0, 1, 2):
        eq.ls_preallocate(restart=30, precond=precond, maxit=60, rtol=1e-12)
        eq.ls_set_lhs(q, dt)
        eq.ls_set_rhs(r)
        dq, it, rr = eq.ls_solve()
        _, jdq = of.jv(case, q, dq)
        Adq = dq / dt - jdq
        rel = np.linalg.norm(Adq - r) / np.linalg.norm(r)
        print("precond", precond, "iters", it, "relres", rr, "true", rel)
        assert abs(rel - rr) <= 1e-3 * rr + 1e-10
        res[precond] = rr
    assert res[2] < res[1] < res[0]
    eq.ls_preallocate(restart=30, precond=4, maxit=60, rtol=1e-12)
    eq.ls_set_lhs(q, dt)
    eq.ls_set_rhs(r)
    dq, it, rr = eq.ls_solve()
    print("precond 4 iters", it, "relres", rr)
    assert rr < res[2]


def test_laminar_plate_converges_and_cf_matches_oracle():
    case = cases.laminar_plate(97, 65)
    eq = lib.Equation(case)
    q0 = case.freestream_state()
    # the reference's CFL ramp (cfl <- cfl^1.2 every step, src/solver/solver.cpp:212-220)
    q, rep, hist = eq.solve(q0, cfl=5.0, cfl_ramp=1, cfl_ramp_iteration=0, cfl_ramp_exponent=1.2, cfl_max=1e10,
                            iteration_max=60, tolerance=1e-10, gmres_restart=30, gmres_maxit=120, gmres_rtol=1e-3,
                            precond=4)
    print({k: rep[k] for k in ("iterations", "converged", "linear_iterations", "wall_s")}, rep["l2_final"])
    assert rep["converged"] == 1
    ro, l2o = of.residual(case, q)
    assert np.all(l2o < 1e-9), l2o
    x, cp, cf, f = eq.wall(q)
    xo, cpo, cfo, fo = of.surface(case, q)
    assert np.abs(cf - cfo).max() <= 1e-12 * np.abs(cfo).max()
    n_up = case.meta["n_up"]
    re = case.fs["u_inf"] / case.fs["mu_inf"]
    xs, cfs = x[n_up:], cf[n_up:]
    sel = (xs > 0.3) & (xs < 1.6)
    blasius = cfs[sel] * np.sqrt(re * xs[sel])
    print("Cf*sqrt(Re_x) on the plate:", blasius.min(), blasius.max())
    assert np.all(np.abs(blasius - 0.664) < 0.1)


@pytest.mark.parametrize("precond", [1, 2, 3, 4, 5])
def test_preconditioner_matches_numpy_model(precond):
    """The GPU preconditioner (blocks by colouring J.v, line factorisation,
    multicolour Gauss-Seidel) equals its numpy restatement built from the
    oracle's dense Jacobian."""
    import precond_model as pm
    case = cases.laminar_plate(13, 9, dy1=1e-2)
    q = case.perturbed_state(seed=6, amp=0.01)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 20.0)
    eq.ls_preallocate(restart=10, precond=precond, maxit=10, rtol=1e-6)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(3).standard_normal(q.shape)
    z = eq.ls_precond_apply(v)
    J = of.jacobian(case, q)
    B = pm.blocks_from_dense(J, case.nic, case.njc, case.nvar)
    zr = pm.apply(B, 1.0 / dt[..., 0], v, precond)
    err = np.abs(z - zr).max() / np.abs(zr).max()
    print("precond", precond, "model err", err)
    assert err <= 1e-10


@pytest.mark.parametrize("precond", [4, 5])
def test_ilu_matches_numpy_model_sa(precond):
    """The ILU preconditioners with the five SA unknowns per cell."""
    import precond_model as pm
    case = cases.sa_bump(13, 9, dy1=1e-3)
    q = case.perturbed_state(seed=6, amp=0.01)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    eq.ls_preallocate(restart=10, precond=precond, maxit=10, rtol=1e-6)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(3).standard_normal(q.shape)
    z = eq.ls_precond_apply(v)
    J = of.jacobian(case, q, wd=eq.walldist())
    B = pm.blocks_from_dense(J, case.nic, case.njc, case.nvar)
    zr = pm.apply(B, 1.0 / dt[..., 0], v, precond)
    err = np.abs(z - zr).max() / np.abs(zr).max()
    print("SA precond", precond, "model err", err)
    assert err <= 1e-9


@pytest.mark.parametrize("width,sa", [(8, False), (16, False), (8, True), (24, True), (32, True)])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("precond", [6, 7])
def test_chunked_ilu_matches_numpy_model(width, sa, transpose, precond):
    """Chunked line ILU(0) (precond 6; precond 7 on the 3x3-box pattern):
    chunks of `width` columns (a partial last chunk at width 8 on 12 columns,
    and at width 32 on 40 columns), M^-1 and M^-T against the model with
    float32 factors.  precond 6 factors the exact Jacobian; precond 7, meant
    for lhs_order 1, the first-order one (whose +-2 blocks are zero, so the
    box pattern drops fill only)."""
    import copy
    import precond_model as pm
    ni = 41 if width >= 24 else 13
    case = cases.sa_bump(ni, 9, dy1=1e-3) if sa else cases.laminar_plate(ni, 9, dy1=1e-2)
    q = case.perturbed_state(seed=6, amp=0.01)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    nine = precond == 7
    eq.ls_preallocate(restart=10, precond=precond, maxit=10, rtol=1e-6)
    eq.ls_set_pc(1 if nine else 0, 1.0, width)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(3).standard_normal(q.shape)
    z = eq.ls_precond_apply(v, transpose=transpose)
    cj = copy.deepcopy(case)
    if nine:
        cj.order = 1
    J = of.jacobian(cj, q, wd=eq.walldist() if sa else None)
    B = pm.blocks_from_dense(J, case.nic, case.njc, case.nvar)
    zr = pm.ilu_chunked(B, 1.0 / dt[..., 0], v, width, transpose=transpose, nine_point=nine)
    z64 = pm.ilu_chunked(B, 1.0 / dt[..., 0], v, width, transpose=transpose, f32=False, nine_point=nine)
    err = float(np.abs(z - zr).max() / np.abs(zr).max())
    print("chunked ILU precond", precond, "width", width, "sa", sa, "transpose", transpose, "model err", err,
          "(float64 factors: %.1e)" % float(np.abs(z - z64).max() / np.abs(z64).max()))
    if nine:  # the 13-point model (which keeps the fill) misses by far more than the bar
        zo = pm.ilu_chunked(B, 1.0 / dt[..., 0], v, width, transpose=transpose)
        assert float(np.abs(z - zo).max() / np.abs(zo).max()) > 1e-7
    # the model stores its factors in float32 like the device; rounding at a
    # float32 tie can differ in the last place, far below 1e-8
    assert err <= 1e-8


def test_first_order_preconditioner_blocks():
    """lhs_order = 1: the preconditioner factors the Jacobian of the
    first-order residual (the oracle's J at order 1), not the exact one."""
    import copy
    import precond_model as pm
    case = cases.laminar_plate(13, 9, dy1=1e-2)
    q = case.perturbed_state(seed=7, amp=0.01)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    eq.ls_preallocate(restart=10, precond=5, maxit=10, rtol=1e-6)
    eq.ls_set_pc(1, 1.0, 0)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(4).standard_normal(q.shape)
    z = eq.ls_precond_apply(v)
    c1 = copy.deepcopy(case)
    c1.order = 1
    B1 = pm.blocks_from_dense(of.jacobian(c1, q), case.nic, case.njc, case.nvar)
    B2 = pm.blocks_from_dense(of.jacobian(case, q), case.nic, case.njc, case.nvar)
    z1 = pm.ilu_natural(B1, 1.0 / dt[..., 0], v)
    z2 = pm.ilu_natural(B2, 1.0 / dt[..., 0], v)
    err = float(np.abs(z - z1).max() / np.abs(z1).max())
    print("first-order preconditioner model err", err, "distance to second order",
          float(np.abs(z2 - z1).max() / np.abs(z1).max()))
    assert err <= 1e-9
    assert np.abs(z2 - z1).max() > 1e-3 * np.abs(z1).max()


def test_first_order_preconditioner_blocks_periodic_channel():
    """lhs_order = 1 on the reference's periodic channel (2 x 100 cells,
    periodic in i, isothermal walls, body force): every ghost cell reads
    interior cells inside the 3x3 box of its readers, so the first-order
    blocks come from the 3 x 3 colouring, with the periodic couplings folded
    into the neighbour blocks exactly as the dense first-order Jacobian has
    them.  The natural-order ILU of those blocks equals the model's."""
    import copy
    import golden_util as gu
    import precond_model as pm
    case, z, _ = gu.load("channel_3x101")
    q = z["q"]
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    eq.ls_preallocate(restart=10, precond=5, maxit=10, rtol=1e-6)
    eq.ls_set_pc(1, 1.0, 0)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(5).standard_normal(q.shape)
    zd = eq.ls_precond_apply(v)
    c1 = copy.deepcopy(case)
    c1.order = 1
    B1 = pm.blocks_from_dense(of.jacobian(c1, q), case.nic, case.njc, case.nvar)
    z1 = pm.ilu_natural(B1, 1.0 / dt[..., 0], v)
    err = float(np.abs(zd - z1).max() / np.abs(z1).max())
    print("periodic channel first-order preconditioner model err", err)
    assert err <= 1e-9


@pytest.mark.parametrize("precond", [6, 7])
def test_precond_apply_dev_async_equals_host_entry(precond):
    """sg_ls_precond_apply_dev_async (device SoA vectors, the context's
    stream) gives the host entry point's z bit for bit, M^-1 and M^-T, and
    refuses overlapping v and z."""
    import ctypes as C
    torch = pytest.importorskip("torch")
    so = lib.load()
    case = cases.sa_bump(41, 17, dy1=1e-3)
    q = case.perturbed_state(seed=6, amp=0.01)
    eq = lib.Equation(case)
    dt = eq.calc_dt(q, 50.0)
    eq.ls_preallocate(restart=10, precond=precond, maxit=10, rtol=1e-6)
    eq.ls_set_pc(1, 1.0, 16)
    eq.ls_set_lhs(q, dt)
    v = np.random.default_rng(5).standard_normal(q.shape)
    dev = torch.device("cuda:0")
    aos = torch.from_numpy(np.ascontiguousarray(v).reshape(-1)).to(dev)
    vd = torch.zeros(eq.nvar * eq.plane, dtype=torch.float64, device=dev)
    eq.aos_to_soa_dev(aos.data_ptr(), vd.data_ptr())
    for transpose in (0, 1):
        zh = eq.ls_precond_apply(v, transpose=bool(transpose))
        zd = torch.full_like(vd, -1.0)
        rc = so.sg_ls_precond_apply_dev_async(eq.h, C.c_void_p(vd.data_ptr()), C.c_void_p(zd.data_ptr()), transpose)
        assert rc == 0, so.sg_last_error(eq.h)
        za = torch.empty_like(aos)
        eq.soa_to_aos_dev(zd.data_ptr(), za.data_ptr())
        eq.synchronize()
        np.testing.assert_array_equal(za.cpu().numpy().reshape(q.shape), zh)
    rc = so.sg_ls_precond_apply_dev_async(eq.h, C.c_void_p(vd.data_ptr()), C.c_void_p(vd.data_ptr() + 8), 0)
    assert lib.ERRORS[rc] == "SG_ERR_ARG"
