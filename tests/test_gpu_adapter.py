"""The C++ adapters of include/structured_gpu_adapter.hpp -- the SA model as
the reference's transport-equation hook (ntrans = 1, nu~_inf from the
configuration), EulerEquation's calc_dt / calc_residual / initialize and the
linear-solver interface -- drive the SA flat plate through the C ABI from a
C++ program (tests/cpp/adapter_driver, built by build()) with the reference's
Solver::step call sequence.  Its first residual equals the oracle's SA
residual of the same case (1e-12), so the adapter passes the freestream nu~,
boundary list and options through unchanged; the state after the Newton
steps equals, bit for bit, the same sequence run through the Python mirror."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as of
from structured_amd import cases, grids, lib

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "adapter_driver")


def test_sa_plate_through_cpp_adapters(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/adapter_driver is not built (run __graft_entry__.build())")
    case = cases.sa_plate(65, 49, dy1=1e-4)
    grid = str(tmp_path / "grid.p3d")
    grids.write_p3d(grid, case.xv, case.yv)
    steps, cfl = 4, 20.0
    prefix = str(tmp_path / "run_")
    mach, re = case.fs["u_inf"], case.fs["u_inf"] / case.fs["mu_inf"]
    out = subprocess.run([DRIVER, grid, str(case.ni), str(case.nj), str(case.meta["n_up"]), repr(mach), repr(re),
                          str(steps), repr(cfl), prefix], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    shape = (case.nic, case.njc, 5)
    r0 = np.fromfile(prefix + "r0.bin").reshape(shape)
    q_cpp = np.fromfile(prefix + "q.bin").reshape(shape)
    ro, _ = of.residual(case, case.freestream_state())
    err = np.abs(r0 - ro).max() / np.abs(ro).max()
    print("adapter first residual vs oracle: rel %.2e" % err)
    assert err <= 1e-12
    # the same calls through the Python mirror of the C ABI
    eq = lib.Equation(case)
    q = eq.initialize()
    for counter in range(steps):
        dt = eq.calc_dt(q, cfl)
        r, _ = eq.calc_residual(q)
        if counter == 0:
            eq.ls_preallocate(restart=30, precond=6, maxit=60, rtol=1e-3)
            eq.ls_set_pc(lhs_order=1, pc_dt_factor=1.0, pc_width=32)
        eq.ls_set_lhs(q, dt)
        eq.ls_set_rhs(r)
        dq, _, _ = eq.ls_solve()
        q = q + dq * 1.0
    eq.close()
    assert np.array_equal(q_cpp, q)
    norms = np.array([[float(v) for v in line.split()[3:]] for line in out.stdout.splitlines()
                      if line.startswith("counter")])
    # the uniform freestream has R_rho = 0 away from the walls; the Newton
    # steps develop the boundary layer and move every unknown
    assert norms.shape == (steps + 1, 5) and np.all(np.isfinite(norms))
    assert np.all(np.abs(q_cpp - case.freestream_state()).max(axis=(0, 1)) > 0)
