"""Adjoint products and the wall objective on the GPU.

J^T psi: the dot-product identity <psi, J v> = <J^T psi, v> against the
matrix-free tangent kernel (1e-12 relative) and the oracle's dense
Jacobian transposed (tiny grids, 1e-10).  dR/dbeta^T psi: the identity
against the oracle's beta-direction derivative.  dF/dq of the wall
objective: central differences of F along random directions."""
import numpy as np
import pytest

import golden_util as gu
import oracle_ffi as of
from structured_amd import cases, lib

pytestmark = pytest.mark.gpu


def small_cases():
    return [cases.laminar_plate(21, 13), cases.sa_plate(21, 15, dy1=1e-3), cases.sa_bump(25, 13, dy1=1e-3)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_jtv_vs_oracle_dense_transpose(k):
    case = small_cases()[k]
    q = case.perturbed_state(seed=12)
    eq = lib.Equation(case)
    psi = np.random.default_rng(5).standard_normal(q.shape)
    jt, db = eq.jtv(q, psi)
    wd = of.wall_distance(case) if case.turbulence else None
    J = of.jacobian(case, q, wd=wd)
    ref = (J.T @ psi.reshape(-1)).reshape(q.shape)
    err = np.abs(jt - ref).max() / np.abs(ref).max()
    print(case.name, "J^T psi vs dense", err)
    assert err <= 1e-10


def test_jtv_rejects_nonlocal_boundaries():
    """The wake boundary copies cells from the far side of the cut; its
    couplings lie outside the 13-point stencil, so the block-based adjoint
    refuses it instead of returning a wrong product."""
    case, z, _ = gu.load("wake_25x11")
    eq = lib.Equation(case)
    with pytest.raises(lib.SgError) as e:
        eq.jtv(z["q"], z["q"])
    assert e.value.code == -1


@pytest.mark.parametrize("name", ["laminar_plate_25x17", "bump_iso_33x17", "channel_3x101"])
def test_jtv_dot_product_identity_golden(name):
    case, z, _ = gu.load(name)
    eq = lib.Equation(case)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(z["q"].shape)
    jt, _ = eq.jtv(z["q"], psi)
    jv = eq.jv(z["q"], z["v"])
    lhs, rhs = np.vdot(psi, jv), np.vdot(jt, z["v"])
    print(name, lhs, rhs)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), np.linalg.norm(psi) * np.linalg.norm(jv))


def test_jtv_dot_product_identity_sa_large():
    case = cases.sa_plate(97, 65, dy1=1e-5)
    q = case.perturbed_state(seed=2)
    eq = lib.Equation(case)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal(q.shape)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    jt, _ = eq.jtv(q, psi)
    jv = eq.jv(q, v)
    lhs, rhs = np.vdot(psi, jv), np.vdot(jt, v)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(psi) * np.linalg.norm(jv)


def test_dbeta_transpose_vs_oracle():
    case = cases.sa_bump(25, 13, dy1=1e-3)
    q = case.perturbed_state(seed=3)
    eq = lib.Equation(case)
    rng = np.random.default_rng(6)
    beta = 1.0 + 0.2 * rng.standard_normal((case.nic, case.njc))
    eq.set_beta(beta)
    wd = eq.walldist()
    psi = rng.standard_normal(q.shape)
    w = rng.standard_normal((case.nic, case.njc))
    _, db = eq.jtv(q, psi)
    dRw = of.dbeta(case, q, w, wd=wd, beta=beta)
    lhs, rhs = np.vdot(psi, dRw), np.vdot(db, w)
    print("dbeta identity", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


@pytest.mark.parametrize("k", [0, 1])
def test_objective_gradient_vs_fd(k):
    case = [cases.laminar_plate(33, 17), cases.sa_bump(33, 17, dy1=1e-3)][k]
    q = case.perturbed_state(seed=9, amp=0.003)
    eq = lib.Equation(case)
    nb = case.ni - 2 * case.tail + 1
    x, cp, cf, _ = eq.wall(q)
    target = cf * (1.0 + 0.1 * np.sin(np.arange(nb)))
    F, g = eq.objective_cf(q, target)
    rng = np.random.default_rng(8)
    for _ in range(3):
        v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
        h = 1e-6
        Fp, _ = eq.objective_cf(q + h * v, target, grad=False)
        Fm, _ = eq.objective_cf(q - h * v, target, grad=False)
        fd = (Fp - Fm) / (2 * h)
        an = np.vdot(g, v)
        print("objective dF.v", an, fd)
        assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-30)
