"""Parity at the full sizes of the benchmark configurations (laminar plate
97x65, SA plate 545x385, SA bumps 705x321 with a seeded beta field, 385x193
and 1409x641): the HIP residual against the CPU oracle on seeded perturbed
states, the fused norms against the norm of the downloaded residual, J.v
against the oracle's dual-number J.v, the wall quantities, and the adjoint
dot-product identity <psi, J v> = <J^T psi, v> at 545x385.

Tolerances as at the small sizes: residual 1e-12 of max|R| (1e-11 per
component), J.v 1e-10, wall Cp/Cf 1e-12 of their range, norms 1e-12, dot
product 1e-12."""
import numpy as np
import pytest

import oracle_ffi as of
from structured_amd import cases, lib

pytestmark = pytest.mark.gpu

CONFIGS = {
    "laminar_plate_97x65": lambda: cases.laminar_plate(97, 65),
    "sa_plate_545x385": lambda: cases.sa_plate(545, 385),
    "sa_bump_705x321": lambda: cases.sa_bump(705, 321),
    "sa_bump_385x193": lambda: cases.sa_bump(385, 193),
    "sa_bump_1409x641": lambda: cases.sa_bump(1409, 641, height=0.065, width=0.25, re=5e6),
}


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup(name):
    case = CONFIGS[name]()
    eq = lib.Equation(case)
    wd = beta = None
    if case.turbulence:
        wd = eq.walldist()
        beta = cases.blob_beta(case, seed=1)
        eq.set_beta(beta)
    return case, eq, wd, beta


@pytest.mark.parametrize("name", list(CONFIGS))
def test_fullsize_residual_and_norms(name):
    case, eq, wd, beta = setup(name)
    q = case.perturbed_state(seed=11)
    r, l2 = eq.calc_residual(q)
    ro, _ = of.residual(case, q, wd=wd, beta=beta)
    err = rel(r, ro)
    print(name, case.nic * case.njc, "cells: residual rel err", err, flush=True)
    assert err <= 1e-12
    for m in range(case.nvar):
        assert rel(r[..., m], ro[..., m]) <= 1e-11
    # the fused fixed-order reduction against the host norm of the same R
    np.testing.assert_allclose(l2, np.sqrt((r ** 2).sum(axis=(0, 1))), rtol=1e-12)


@pytest.mark.parametrize("name", ["sa_plate_545x385", "sa_bump_1409x641"])
def test_fullsize_jv_and_wall(name):
    case, eq, wd, beta = setup(name)
    q = case.perturbed_state(seed=12)
    v = np.random.default_rng(5).standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    jv = eq.jv(q, v)
    _, jvo = of.jv(case, q, v, wd=wd, beta=beta)
    err = rel(jv, jvo)
    print(name, "J.v rel err", err, flush=True)
    assert err <= 1e-10
    x, cp, cf, f = eq.wall(q)
    xo, cpo, cfo, fo = of.surface(case, q, wd=wd, beta=beta)
    np.testing.assert_array_equal(x, xo)
    assert np.abs(cp - cpo).max() <= 1e-12 * np.abs(cpo).max()
    assert np.abs(cf - cfo).max() <= 1e-12 * np.abs(cfo).max()


@pytest.mark.parametrize("name", ["sa_plate_545x385", "sa_bump_705x321"])
def test_fullsize_dot_product_identity(name):
    """<psi, J v> = <J^T psi, v> at the SA plate and at the adjoint
    configuration's bump (705x321 with the seeded 8-blob beta field)."""
    case, eq, _, _ = setup(name)
    q = case.perturbed_state(seed=13)
    rng = np.random.default_rng(6)
    psi = rng.standard_normal(q.shape)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    jt, _ = eq.jtv(q, psi)
    jv = eq.jv(q, v)
    a, b = np.vdot(psi, jv), np.vdot(jt, v)
    print(name, "<psi, J v> %.17e  <J^T psi, v> %.17e" % (a, b), flush=True)
    assert abs(a - b) <= 1e-12 * np.linalg.norm(psi) * np.linalg.norm(jv)


def test_fullsize_mlp_closure_residual_and_jv():
    """The closure network fused into the residual and J.v kernels at the
    largest configuration (1409x641 bump): beta against the oracle's FP64
    network with the kernel's BF16 rounding points, the residual against the
    oracle's at the kernel's beta (everything but the network, 1e-12), J.v
    components 0-3 (1e-10) and the closure's part of the nu~ component."""
    import mlp_model as mm
    case, eq, wd, _ = setup("sa_bump_1409x641")
    q = case.perturbed_state(seed=14)
    p = mm.glorot(seed=4).astype(np.float32)
    p[mm.OFF["b3"][0]] = 0.05
    eq.set_closure(2, params=p)
    of.set_closure(2, params=p)
    try:
        _, b = eq.closure_eval(q)
        _, bo = of.closure_eval(case, q, wd=wd)
        eb = np.abs(b - bo)
        print("1409x641 mlp beta vs oracle: median %.2e p99.9 %.2e max %.2e" % (
            np.median(eb), np.quantile(eb, 0.999), eb.max()), flush=True)
        assert np.median(eb) <= 1e-6 and eb.max() <= 2e-3
        r, _ = eq.calc_residual(q)
        of.set_closure(0)
        ro, _ = of.residual(case, q, wd=wd, beta=b)
        print("1409x641 mlp residual rel err %.2e" % rel(r, ro), flush=True)
        assert rel(r, ro) <= 1e-12
        v = np.random.default_rng(15).standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
        jv = eq.jv(q, v)
        _, jv_fixed = of.jv(case, q, v, wd=wd, beta=b)
        of.set_closure(2, params=p)
        _, jvo = of.jv(case, q, v, wd=wd)
        co, cg = (jvo - jv_fixed)[..., 4], (jv - jv_fixed)[..., 4]
        e = np.abs(cg - co) / np.abs(co).max()
        print("1409x641 closure part of J.v: median %.2e p99.9 %.2e max %.2e; rest %.2e" % (
            np.median(e), np.quantile(e, 0.999), e.max(), rel(jv[..., :4], jvo[..., :4])), flush=True)
        assert rel(jv[..., :4], jvo[..., :4]) <= 1e-10
        # measured: median 9e-18, p99.9 1e-9, max 1.1e-6
        assert np.median(e) <= 1e-8 and np.quantile(e, 0.999) <= 1e-6 and e.max() <= 1e-4
    finally:
        of.set_closure(0)
