"""beta closures fused into the residual / J.v kernels, against the CPU
oracle (same grids, states, wall distance and closure parameters).

Analytic closure (FP64): features, beta, residual and J.v to the same bars
as the plain residual (1e-12 / 1e-10).  MLP closure (BF16 matrix cores, FP32
accumulation): beta against the oracle's FP64 network with the same BF16
rounding points (median 1e-6, max 2e-3: rare rounding flips), the residual
exactly equal to the field-beta residual at the beta the kernel produced, and
J.v whose closure contribution matches the oracle's dual-number one (max
1e-4 of its range, median 1e-8, p99 1e-5: the derivative of the network at
the same BF16 rounding points) while everything else matches to 1e-10."""
import numpy as np
import pytest

import mlp_model as mm
import oracle_ffi as of
from structured_amd import cases, lib

pytestmark = pytest.mark.gpu

COEF = np.array([0.8, -0.5, 1.2, 2.0, -0.7, 0.1, 0.3])


@pytest.fixture(autouse=True)
def _reset_oracle():
    of.set_closure(0)
    yield
    of.set_closure(0)


def sa_case(k):
    return [cases.sa_plate(33, 25, dy1=1e-4), cases.sa_bump(41, 21, dy1=1e-3)][k]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def mlp_params(seed=4):
    p = mm.glorot(seed=seed).astype(np.float32)
    p[mm.OFF["b3"][0]] = 0.05
    return p


@pytest.mark.parametrize("k", [0, 1])
def test_features_and_analytic_closure_vs_oracle(k):
    case = sa_case(k)
    q = case.perturbed_state(seed=3)
    eq = lib.Equation(case)
    wd = eq.walldist()
    eq.set_closure(1, coef=COEF)
    of.set_closure(1, coef=COEF)
    f, b = eq.closure_eval(q)
    fo, bo = of.closure_eval(case, q, wd=wd)
    # features are O(1) functions of Green-Gauss cell gradients, whose own
    # relative rounding is ~1e-16 |q| / |dq| (1e-11 on smooth near-freestream cells)
    for m in range(5):
        err = float(np.abs(f[..., m] - fo[..., m]).max())
        print("feature", m, "max abs err %.2e range %.2e" % (err, np.ptp(fo[..., m])))
        assert err <= 1e-9, (m, err)
    assert float(np.abs(b - bo).max()) <= 1e-9  # follows the features
    r, _ = eq.calc_residual(q)
    ro, _ = of.residual(case, q, wd=wd)
    assert rel(r, ro) <= 1e-12
    rng = np.random.default_rng(5)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    jv = eq.jv(q, v)
    _, jvo = of.jv(case, q, v, wd=wd)
    print("analytic closure: residual rel %.2e jv rel %.2e" % (rel(r, ro), rel(jv, jvo)))
    assert rel(jv, jvo) <= 1e-10


@pytest.mark.parametrize("k", [0, 1])
def test_mlp_closure_beta_vs_oracle_and_network(k):
    case = sa_case(k)
    q = case.perturbed_state(seed=4)
    eq = lib.Equation(case)
    wd = eq.walldist()
    p = mlp_params()
    eq.set_closure(2, params=p)
    of.set_closure(2, params=p)
    f, b = eq.closure_eval(q)
    fo, bo = of.closure_eval(case, q, wd=wd)
    assert float(np.abs(f - fo).max()) <= 1e-9
    err = np.abs(b - bo)
    print("mlp beta vs oracle: median %.2e max %.2e (range %.2e)" % (np.median(err), err.max(), np.ptp(bo)))
    assert np.median(err) <= 1e-6
    assert err.max() <= 2e-3
    assert np.ptp(bo) > 1e-3
    # the same network through the training kernels' forward pass
    net = lib.Mlp(max_rows=4096, seed=0)
    net.params = p
    y = net.forward(f.reshape(-1, 5).astype(np.float32)).reshape(b.shape)
    assert np.abs(b - (1.0 + y.astype(np.float64))).max() <= 1e-5
    # the context picks the closure up from the network object as well
    eq.set_closure_mlp(net)
    _, b2 = eq.closure_eval(q)
    assert np.array_equal(b, b2)


def test_set_mlp_weights_equals_set_closure():
    """sg_set_mlp_weights is sg_set_closure with the MLP kind: the same
    residual bit for bit; a NULL weight vector is an argument error."""
    import ctypes as C
    case = sa_case(0)
    q = case.perturbed_state(seed=6)
    p = mlp_params(seed=8)
    eq = lib.Equation(case)
    eq.set_closure(2, params=p)
    r1, _ = eq.calc_residual(q)
    eq.set_closure(0)
    so = lib.load()
    assert so.sg_set_mlp_weights(eq.h, C.c_void_p(p.ctypes.data)) == 0
    r2, _ = eq.calc_residual(q)
    assert np.array_equal(r1, r2)
    assert so.sg_set_mlp_weights(eq.h, None) != 0


def test_mlp_closure_residual_is_field_residual_at_its_beta():
    case = sa_case(0)
    q = case.perturbed_state(seed=6)
    eq = lib.Equation(case)
    eq.set_closure(2, params=mlp_params(seed=7))
    _, b = eq.closure_eval(q)
    r_mlp, _ = eq.calc_residual(q)
    eq.set_closure(0)
    eq.set_beta(b)
    r_field, _ = eq.calc_residual(q)
    assert np.array_equal(r_mlp, r_field)
    eq.set_beta(np.ones_like(b))
    r_plain, _ = eq.calc_residual(q)
    assert np.abs(r_plain[..., 4] - r_mlp[..., 4]).max() > 0


@pytest.mark.parametrize("k", [0, 1])
def test_mlp_closure_jv_vs_oracle_dual(k):
    case = sa_case(k)
    q = case.perturbed_state(seed=8)
    eq = lib.Equation(case)
    wd = eq.walldist()
    p = mlp_params(seed=9)
    rng = np.random.default_rng(10)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    eq.set_closure(2, params=p)
    jv = eq.jv(q, v)
    _, b_gpu = eq.closure_eval(q)
    # oracle, fixed beta = the kernel's beta: the part J.v has without the closure derivative
    of.set_closure(0)
    _, jv_fixed = of.jv(case, q, v, wd=wd, beta=b_gpu)
    of.set_closure(2, params=p)
    _, jvo = of.jv(case, q, v, wd=wd)
    contrib_o = jvo - jv_fixed
    contrib = jv - jv_fixed
    e = np.abs(contrib[..., 4] - contrib_o[..., 4]) / np.abs(contrib_o[..., 4]).max()
    print("closure part of J.v: max %.3e, rel err %.2e (median %.2e p99 %.2e); rest rel err %.2e" % (
        np.abs(contrib_o).max(), rel(contrib[..., 4], contrib_o[..., 4]), np.median(e), np.quantile(e, 0.99),
        rel(jv[..., :4], jvo[..., :4])))
    assert np.abs(contrib_o[..., 4]).max() > 1e-4 * np.abs(jvo[..., 4]).max()
    # measured: max 1.7e-6, median 8.5e-10 (BF16 rounding flips of single cells)
    assert rel(contrib[..., 4], contrib_o[..., 4]) <= 1e-4
    assert np.median(e) <= 1e-8 and np.quantile(e, 0.99) <= 1e-5
    assert rel(jv[..., :4], jvo[..., :4]) <= 1e-10
    assert np.abs(contrib[..., :4]).max() <= 1e-10 * np.abs(jvo[..., :4]).max()


def test_newton_converges_with_mlp_closure():
    case = cases.sa_plate(65, 49, dy1=1e-5, re=1e6)
    eq = lib.Equation(case)
    opts = dict(cfl=5.0, ser=1, ser_exponent=1.5, cfl_max=1e10, gmres_rtol=1e-4, precond=4, pc_cfl_max=300.0,
                gmres_restart=40, gmres_maxit=200, iteration_max=300, norm_weighted=1, tolerance=1e-11)
    q0, rep0, _ = eq.solve(case.freestream_state(), **opts)
    assert rep0["converged"], (rep0["iterations"], rep0["l2w_final"])
    p = mlp_params(seed=12)
    eq.set_closure(2, params=p)
    # warm start at a larger CFL; the BF16 network's rounding points set a
    # floor near 1.5e-11 on this norm, so the closure solve stops at 1e-10
    q1, rep1, hist = eq.solve(q0, **dict(opts, cfl=100.0, tolerance=1e-10))
    print("mlp closure solve", rep1["iterations"], rep1["l2w_final"])
    assert rep1["converged"], (rep1["iterations"], rep1["l2w_final"])
    _, _, cf0, _ = eq.wall(q0)
    _, _, cf1, _ = eq.wall(q1)
    assert np.abs(cf1 - cf0).max() > 1e-3 * np.abs(cf0).max()
    of.set_closure(2, params=p)
    ro, _ = of.residual(case, q1, wd=eq.walldist())
    r1, _ = eq.calc_residual(q1)
    r_start, _ = eq.calc_residual(case.freestream_state())
    # at the GPU solution the oracle's mean-flow residual (beta enters only
    # the SA equation) equals the GPU's to rounding of the large near-wall
    # terms; the SA row differs by the BF16 rounding of the network
    d4 = float(np.abs(ro[..., :4] - r1[..., :4]).max())
    print("oracle vs gpu residual at the GPU solution: mean flow %.3e, scale %.3e" % (d4, np.abs(r_start).max()))
    assert d4 <= 1e-12 * np.abs(r_start).max()
