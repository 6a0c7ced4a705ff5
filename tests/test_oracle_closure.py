"""Pins for the oracle's beta closures (no counterpart in the reference):
features against closed forms on a uniform shear flow, the closures entering
the residual only through beta, the dual-number J.v of the analytic closure
against central differences, the oracle's MLP against the independent numpy
model, and the numpy model's backward pass against finite differences."""
import numpy as np
import pytest

import mlp_model as mm
import oracle_ffi as of
from structured_amd import cases
from test_oracle_sa import CB1, CT3, CV1, CV2, CV3, CW1, CW2, CW3, KAP, shear_case


@pytest.fixture(autouse=True)
def _reset_closure():
    of.set_closure(0)
    yield
    of.set_closure(0)


def sa_p_d(rho, nt, mu, omega, d):
    """(P1, D) of SA-noft2 with beta = 1 (closed form, per unit volume)."""
    nu = mu / rho
    if nt >= 0:
        chi = nt / nu
        fv1 = chi ** 3 / (chi ** 3 + CV1 ** 3)
        fv2 = 1 - chi / (1 + chi * fv1)
        sbar = nt * fv2 / (KAP ** 2 * d ** 2)
        st = omega + sbar if sbar >= -CV2 * omega else \
            omega + omega * (CV2 ** 2 * omega + CV3 * sbar) / ((CV3 - 2 * CV2) * omega - sbar)
        r = min(nt / (st * KAP ** 2 * d ** 2), 10.0) if st > 0 else 10.0
        g = r + CW2 * (r ** 6 - r)
        fw = g * ((1 + CW3 ** 6) / (g ** 6 + CW3 ** 6)) ** (1 / 6)
        return CB1 * st * rho * nt, CW1 * fw * rho * (nt / d) ** 2
    return CB1 * (1 - CT3) * omega * rho * nt, -CW1 * rho * (nt / d) ** 2


def shear_state(case, yc, a, nt):
    q = case.freestream_state()
    u = a * yc
    q[..., 0] = 1.0
    q[..., 1] = u
    q[..., 2] = 0.0
    q[..., 3] = (1 / 1.4) / 0.4 + 0.5 * u ** 2
    q[..., 4] = nt
    return q


@pytest.mark.parametrize("nt,dval", [(3e-3, 0.05), (4e-2, 0.02), (-2e-3, 0.1), (5e-4, 1.0)])
def test_features_closed_form_shear(nt, dval):
    a = 3.0
    case, yc = shear_case(a=a)
    q = shear_state(case, yc, a, nt)
    wd = np.full((case.nic, case.njc), dval)
    f, b = of.closure_eval(case, q, wd=wd)
    mu = case.fs["mu_inf"]  # T = T_inf
    nu = mu
    U2 = case.fs["u_inf"] ** 2
    smag = np.sqrt(a * a + 1e-12 * U2)
    P1, D = sa_p_d(1.0, nt, mu, a, dval)
    expect = [a / (a + smag), np.log1p(a * dval ** 2 / nu) / 10, np.arcsinh(nt / nu) / 8, 0.0,
              P1 / (abs(P1) + abs(D) + 1e-8 * U2 ** 1.5)]
    inner = (slice(2, case.nic - 2), slice(2, case.njc - 2))
    for k in range(5):
        np.testing.assert_allclose(f[inner][..., k], expect[k], rtol=1e-9, atol=1e-12, err_msg="feature %d" % k)
    assert np.all(b == 1.0)


def test_pressure_gradient_feature_sign():
    """f3 follows the sign of the streamwise pressure gradient: adverse
    (dp/dx > 0 along u > 0) gives f3 > 0, favourable f3 < 0."""
    case, yc = shear_case(a=0.0)
    xc = 0.25 * (case.xv[:-1, :-1] + case.xv[1:, :-1] + case.xv[:-1, 1:] + case.xv[1:, 1:])
    wd = np.full((case.nic, case.njc), 0.1)
    for sgn in (1.0, -1.0):
        q = case.freestream_state()
        p = 1 / 1.4 + sgn * 1e-3 * xc
        q[..., 3] = p / 0.4 + 0.5 * q[..., 1] ** 2
        f, _ = of.closure_eval(case, q, wd=wd)
        inner = f[2:-2, 2:-2, 3]
        assert np.all(np.sign(inner) == sgn)
        # tanh(d u dp/dx / (|u| (q_dyn + 0.025 rho U^2)))
        u = case.fs["u_inf"]
        expect = np.tanh(0.1 * sgn * 1e-3 / (0.5 * u * u + 0.025 * u * u) * u / np.sqrt(u * u * (1 + 1e-6)))
        np.testing.assert_allclose(inner, expect, rtol=1e-8)


def analytic_coef():
    return np.array([0.8, -0.5, 1.2, 2.0, -0.7, 0.1, 0.3])


def test_analytic_closure_enters_only_through_beta():
    case = cases.sa_plate(25, 17, dy1=1e-3)
    q = case.perturbed_state(seed=5)
    wd = of.wall_distance(case)
    r0, _ = of.residual(case, q, wd=wd)
    c = analytic_coef()
    of.set_closure(1, coef=np.r_[c[:6], 0.0])  # amp 0: plain SA, bitwise
    r1, _ = of.residual(case, q, wd=wd)
    assert np.array_equal(r0, r1)
    of.set_closure(1, coef=c)
    f, b = of.closure_eval(case, q, wd=wd)
    np.testing.assert_allclose(b, 1 + c[6] * np.tanh(f @ c[:5] + c[5]), rtol=1e-14)
    assert np.ptp(b) > 0.05
    rc, _ = of.residual(case, q, wd=wd)
    of.set_closure(0)
    rb, _ = of.residual(case, q, wd=wd, beta=b)
    np.testing.assert_allclose(rc, rb, rtol=1e-13, atol=1e-13 * np.abs(rb).max())
    assert np.abs(rc[..., 4] - r0[..., 4]).max() > 1e-6 * np.abs(r0[..., 4]).max()


def test_analytic_closure_jv_matches_fd():
    case = cases.sa_plate(21, 15, dy1=1e-3)
    q = case.perturbed_state(seed=8)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    wd = of.wall_distance(case)
    of.set_closure(1, coef=analytic_coef())
    _, jv = of.jv(case, q, v, wd=wd)
    h = 1e-7
    rp, _ = of.residual(case, q + h * v, wd=wd)
    rm, _ = of.residual(case, q - h * v, wd=wd)
    fd = (rp - rm) / (2 * h)
    of.set_closure(0)
    _, jv0 = of.jv(case, q, v, wd=wd)
    for m in range(5):
        scale = np.abs(fd[..., m]).max()
        assert np.abs(jv[..., m] - fd[..., m]).max() <= 1e-5 * scale, m
    # the closure's own contribution is resolved, not just the fixed-beta part
    contrib = jv[..., 4] - jv0[..., 4]
    assert np.abs(contrib).max() > 1e-3 * np.abs(jv[..., 4]).max()
    assert np.abs((jv - fd)[..., 4]).max() <= 1e-4 * np.abs(contrib).max()


def test_mlp_oracle_matches_numpy_model():
    case = cases.sa_plate(25, 17, dy1=1e-3)
    q = case.perturbed_state(seed=6)
    wd = of.wall_distance(case)
    p = mm.glorot(seed=4).astype(np.float32)
    p[mm.OFF["b3"][0]] = 0.05
    of.set_closure(2, params=p)
    f, b = of.closure_eval(case, q, wd=wd)
    y, _ = mm.forward(p.astype(np.float64), f.reshape(-1, 5).astype(np.float32), bf=True)
    np.testing.assert_allclose(b.reshape(-1), 1 + y, rtol=0, atol=1e-12)
    assert np.ptp(b) > 1e-3


def test_mlp_model_backward_matches_fd():
    rng = np.random.default_rng(3)
    p = mm.glorot(seed=1) + 0.05 * rng.standard_normal(mm.NPARAM)
    x = rng.standard_normal((40, 5))
    t = rng.standard_normal(40) * 0.3
    _, g = mm.loss_grad(p, x, t, bf=False)
    idx = rng.choice(mm.NPARAM, 60, replace=False)
    idx = np.r_[idx, [mm.OFF["b3"][0], mm.OFF["b2"][0] + 3, mm.OFF["W1"][0] + 7]]
    h = 1e-6
    for k in idx:
        e = np.zeros(mm.NPARAM)
        e[k] = h
        fd = (mm.loss_grad(p + e, x, t, bf=False)[0] - mm.loss_grad(p - e, x, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, np.abs(g).max()), k


def test_bf16_rounding_model():
    """Round-to-nearest-even at 8 significand bits, through float32."""
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.140625, 1e-3])
    r = mm.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0            # halfway, ties to even (down)
    assert r[2] == 1.0 + 2 ** -6  # halfway, ties to even (up)
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert r[5] == 3.140625
    assert abs(r[6] - 1e-3) <= 2 ** -9 * 1e-3


def test_adam_model_matches_torch_adam():
    """The model's Adam step (Kingma & Ba, bias-corrected moments) against
    torch.optim.Adam with the same hyper-parameters, over several steps in
    FP64 on the CPU."""
    import torch
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(4)]
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step, g in enumerate(grads, start=1):
        p, m, v = mm.adam(p, m, v, g, step, lr=1e-3)
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=0, atol=1e-15)
    assert np.abs(p - p0).max() > 1e-4
