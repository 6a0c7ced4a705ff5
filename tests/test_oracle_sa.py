"""Pins for the oracle's Spalart-Allmaras terms, which have no counterpart in
the reference (so no reference output to compare with; they are pinned by
value instead): the SA-noft2 source against a closed-form evaluation of the
published model (Spalart & Allmaras 1994 / NASA TMR, SA-neg branch for
nu~ < 0), and the dual-number J.v against central differences of the oracle
residual.  The diffusion, c_b2, density-gradient, eddy-viscosity and ghost
terms are pinned in closed form in test_oracle_pins.py."""
import numpy as np
import pytest

import oracle_ffi as of
from structured_amd import cases

CB1, SIG, CB2, KAP, CW2, CW3, CV1, CV2, CV3, CT3 = 0.1355, 2 / 3, 0.622, 0.41, 0.3, 2.0, 7.1, 0.7, 0.9, 1.2
CW1 = CB1 / KAP ** 2 + (1 + CB2) / SIG


def sa_source_closed_form(rho, nt, mu, omega, d, beta):
    """rho*(P - D) of SA-noft2 (TMR note 1(c) limiter on S~), per unit volume."""
    nu = mu / rho
    if nt >= 0:
        chi = nt / nu
        fv1 = chi ** 3 / (chi ** 3 + CV1 ** 3)
        fv2 = 1 - chi / (1 + chi * fv1)
        sbar = nt * fv2 / (KAP ** 2 * d ** 2)
        if sbar >= -CV2 * omega:
            st = omega + sbar
        else:
            st = omega + omega * (CV2 ** 2 * omega + CV3 * sbar) / ((CV3 - 2 * CV2) * omega - sbar)
        r = min(nt / (st * KAP ** 2 * d ** 2), 10.0) if st > 0 else 10.0
        g = r + CW2 * (r ** 6 - r)
        fw = g * ((1 + CW3 ** 6) / (g ** 6 + CW3 ** 6)) ** (1 / 6)
        return beta * CB1 * st * rho * nt - CW1 * fw * rho * (nt / d) ** 2
    return beta * CB1 * (1 - CT3) * omega * rho * nt + CW1 * rho * (nt / d) ** 2


def shear_case(nx=12, ny=10, a=3.0):
    """Uniform Cartesian grid, u = a*y, v = 0, constant rho, p and nu~."""
    xs = np.linspace(0.0, 1.0, nx + 1)
    ys = np.linspace(0.0, 0.8, ny + 1)
    xv = np.repeat(xs[:, None], ny + 1, axis=1)
    yv = np.repeat(ys[None, :], nx + 1, axis=0)
    fs = dict(rho_inf=1.0, u_inf=0.2, v_inf=0.0, p_inf=1 / 1.4, T_inf=1 / 1.4, mu_inf=1e-3, pr_inf=0.72,
              aoa_deg=0.0, nut_inf=3e-3)
    bcs = [dict(type="freestream", face="bottom", start=0, end=-1), dict(type="freestream", face="top", start=0, end=-1),
           dict(type="freestream", face="left", start=0, end=-1), dict(type="freestream", face="right", start=0, end=-1)]
    c = cases.Case("shear", xv, yv, fs, bcs, order=2, turbulence=1, tail=1)
    yc = 0.25 * (yv[:-1, :-1] + yv[1:, :-1] + yv[:-1, 1:] + yv[1:, 1:])
    return c, yc


@pytest.mark.parametrize("nt,dval,beta", [(3e-3, 0.05, 1.0), (3e-3, 0.2, 1.3), (4e-2, 0.02, 0.7),
                                          (5e-4, 1.0, 1.0), (-2e-3, 0.1, 1.0), (0.02, 0.01, 1.0)])
def test_sa_production_destruction_closed_form(nt, dval, beta):
    a = 3.0
    case, yc = shear_case(a=a)
    q = case.freestream_state()
    u = a * yc
    rho, p = 1.0, 1 / 1.4
    q[..., 0] = rho
    q[..., 1] = rho * u
    q[..., 2] = 0.0
    q[..., 3] = p / 0.4 + 0.5 * rho * u ** 2
    q[..., 4] = rho * nt
    wd = np.full((case.nic, case.njc), dval)
    bb = np.full((case.nic, case.njc), beta)
    r, _ = of.residual(case, q, wd=wd, beta=bb)
    T = p / rho
    mu = case.fs["mu_inf"] * (T / case.fs["T_inf"]) ** (2 / 3)
    expect = sa_source_closed_form(rho, nt, mu, a, dval, beta)
    inner = (slice(2, case.nic - 2), slice(2, case.njc - 2))
    got = (r[..., 4] - nt * r[..., 0])[inner]
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-12 * abs(expect))


@pytest.mark.parametrize("k", [0, 1])
def test_sa_jv_dual_matches_fd(k):
    case = [cases.sa_plate(21, 15, dy1=1e-3), cases.sa_bump(21, 13, dy1=1e-3)][k]
    q = case.perturbed_state(seed=8)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
    wd = of.wall_distance(case)
    _, jv = of.jv(case, q, v, wd=wd)
    h = 1e-7
    rp, _ = of.residual(case, q + h * v, wd=wd)
    rm, _ = of.residual(case, q - h * v, wd=wd)
    fd = (rp - rm) / (2 * h)
    for m in range(5):
        scale = np.abs(fd[..., m]).max()
        assert np.abs(jv[..., m] - fd[..., m]).max() <= 1e-5 * scale, m


def test_sa_freestream_preservation_without_walls():
    """Uniform flow, no walls, d -> large: every residual is zero up to
    rounding (the SA source vanishes for omega = 0 only through destruction,
    which is O(1/d^2) -> use a huge wall distance)."""
    case, _ = shear_case(a=0.0)
    q = case.freestream_state()
    wd = np.full((case.nic, case.njc), 1e12)
    r, _ = of.residual(case, q, wd=wd)
    assert np.abs(r).max() < 1e-12
