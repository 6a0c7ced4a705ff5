"""Value pins for the oracle functions that the golden residual fixtures do
not reach.

- Wall force integrals: against the coefficients the compiled reference
  prints after writing its surface file (src/utils/io.cpp:239-251), kept in
  tests/golden/ref_scalars.json to the 6 digits it prints.
- Initial state: against the reference's EulerEquation::initialize
  (src/model/eulerequation.cpp:262-277), same file.
- Spalart-Allmaras transport (no counterpart in the reference, so these are
  closed forms on uniform Cartesian grids where the discrete operators are
  exact): the nu~ diffusion (1/sigma) div(eta grad nu~) with eta = mu + rho
  nu~, the c_b2 term, the -(1/sigma)(nu + nu~) grad rho . grad nu~ term; the
  eddy viscosity mu_t = rho nu~ f_v1 in the momentum and energy viscous
  fluxes and k = c_p (mu/Pr + mu_t/0.9) in the heat flux; the nu~ ghost
  values of every boundary type; the wall distance.
Each test also checks that the term it pins is a sizeable part of the
result, so that dropping it or flipping its sign fails."""
import json
import os

import numpy as np
import pytest

import golden_util as gu
import oracle_ffi as of
from structured_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))
SIG, CB2, CV1, PRT, GAMMA = 2.0 / 3.0, 0.622, 7.1, 0.9, 1.4

with open(os.path.join(HERE, "golden", "ref_scalars.json")) as _f:
    REF_SCALARS = json.load(_f)


def _printed_close(got, printed):
    """got equals a value the reference printed with 6 significant digits."""
    got, printed = np.asarray(got, float), np.asarray(printed, float)
    tol = 5.0001e-6 * np.abs(printed) + 1e-15 * np.abs(got).max()
    return np.all(np.abs(got - printed) <= tol)


def _case_and_state(name):
    if name == "naca0012_281x151":
        case, q, _ = gu.load_naca()
        return case, q
    case, z, _ = gu.load(name)
    return case, z["q"]


@pytest.mark.parametrize("name", sorted(REF_SCALARS))
def test_oracle_forces_match_reference_printout(name):
    case, q = _case_and_state(name)
    _, _, _, forces = of.surface(case, q)
    ref = REF_SCALARS[name]["forces"]
    assert _printed_close(forces, ref), (forces, ref)


def test_forces_printout_discriminates():
    """A sign flip of the viscous lift or a swapped pressure/viscous pair
    misses the printed values."""
    case, q = _case_and_state("bump_iso_33x17")
    _, _, _, f = of.surface(case, q)
    ref = REF_SCALARS["bump_iso_33x17"]["forces"]
    flipped = f.copy()
    flipped[2] = -flipped[2]
    assert not _printed_close(flipped, ref)
    assert not _printed_close(f[[2, 3, 0, 1, 4, 5]], ref)


@pytest.mark.parametrize("name", sorted(REF_SCALARS))
def test_freestream_state_matches_reference_initialize(name):
    case, _ = _case_and_state(name)
    q0 = case.freestream_state()
    assert np.array_equal(q0[..., :4], np.broadcast_to(REF_SCALARS[name]["init"], q0[..., :4].shape))


# ---------------------------------------------------------------- SA pins --
def cartesian_case(nx=14, ny=12, lx=1.0, ly=0.8, mu=1e-3, nut_inf=3e-3, bcs=None):
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    xv = np.repeat(xs[:, None], ny + 1, axis=1)
    yv = np.repeat(ys[None, :], nx + 1, axis=0)
    fs = dict(rho_inf=1.0, u_inf=0.2, v_inf=0.0, p_inf=1 / 1.4, T_inf=1 / 1.4, mu_inf=mu, pr_inf=0.72,
              aoa_deg=0.0, nut_inf=nut_inf)
    if bcs is None:
        bcs = [dict(type="freestream", face=f, start=0, end=-1) for f in ("bottom", "top", "left", "right")]
    c = cases.Case("cart", xv, yv, fs, bcs, order=2, turbulence=1, tail=1)
    yc = 0.25 * (yv[:-1, :-1] + yv[1:, :-1] + yv[:-1, 1:] + yv[1:, 1:])
    return c, yc, ly / ny


def fv1(chi):
    return chi ** 3 / (chi ** 3 + CV1 ** 3)


def state(case, rho, u, T, nut):
    """Conservative AoS state from cell primitives (R = 1, v = 0)."""
    q = np.zeros((case.nic, case.njc, 5))
    p = rho * T
    q[..., 0] = rho
    q[..., 1] = rho * u
    q[..., 3] = p / (GAMMA - 1.0) + 0.5 * rho * u ** 2
    q[..., 4] = rho * nut
    return q


INNER = (slice(2, -2), slice(2, -2))


def test_sa_diffusion_cb2_and_density_gradient_terms():
    """nu~ = n0 + b y and rho = r0 + c y at rest with uniform T (so mu is
    constant and the flow carries no mass flux) and d -> infinity (no
    production or destruction): the rho nu~ residual is
        (1/sigma) b d/dy[rho nu~]          (diffusion of eta = mu + rho nu~)
      + (c_b2/sigma) rho b^2                (c_b2 term)
      - (1/sigma) (nu + nu~) c b            (density-gradient term)."""
    case, yc, _ = cartesian_case()
    n0, b, r0, c, T0 = 2e-3, 4e-3, 1.0, 0.5, 1 / 1.4
    rho, nut = r0 + c * yc, n0 + b * yc
    q = state(case, rho, 0.0 * yc, T0, nut)
    wd = np.full((case.nic, case.njc), 1e12)
    r, _ = of.residual(case, q, wd=wd)
    mu = case.fs["mu_inf"] * (T0 / case.fs["T_inf"]) ** (2 / 3)
    diff = b * (c * nut + b * rho) / SIG
    cb2 = CB2 / SIG * rho * b * b
    cross = -(mu / rho + nut) * c * b / SIG
    expect = diff + cb2 + cross
    np.testing.assert_allclose(r[..., 4][INNER], expect[INNER], rtol=1e-10, atol=0)
    # each term moves the result by far more than the tolerance
    for t in (diff, cb2, cross):
        assert np.all(np.abs(t[INNER]) > 0.2 * np.abs(expect[INNER]))


def test_sa_eddy_viscosity_in_momentum_and_energy():
    """Shear u = a y with nu~ = n0 + b y (uniform rho, p, T): the x-momentum
    and energy residuals are the divergence of mu_eff u_y and u mu_eff u_y
    with mu_eff = mu + rho nu~ f_v1(rho nu~ / mu) averaged onto the faces:
        R1 = a (mu_n - mu_s) / dy,  R3 = a (u_n mu_n - u_s mu_s) / dy."""
    case, yc, dy = cartesian_case()
    a, n0, b, T0 = 0.3, 2e-3, 1.2e-2, 1 / 1.4
    nut = n0 + b * yc
    q = state(case, np.ones_like(yc), a * yc, T0, nut)
    wd = np.full((case.nic, case.njc), 1e12)
    r, _ = of.residual(case, q, wd=wd)
    mu = case.fs["mu_inf"]
    meff = mu + nut * fv1(nut / mu)
    mn = 0.5 * (meff[:, 1:-1] + meff[:, 2:])
    ms = 0.5 * (meff[:, 1:-1] + meff[:, :-2])
    un, us = a * (yc[:, 1:-1] + 0.5 * dy), a * (yc[:, 1:-1] - 0.5 * dy)
    r1 = a * (mn - ms) / dy
    r3 = a * (un * mn - us * ms) / dy
    inner = (slice(2, -2), slice(1, -1))
    np.testing.assert_allclose(r[..., 1][2:-2, 2:-2], r1[inner], rtol=1e-9, atol=0)
    np.testing.assert_allclose(r[..., 3][2:-2, 2:-2], r3[inner], rtol=1e-9, atol=0)
    assert np.abs(r[..., 2][2:-2, 2:-2]).max() < 1e-9 * np.abs(r1).max()
    # without mu_t, R1 would vanish (mu is uniform)
    assert np.abs(r1[inner]).min() > 0.1 * np.abs(r1[inner]).max()


def test_sa_eddy_conductivity():
    """At rest with uniform p and T = T0 + t y (rho = p / T) and uniform
    nu~: the energy residual is the heat-flux divergence
        R3 = t (k_n - k_s) / dy,  k = c_p (mu / Pr + mu_t / 0.9),
    mu = mu_inf (T / T_inf)^(2/3), mu_t = rho nu~ f_v1."""
    case, yc, dy = cartesian_case(mu=2e-3)
    T0, t, p0, n0 = 0.7, 0.4, 1 / 1.4, 6e-3
    T = T0 + t * yc
    rho = p0 / T
    q = state(case, rho, 0.0 * yc, T, np.full_like(yc, n0))
    wd = np.full((case.nic, case.njc), 1e12)
    r, _ = of.residual(case, q, wd=wd)
    fs = case.fs
    mu = fs["mu_inf"] * (T / fs["T_inf"]) ** (2 / 3)
    mut = rho * n0 * fv1(rho * n0 / mu)
    cp = GAMMA / (GAMMA - 1.0)
    k = cp * (mu / fs["pr_inf"] + mut / PRT)
    kn = 0.5 * (k[:, 1:-1] + k[:, 2:])
    ks = 0.5 * (k[:, 1:-1] + k[:, :-2])
    r3 = t * (kn - ks) / dy
    np.testing.assert_allclose(r[..., 3][2:-2, 2:-2], r3[2:-2, 1:-1], rtol=1e-9, atol=0)
    # the turbulent part is a sizeable share of the conductivity difference
    kl = cp * mu / fs["pr_inf"]
    r3_lam = t * (0.5 * (kl[:, 1:-1] + kl[:, 2:]) - 0.5 * (kl[:, 1:-1] + kl[:, :-2])) / dy
    assert np.all(np.abs(r3 - r3_lam)[2:-2, 1:-1] > 0.2 * np.abs(r3)[2:-2, 1:-1])


def _ghost_state(case, seed):
    rng = np.random.default_rng(seed)
    shape = (case.nic, case.njc)
    rho = 1.0 + 0.1 * rng.random(shape)
    u = 0.2 + 0.05 * rng.random(shape)
    T = 0.7 + 0.05 * rng.random(shape)
    nut = 1e-3 * (1.0 + rng.random(shape))
    q = state(case, rho, u, T, nut)
    return q, q[..., 4] / q[..., 0]  # nu~ as primvars recovers it


def test_sa_ghost_values_walls_slip_freestream_outflow():
    """nu~ ghosts: no-slip and isothermal walls -(1.5 a - 0.5 b) (a, b the
    first two cells inward), slip wall 1.5 a - 0.5 b, freestream nu~_inf,
    outflow the adjacent cell."""
    bcs = [dict(type="wall", face="bottom", start=1, end=6, u=0.0, v=0.0),
           dict(type="isothermalwall", face="bottom", start=7, end=-2, u=0.0, v=0.0, T=0.7),
           dict(type="slipwall", face="top", start=1, end=-2),
           dict(type="freestream", face="left", start=0, end=-1),
           dict(type="outflow", face="right", start=0, end=-1)]
    case, _, _ = cartesian_case(bcs=bcs)
    q, nut = _ghost_state(case, 3)
    g = of.ghosts(case, q)[5]
    nic, njc = case.nic, case.njc
    ext_b = 1.5 * nut[:, 0] - 0.5 * nut[:, 1]
    ext_t = 1.5 * nut[:, -1] - 0.5 * nut[:, -2]
    np.testing.assert_array_equal(g[1:nic + 1, 0], -ext_b)
    np.testing.assert_array_equal(g[1:nic + 1, njc + 1], ext_t)
    np.testing.assert_array_equal(g[0, :], case.fs["nut_inf"])
    np.testing.assert_array_equal(g[nic + 1, 1:njc + 1], nut[-1, :])
    np.testing.assert_array_equal(g[1:nic + 1, 1:njc + 1], nut)


def test_sa_ghost_values_periodic_and_wake():
    """nu~ ghosts: periodic copies the opposite interior column; the wake cut
    mirrors the first row about the cut (ghost i <- cell nic + 1 - i)."""
    bcs = [dict(type="periodic", face="left", start=0, end=-1),
           dict(type="wake", face="bottom", start=1, end=5),
           dict(type="wall", face="bottom", start=6, end=9, u=0.0, v=0.0),
           dict(type="slipwall", face="top", start=1, end=-2)]
    case, _, _ = cartesian_case(bcs=bcs)
    q, nut = _ghost_state(case, 4)
    g = of.ghosts(case, q)[5]
    nic, njc = case.nic, case.njc
    np.testing.assert_array_equal(g[0, 1:njc + 1], nut[-1, :])
    np.testing.assert_array_equal(g[nic + 1, 1:njc + 1], nut[0, :])
    for i in range(1, 6):
        m = nic + 1 - i
        assert g[i, 0] == nut[m - 1, 0]
        assert g[m, 0] == nut[i - 1, 0]


def test_wall_distance_flat_plate_closed_form():
    """Flat plate with the no-slip wall from x = 0: d = y_c above the plate,
    the distance to the leading edge (0, 0) upstream of it."""
    case = cases.sa_plate(41, 21, dy1=1e-3)
    d = of.wall_distance(case)
    xc = 0.25 * (case.xv[:-1, :-1] + case.xv[1:, :-1] + case.xv[:-1, 1:] + case.xv[1:, 1:])
    yc = 0.25 * (case.yv[:-1, :-1] + case.yv[1:, :-1] + case.yv[:-1, 1:] + case.yv[1:, 1:])
    expect = np.where(xc >= 0.0, yc, np.hypot(xc, yc))
    np.testing.assert_allclose(d, expect, rtol=1e-14, atol=0)
    assert np.any(xc < 0) and np.any(xc > 0)


def test_wall_distance_bump_brute_force():
    """Bump channel: the minimum over the lower-wall edges of the
    point-to-segment distance, by brute force."""
    case = cases.sa_bump(41, 21, dy1=1e-3)
    d = of.wall_distance(case)
    xc = 0.25 * (case.xv[:-1, :-1] + case.xv[1:, :-1] + case.xv[:-1, 1:] + case.xv[1:, 1:])
    yc = 0.25 * (case.yv[:-1, :-1] + case.yv[1:, :-1] + case.yv[:-1, 1:] + case.yv[1:, 1:])
    ax, ay = case.xv[:-1, 0], case.yv[:-1, 0]
    bx, by = case.xv[1:, 0], case.yv[1:, 0]
    px, py = xc.reshape(-1, 1), yc.reshape(-1, 1)
    ex, ey = bx - ax, by - ay
    t = np.clip(((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey), 0.0, 1.0)
    dist = np.hypot(px - (ax + t * ex), py - (ay + t * ey)).min(axis=1).reshape(xc.shape)
    np.testing.assert_allclose(d, dist, rtol=1e-12, atol=0)
    # the bump makes it differ from the vertical distance to y = 0
    assert np.abs(dist - yc).max() > 1e-3
