"""Adjoint beta-gradient at the adjoint configuration's size against finite
differences of re-converged solves: the 705x321 SA channel bump (Re 3e6,
first spacing 1e-6), beta = 1 + 8 seeded blobs, the target wall Cf from a
second seeded beta*, F = sum_i ds_i (Cf_i - Cf*_i)^2, 16 seeded cells.

The gradient g comes from one adjoint solve (J^T psi = dF/dq, dF/dbeta =
-psi^T dR/dbeta, GMRES to 1e-12).  Around each of 16 seeded cells of the
sensitive region (|g| >= 1e-2 max|g|) the direction d is a Gaussian bump of
beta of radius three cells there, and the adjoint's directional derivative
g . d is compared with central differences of F(beta + e d) re-converged to
the roundoff floor at e = +-0.05, +-0.1, Richardson-extrapolated (O(e^4)).
The bar is 1e-6 relative.  Where the perturbation moves a face across a kink
of the discretisation (the Roe |u_n| where u_n ~ 0, the limiter's branches)
the left and right derivatives differ; there g . d, which takes one branch,
must lie between the one-sided differences instead.

Why directions and not single cells: F is a sum of squared differences of
two nearby Cf distributions, so its value is resolved only to about 1e-9
relative (printed below as the noise floor: the same beta re-converged from
another state); one cell of this grid moves F by about 1e-5 relative, too
little for a finite difference to resolve 1e-6, while a bump of ~30 cells
moves it by enough.  The 16 directions run as four tests of four and a
summary test."""
import numpy as np
import pytest

from structured_amd import cases, fiml, lib

pytestmark = pytest.mark.gpu

RESOLVE = dict(fiml.WARM, tolerance=1e-17, stall_window=4, iteration_max=40)
GROUPS = 4        # the 16 directions run as 4 tests of 4, so that each test ends within a minute or so
_RESULTS = {}


@pytest.fixture(scope="module")
def base():
    case = cases.sa_bump(705, 321)
    eq = lib.Equation(case)
    q1, r1, _ = eq.solve(case.freestream_state(), **fiml.NEWTON)
    assert r1["converged"], r1
    eq.set_beta(cases.blob_beta(case, seed=2))
    qs, r2, _ = eq.solve(q1, **fiml.WARM)
    assert r2["converged"], r2
    _, _, cf_target, _ = eq.wall(qs)
    beta = cases.blob_beta(case, seed=1)
    eq.set_beta(beta)
    q, r3, _ = eq.solve(q1, **RESOLVE)
    print("base state: newton", r3["iterations"], "l2w", max(r3["l2w_final"]), flush=True)
    g, F0, arep = eq.adjoint_gradient(q, cf_target, **dict(fiml.ADJOINT, rtol=1e-12, restart=120, maxit=6000))
    print("F", F0, "adjoint", arep, flush=True)
    assert arep["converged"]
    # noise floor of F: the same beta re-converged from another state
    qa, _, _ = eq.solve(qs, **RESOLVE)
    noise = abs(eq.objective_cf(qa, cf_target, grad=False)[0] - F0)
    print("F noise floor %.3e (relative %.2e)" % (noise, noise / F0), flush=True)
    gmax = np.abs(g).max()
    sens = np.argwhere(np.abs(g) >= 1e-2 * gmax)
    rng = np.random.default_rng(16)
    cells = [tuple(int(v) for v in sens[k]) for k in rng.choice(len(sens), 16, replace=False)]
    yield dict(case=case, eq=eq, q=q, beta=beta, g=g, F0=F0, noise=noise, cells=cells, cf_target=cf_target)
    eq.close()


@pytest.mark.parametrize("group", range(GROUPS))
def test_adjoint_gradient_fd_705x321_directions(base, group):
    eq, case, beta, g, F0, q = base["eq"], base["case"], base["beta"], base["g"], base["F0"], base["q"]
    ii, jj = np.meshgrid(np.arange(case.nic), np.arange(case.njc), indexing="ij")

    def F_at(d, e):
        eq.set_beta(beta + e * d)
        qp, rp, _ = eq.solve(q, **RESOLVE)
        return eq.objective_cf(qp, base["cf_target"], grad=False)[0], rp

    n = len(base["cells"]) // GROUPS
    for (i, j) in base["cells"][group * n:(group + 1) * n]:
        d = np.exp(-((ii - i) ** 2 + (jj - j) ** 2) / 9.0)
        d[d < 1e-6] = 0.0
        gd = float(np.sum(g * d))
        e = 0.05
        Fp1, rp = F_at(d, e)
        Fm1, _ = F_at(d, -e)
        Fp2, _ = F_at(d, 2 * e)
        Fm2, _ = F_at(d, -2 * e)
        c1, c2 = (Fp1 - Fm1) / (2 * e), (Fp2 - Fm2) / (4 * e)
        fd = (4.0 * c1 - c2) / 3.0
        rel = abs(gd - fd) / abs(fd)
        fwd, bwd = (Fp1 - F0) / e, (F0 - Fm1) / e
        bracketed = min(fwd, bwd) <= gd <= max(fwd, bwd)
        _RESULTS[(i, j)] = (rel, bracketed)
        print("cell", (i, j), "g.d %.12e fd %.12e rel %.2e | one-sided %.9e %.9e bracketed %s | dF/F %.1e "
              "fd noise bound %.1e | resolve %d" % (gd, fd, rel, fwd, bwd, bracketed, abs(Fp1 - F0) / F0,
                                                  1.5 * base["noise"] / e / abs(fd), rp["iterations"]), flush=True)
        if rel > 1e-6:
            assert bracketed, ((i, j), rel)
    eq.set_beta(beta)


def test_adjoint_gradient_fd_705x321_summary(base):
    """Over the 16 directions: median relative difference <= 1e-6 and at
    most 4 directions across a kink (each bracketed, checked above)."""
    assert len(_RESULTS) == len(base["cells"]), "the direction groups must run first"
    rels = [r for r, _ in _RESULTS.values()]
    kinks = [c for c, (r, _) in _RESULTS.items() if r > 1e-6]
    print("median rel %.2e max rel %.2e; directions across a kink (bracketed): %s" % (
        np.median(rels), max(rels), kinks))
    assert np.median(rels) <= 1e-6
    assert len(kinks) <= 4
