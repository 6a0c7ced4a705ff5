"""Writes the golden fixtures under tests/golden/ by running the compiled
reference (oracle/_ref/ref_harness, built by `make -C oracle ref` from
/root/reference) on small grids and seeded states.

Each fixture .npz holds the case inputs (grid, freestream, ordered boundary
list, options, state, J.v direction) and the reference's outputs:
  R        calc_residual                       (src/model/eulerequation.cpp:202)
  dt       calc_dt at cfl=3                    (:237)
  grads    grad_{u,v,T}_{chi,eta}              (src/utils/mesh.cpp:34)
  surface  write_surface text (x, Cp, Cf)      (src/utils/io.cpp:181)
  jv_fd    central difference of the reference residual along v
           (h = 1e-6 * |q|-scaled), a derivative check of the tangent.
explicit_*.npz hold the reference's Solver::solve result for the explicit
schemes (forward Euler, RK4-Jameson) from a fixture's state (explicit.json).
naca0012_281x151.npz holds the reference's outputs on its own NACA 0012
example grid (naca_fixture).
ref_scalars.json holds the reference's printed force coefficients and its
initial state per viscous fixture (ref_scalars).
Run:  python tests/golden/make_golden.py [case names]
      python tests/golden/make_golden.py explicit
      python tests/golden/make_golden.py naca
      python tests/golden/make_golden.py scalars
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from structured_amd import cases, grids  # noqa: E402
import oracle_ffi as of  # noqa: E402


def wake_case():
    """Small C-grid-like topology exercising the wake mirror: bottom wake on
    ghosted 1..8 (mirrors onto 17..24), wall on 9..16, freestream around."""
    xv, yv, _ = grids.flat_plate(25, 11, 1e-2)
    bcs = [dict(type="freestream", face="top", start=0, end=-1),
           dict(type="freestream", face="left", start=0, end=-1),
           dict(type="freestream", face="right", start=0, end=-1),
           dict(type="wake", face="bottom", start=1, end=8),
           dict(type="wall", face="bottom", start=9, end=16, u=0.0, v=0.0)]
    fs = dict(rho_inf=1.0, u_inf=0.5, v_inf=0.01, p_inf=1 / 1.4, T_inf=1 / 1.4, mu_inf=1e-3, pr_inf=0.7,
              aoa_deg=1.0, nut_inf=0.0)
    return cases.Case("wake_25x11", xv, yv, fs, bcs, order=2, turbulence=0, tail=9)


def bump_iso_case():
    xv, yv = grids.bump_channel(33, 17, 1e-3, 0.05, 0.2)
    bcs = [dict(type="isothermalwall", face="bottom", start=1, end=-2, u=0.0, v=0.0, T=0.75),
           dict(type="slipwall", face="top", start=1, end=-2),
           dict(type="freestream", face="left", start=0, end=-1),
           dict(type="outflow", face="right", start=0, end=-1)]
    fs = dict(rho_inf=1.0, u_inf=0.2, v_inf=0.0, p_inf=1 / 1.4, T_inf=1 / 1.4, mu_inf=2e-5, pr_inf=0.72,
              aoa_deg=0.0, nut_inf=0.0)
    return cases.Case("bump_iso_33x17", xv, yv, fs, bcs, order=2, turbulence=0, tail=1)


def ausm_plate_case():
    c = cases.laminar_plate(25, 17)
    c.flux = "ausm"
    c.name = "laminar_plate_25x17_ausm"
    return c


def ausm_supersonic_case():
    """Inviscid Mach 1.6 bump channel: AUSM faces on both branches of the
    split Mach numbers (|M| > 1 along i, |M| < 1 along j) and both signs of
    the interface Mach number."""
    xv, yv = grids.bump_channel(33, 17, 1e-2, 0.04, 0.2)
    bcs = [dict(type="slipwall", face="bottom", start=1, end=-2),
           dict(type="slipwall", face="top", start=1, end=-2),
           dict(type="freestream", face="left", start=0, end=-1),
           dict(type="outflow", face="right", start=0, end=-1)]
    fs = dict(rho_inf=1.0, u_inf=1.6, v_inf=0.0, p_inf=1 / 1.4, T_inf=1 / 1.4, mu_inf=0.0, pr_inf=0.72,
              aoa_deg=0.0, nut_inf=0.0)
    c = cases.Case("inviscid_bump_ausm_33x17", xv, yv, fs, bcs, order=2, turbulence=0, tail=1)
    c.flux = "ausm"
    return c


def all_cases():
    out = [cases.channel_example(), cases.laminar_plate(25, 17), wake_case(), bump_iso_case()]
    fo = cases.laminar_plate(21, 13)
    fo.order = 1
    fo.name = "laminar_plate_21x13_o1"
    out.append(fo)
    inv = cases.laminar_plate(21, 13)
    inv.fs = dict(inv.fs, mu_inf=0.0)
    inv.name = "inviscid_plate_21x13"
    out.append(inv)
    out += [ausm_plate_case(), ausm_supersonic_case()]
    return out


def main(only=None):
    """Writes every fixture, or only those named in `only` (the other entries
    of index.json are kept)."""
    if not of.have_ref():
        sys.exit("reference harness missing: make -C oracle ref")
    index = {}
    if only:
        with open(os.path.join(HERE, "index.json")) as f:
            index = json.load(f)
    for case in all_cases():
        if only and case.name not in only:
            continue
        q = case.perturbed_state(seed=7)
        rng = np.random.default_rng(11)
        v = rng.standard_normal(q.shape) * np.abs(q).max(axis=(0, 1))
        ref = of.RefRun(case)
        R = ref.residual(q)
        dt = ref.dt(q, 3.0)
        surf = ref.surface(q) if case.fs["mu_inf"] > 1e-15 else np.zeros((0, 3))
        grads = ref.grads(q) if case.fs["mu_inf"] > 1e-15 else [np.zeros(0)] * 6
        h = 1e-6
        jv_fd = (ref.residual(q + h * v) - ref.residual(q - h * v)) / (2 * h)
        path = os.path.join(HERE, case.name + ".npz")
        np.savez_compressed(path, xv=case.xv, yv=case.yv, q=q, v=v, R=R, dt=dt, surface=surf, jv_fd=jv_fd,
                            gu_c=grads[0], gv_c=grads[1], gT_c=grads[2], gu_e=grads[3], gv_e=grads[4],
                            gT_e=grads[5])
        index[case.name] = dict(fs=case.fs, bcs=case.bcs, order=case.order, turbulence=case.turbulence,
                                tail=case.tail, dpdx=case.dpdx, dpdy=case.dpdy, cfl=3.0, fd_h=h)
        if case.flux != "roe":
            index[case.name]["flux"] = case.flux
        print(case.name, "max|R| %.3e" % np.abs(R).max())
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


# Explicit pseudo-time runs of the reference's own Solver::solve
# (src/solver/solver.cpp:103-121,197-236) from a fixture's state: name ->
# (base fixture, [solver] keys).  The channel run uses the reference's
# examples/laminar/channel/explicit.inp settings with fewer iterations.
EXPLICIT = {
    "explicit_rk4_channel_3x101": ("channel_3x101", dict(scheme="rk4_jameson", cfl=0.1, iteration_max=20)),
    "explicit_fe_ramp_laminar_plate_25x17": ("laminar_plate_25x17",
                                             dict(scheme="forward_euler", cfl=0.8, iteration_max=3, cfl_ramp=True,
                                                  cfl_ramp_iteration=1, cfl_ramp_exponent=1.2)),
    "explicit_rk4_laminar_plate_25x17_ausm": ("laminar_plate_25x17_ausm",
                                              dict(scheme="rk4_jameson", cfl=1.5, iteration_max=3)),
}


def explicit_fixtures():
    """Writes explicit.json and explicit_*.npz (q0 = the base fixture's
    state, q = the reference's state after Solver::solve)."""
    import golden_util as gu
    index = {}
    for name, (base, solver) in EXPLICIT.items():
        case, z, _ = gu.load(base)
        ref = of.RefRun(case)
        q = ref.solve(z["q"], **solver)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), q=q)
        if name in TECPLOT:
            # the reference's own Tecplot and restart files of the final state
            # (IOManager::write at the last step, src/solver/solver.cpp:136-143)
            shutil.copy(os.path.join(ref.dir, "solve.tec"), os.path.join(HERE, name + ".tec"))
            with open(os.path.join(ref.dir, "solve.out"), "rb") as f:
                assert np.array_equal(np.frombuffer(f.read(), dtype=np.float64), q.reshape(-1))
        index[name] = dict(base=base, solver=solver)
        print(name, "max|dq| %.3e" % np.abs(q - z["q"]).max())
    with open(os.path.join(HERE, "explicit.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


TECPLOT = ("explicit_fe_ramp_laminar_plate_25x17",)


NACA_GRID = "/root/reference/examples/laminar/naca0012/grid.unf2"
NACA_SUB = 3  # every third cell in i and j of R and dt is kept


def naca_fixture():
    """naca0012_281x151.npz: the reference's own NACA 0012 example (its grid
    file and config values, cases.naca0012_example) on a seeded perturbed
    state (seed 7): the grid, the reference's R and dt at cfl 3 on every
    third cell, and its surface file."""
    from structured_amd import lib
    xv, yv = lib.read_grid(NACA_GRID, "simple", 281, 151)
    case = cases.naca0012_example(xv, yv)
    q = case.perturbed_state(seed=7)
    ref = of.RefRun(case)
    R = ref.residual(q)
    dt = ref.dt(q, 3.0)
    surf = ref.surface(q)
    np.savez_compressed(os.path.join(HERE, "naca0012_281x151.npz"), xv=xv, yv=yv, seed=7, sub=NACA_SUB,
                        R_sub=R[::NACA_SUB, ::NACA_SUB], dt_sub=dt[::NACA_SUB, ::NACA_SUB], surface=surf,
                        l2=np.sqrt((R ** 2).sum(axis=(0, 1))))
    print("naca0012_281x151 max|R| %.3e" % np.abs(R).max())


def ref_scalars():
    """ref_scalars.json: per viscous fixture (and the NACA 0012 example),
    the reference's six force coefficients on the fixture's state as its
    write_surface prints them (src/utils/io.cpp:239-251), and the state its
    EulerEquation::initialize sets (one cell: every cell is the same,
    checked here; src/model/eulerequation.cpp:262-277)."""
    import golden_util as gu
    out = {}
    todo = [(n,) + gu.load(n)[:2] for n in gu.names()]
    from structured_amd import lib
    xv, yv = lib.read_grid(NACA_GRID, "simple", 281, 151)
    naca = cases.naca0012_example(xv, yv)
    todo.append(("naca0012_281x151", naca, {"q": naca.perturbed_state(seed=7)}))
    for name, case, z in todo:
        if case.fs["mu_inf"] <= 1e-15:
            continue
        ref = of.RefRun(case)
        f = ref.forces(z["q"])
        q0 = ref.init()
        assert np.all(q0 == q0[0, 0]), name
        out[name] = dict(forces=[float(x) for x in f], init=[float(x) for x in q0[0, 0]])
        print(name, "forces", f)
    with open(os.path.join(HERE, "ref_scalars.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["scalars"]:
        ref_scalars()
    elif sys.argv[1:] == ["explicit"]:
        explicit_fixtures()
    elif sys.argv[1:] == ["naca"]:
        naca_fixture()
    else:
        main(sys.argv[1:])
