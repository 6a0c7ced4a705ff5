"""ctypes mirrors of the plain-data structs in include/structured_gpu.h."""
from __future__ import annotations

import ctypes as C

BC_TYPES = {"freestream": 0, "outflow": 1, "slipwall": 2, "wall": 3, "isothermalwall": 4, "periodic": 5, "wake": 6}
FACES = {"bottom": 0, "right": 1, "top": 2, "left": 3}  # src/model/bc.h:8-11


class Freestream(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("rho_inf", "u_inf", "v_inf", "p_inf", "T_inf", "mu_inf", "pr_inf", "aoa_deg", "nut_inf")]


class Bc(C.Structure):
    _fields_ = [("type", C.c_int), ("face", C.c_int), ("start", C.c_longlong), ("end", C.c_longlong),
                ("u", C.c_double), ("v", C.c_double), ("T", C.c_double)]


FLUXES = {"roe": 0, "ausm": 1}


class Opts(C.Structure):
    _fields_ = [("order", C.c_int), ("turbulence", C.c_int), ("tail", C.c_int), ("flux", C.c_int),
                ("dpdx", C.c_double), ("dpdy", C.c_double)]


def make_bcs(bcs):
    """bcs: list of dicts with keys type, face, start, end (+ u, v, T)."""
    arr = (Bc * max(1, len(bcs)))()
    for k, b in enumerate(bcs):
        arr[k].type = BC_TYPES[b["type"]]
        arr[k].face = FACES[b["face"]]
        arr[k].start = int(b.get("start", 0))
        arr[k].end = int(b.get("end", 0))
        arr[k].u = float(b.get("u", 0.0))
        arr[k].v = float(b.get("v", 0.0))
        arr[k].T = float(b.get("T", 0.0))
    return arr


def dptr(a):
    """double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


class NewtonOpts(C.Structure):
    _fields_ = [("cfl", C.c_double), ("cfl_ramp", C.c_int), ("cfl_ramp_iteration", C.c_int),
                ("cfl_ramp_exponent", C.c_double), ("cfl_max", C.c_double), ("under_relaxation", C.c_double),
                ("under_relaxation_ramp", C.c_int), ("under_relaxation_ramp_iteration", C.c_int),
                ("under_relaxation_ramp_exponent", C.c_double), ("iteration_max", C.c_int), ("tolerance", C.c_double),
                ("gmres_restart", C.c_int), ("gmres_maxit", C.c_int), ("gmres_rtol", C.c_double), ("precond", C.c_int),
                ("ser", C.c_int), ("ser_exponent", C.c_double), ("verbose", C.c_int), ("pc_cfl_max", C.c_double),
                ("norm_weighted", C.c_int), ("stall_window", C.c_int),
                ("lhs_order", C.c_int), ("pc_width", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("linear_iterations", C.c_int),
                ("stalled", C.c_int), ("l2_initial", C.c_double * 5), ("l2_final", C.c_double * 5),
                ("l2w_final", C.c_double * 5), ("cfl_final", C.c_double), ("wall_s", C.c_double), ("residual_s", C.c_double),
                ("jacobian_s", C.c_double), ("linsolve_s", C.c_double)]


# Defaults of the reference's [solver] table (src/utils/config.cpp:44-73) where
# they exist; the Krylov settings and the tolerance test are this build's.
NEWTON_DEFAULTS = dict(cfl=1.0, cfl_ramp=0, cfl_ramp_iteration=20, cfl_ramp_exponent=1.1, cfl_max=1e12,
                       under_relaxation=1.0, under_relaxation_ramp=0, under_relaxation_ramp_iteration=20,
                       under_relaxation_ramp_exponent=1.1, iteration_max=1, tolerance=1e-10, gmres_restart=30,
                       gmres_maxit=60, gmres_rtol=1e-3, precond=4, ser=0, ser_exponent=1.0, verbose=0,
                       pc_cfl_max=0.0, norm_weighted=0, stall_window=0, lhs_order=0, pc_width=0)


def newton_opts(**kw):
    d = dict(NEWTON_DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError("unknown Newton options: %s" % sorted(unknown))
    d.update(kw)
    return NewtonOpts(**d)


SCHEMES = {"forward_euler": 0, "rk4_jameson": 1}


class ExplicitOpts(C.Structure):
    _fields_ = [("scheme", C.c_int), ("iteration_max", C.c_int), ("cfl", C.c_double), ("cfl_ramp", C.c_int),
                ("cfl_ramp_iteration", C.c_int), ("cfl_ramp_exponent", C.c_double), ("tolerance", C.c_double),
                ("verbose", C.c_int), ("reserved", C.c_int)]


def explicit_opts(scheme="forward_euler", iteration_max=1, cfl=1.0, cfl_ramp=False, cfl_ramp_iteration=20,
                  cfl_ramp_exponent=1.1, tolerance=0.0, verbose=0):
    """The reference's [solver] defaults (src/utils/config.cpp:46-61)."""
    return ExplicitOpts(scheme=SCHEMES[scheme] if isinstance(scheme, str) else int(scheme),
                        iteration_max=iteration_max, cfl=cfl, cfl_ramp=int(bool(cfl_ramp)),
                        cfl_ramp_iteration=cfl_ramp_iteration, cfl_ramp_exponent=cfl_ramp_exponent,
                        tolerance=tolerance, verbose=verbose, reserved=0)


class AdjointOpts(C.Structure):
    _fields_ = [("restart", C.c_int), ("maxit", C.c_int), ("rtol", C.c_double), ("precond", C.c_int),
                ("lhs_order", C.c_int), ("pc_width", C.c_int), ("pc_cfl", C.c_double)]


class AdjointReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("reorthogonalised", C.c_int), ("relres", C.c_double),
                ("wall_s", C.c_double)]


class InversionOpts(C.Structure):
    _fields_ = [("iterations", C.c_int), ("memory", C.c_int), ("reg", C.c_double), ("step0", C.c_double),
                ("beta_min", C.c_double), ("beta_max", C.c_double), ("max_backtracks", C.c_int),
                ("reserved", C.c_int), ("newton", NewtonOpts), ("adjoint", AdjointOpts)]


class InversionReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("evaluations", C.c_int), ("failed_solves", C.c_int),
                ("adjoint_unconverged", C.c_int),
                ("J0", C.c_double), ("J", C.c_double), ("F0", C.c_double), ("F", C.c_double),
                ("wall_s", C.c_double), ("primal_s", C.c_double), ("adjoint_s", C.c_double),
                ("adjoint_relres_max", C.c_double), ("adjoint_iterations", C.c_int), ("steepest_restarts", C.c_int)]
