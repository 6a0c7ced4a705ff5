"""Python mirror of the C ABI (include/structured_gpu.h) over ctypes.

``Equation`` mirrors the reference's EulerEquation surface used by the
Newton loop (src/model/eulerequation.h:53-55): ``calc_residual``,
``calc_dt`` plus the new tangent/adjoint products.  Arrays are numpy float64
in the reference layout q[i][j][k].  The HIP library must be built
(``python -m structured_amd.build``); without it every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstructured_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "structured_gpu.h")

ERRORS = {0: "SG_OK", -1: "SG_ERR_ARG", -2: "SG_ERR_BC", -3: "SG_ERR_NAN", -4: "SG_ERR_HIP", -5: "SG_ERR_NOCONV",
          -6: "SG_ERR_IO", -7: "SG_ERR_STATE"}

_lib = None


class SgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library missing: build it with `python -m structured_amd.build`")
        # torch wheels bundle their own HIP runtime under the same soname: load
        # torch first so that the process keeps a single runtime whichever of
        # the two touches the GPU first
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(LIB_PATH)
        _lib.sg_last_error.restype = C.c_char_p
        _lib.sg_last_error.argtypes = [C.c_void_p]
        _lib.sg_create.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p, C.POINTER(C.c_void_p)]
        _lib.sg_destroy.argtypes = [C.c_void_p]
        _lib.sg_destroy.restype = None
        for name in ("sg_nvar", "sg_ldi"):
            getattr(_lib, name).argtypes = [C.c_void_p]
    return _lib


def read_grid(path, fmt, ni, nj):
    """Mesh::simple_loader / plot3d_loader (src/utils/mesh.cpp:133-169):
    returns xv, yv [ni, nj]."""
    lib = load()
    xv, yv = np.empty((ni, nj)), np.empty((ni, nj))
    rc = lib.sg_read_grid(os.fsencode(path), fmt.encode(), ni, nj, _p(xv), _p(yv))
    if rc != 0:
        raise SgError(rc, lib.sg_last_error(None).decode())
    return xv, yv


def exported_symbols():
    """Function names declared in include/structured_gpu.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", text)))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


class Equation:
    """One grid + flow setup on the GPU (sg_create ... sg_destroy)."""

    def __init__(self, case):
        lib = load()
        self.case = case
        self.nic, self.njc, self.nvar = case.nic, case.njc, case.nvar
        xv = _f64(case.xv)
        yv = _f64(case.yv)
        fs = abi.Freestream(**case.fs)
        bcs = abi.make_bcs(case.bcs)
        o = abi.Opts(order=case.order, turbulence=case.turbulence, tail=case.tail, flux=abi.FLUXES[case.flux],
                     dpdx=case.dpdx, dpdy=case.dpdy)
        h = C.c_void_p()
        rc = lib.sg_create(case.ni, case.nj, _p(xv), _p(yv), C.byref(fs), C.cast(bcs, C.c_void_p), len(case.bcs),
                           C.byref(o), C.byref(h))
        if rc != 0:
            raise SgError(rc, lib.sg_last_error(None).decode())
        self.h = h
        self.ldi = lib.sg_ldi(h)
        self.plane = self.njc * self.ldi

    def close(self):
        if getattr(self, "h", None):
            load().sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise SgError(rc, load().sg_last_error(self.h).decode())

    @property
    def shape(self):
        return (self.nic, self.njc, self.nvar)

    # ---- reference-shaped operations (host buffers, AoS)
    def calc_residual(self, q):
        q = _f64(q, self.shape)
        r = np.empty_like(q)
        l2 = np.zeros(self.nvar)
        rc = load().sg_residual(self.h, _p(q), _p(r), _p(l2))
        if rc not in (0, -3):
            self._check(rc)
        return r, l2

    def jv(self, q, v):
        q = _f64(q, self.shape)
        v = _f64(v, self.shape)
        out = np.empty_like(q)
        self._check(load().sg_jv(self.h, _p(q), _p(v), _p(out)))
        return out

    def jtv(self, q, psi):
        q = _f64(q, self.shape)
        psi = _f64(psi, self.shape)
        out = np.empty_like(q)
        db = np.empty((self.nic, self.njc))
        self._check(load().sg_jtv(self.h, _p(q), _p(psi), _p(out), _p(db)))
        return out, db

    def calc_dt(self, q, cfl):
        q = _f64(q, self.shape)
        out = np.empty_like(q)
        self._check(load().sg_dt(self.h, _p(q), C.c_double(cfl), _p(out)))
        return out

    def wall(self, q):
        q = _f64(q, self.shape)
        nb = self.case.ni - 2 * self.case.tail + 1
        x, cp, cf, f = np.empty(nb), np.empty(nb), np.empty(nb), np.empty(6)
        self._check(load().sg_wall(self.h, _p(q), _p(x), _p(cp), _p(cf), _p(f)))
        return x, cp, cf, f

    # ---- files (src/utils/io.cpp writers, same layouts)
    def write_restart(self, path, q):
        q = _f64(q, self.shape)
        self._check(load().sg_write_restart(self.h, os.fsencode(path), _p(q)))

    def read_restart(self, path):
        q = np.empty(self.shape)
        self._check(load().sg_read_restart(self.h, os.fsencode(path), _p(q)))
        return q

    def write_surface(self, path, q):
        q = _f64(q, self.shape)
        f = np.empty(6)
        self._check(load().sg_write_surface(self.h, os.fsencode(path), _p(q), _p(f)))
        return f

    def write_tecplot(self, path, q):
        q = _f64(q, self.shape)
        self._check(load().sg_write_tecplot(self.h, os.fsencode(path), _p(q)))

    def write_npz(self, path, q):
        """<label>.npz of IOManager::write_npz (src/utils/io.cpp:103-126):
        xc, yc [nic, njc], q [nic, njc, nvar] and the primitives rho, u, v,
        p, T [nic, njc].  The reference stores its ghosted primitive arrays
        under an unghosted shape (their leading nic*njc entries); the cell
        values are written here instead."""
        q = _f64(q, self.shape)
        xv, yv = self.case.xv, self.case.yv
        xc = 0.25 * (xv[:-1, :-1] + xv[1:, :-1] + xv[:-1, 1:] + xv[1:, 1:])
        yc = 0.25 * (yv[:-1, :-1] + yv[1:, :-1] + yv[:-1, 1:] + yv[1:, 1:])
        r = q[..., 0]
        u, v = q[..., 1] / r, q[..., 2] / r
        p = (q[..., 3] - 0.5 * r * (u * u + v * v)) * 0.4
        T = p / r / (self.case.fs["p_inf"] / self.case.fs["rho_inf"] / self.case.fs["T_inf"])
        np.savez(path, xc=xc, yc=yc, q=q, rho=r, u=u, v=v, p=p, T=T)

    def set_beta(self, beta):
        beta = _f64(beta, (self.nic, self.njc))
        self._check(load().sg_set_beta(self.h, _p(beta)))

    def set_walldist(self, d):
        d = _f64(d, (self.nic, self.njc))
        self._check(load().sg_set_walldist(self.h, _p(d)))

    def walldist(self):
        d = np.empty((self.nic, self.njc))
        self._check(load().sg_get_walldist(self.h, _p(d)))
        return d

    # ---- device-resident (SoA) operations on raw device pointers
    def residual_dev_async(self, q_ptr: int, r_ptr: int, l2_ptr: int | None = None):
        self._check(load().sg_residual_dev_async(self.h, C.c_void_p(q_ptr), C.c_void_p(r_ptr),
                                                 None if l2_ptr is None else C.c_void_p(l2_ptr)))

    def stream(self) -> int:
        lib = load()
        lib.sg_get_stream.restype = C.c_void_p
        return int(lib.sg_get_stream(self.h) or 0)

    def residual_dev(self, q_ptr: int, r_ptr: int):
        lib = load()
        l2 = np.zeros(self.nvar)
        rc = lib.sg_residual_dev(self.h, C.c_void_p(q_ptr), C.c_void_p(r_ptr), _p(l2))
        if rc not in (0, -3):
            self._check(rc)
        return l2

    def jv_dev(self, q_ptr: int, v_ptr: int, out_ptr: int):
        self._check(load().sg_jv_dev(self.h, C.c_void_p(q_ptr), C.c_void_p(v_ptr), C.c_void_p(out_ptr)))

    def jv_dev_async(self, q_ptr: int, v_ptr: int, out_ptr: int):
        self._check(load().sg_jv_dev_async(self.h, C.c_void_p(q_ptr), C.c_void_p(v_ptr), C.c_void_p(out_ptr)))

    def initialize(self):
        """EulerEquation::initialize (src/model/eulerequation.cpp:262-277):
        the freestream state [nic, njc, nvar] (rho nu~_inf fifth with SA)."""
        q = np.empty(self.shape)
        self._check(load().sg_initialize(self.h, _p(q)))
        return q

    def synchronize(self):
        self._check(load().sg_synchronize(self.h))

    def aos_to_soa_dev(self, aos_ptr: int, soa_ptr: int, nvar: int | None = None):
        self._check(load().sg_aos_to_soa_dev(self.h, C.c_void_p(aos_ptr), C.c_void_p(soa_ptr), nvar or self.nvar))

    def soa_to_aos_dev(self, soa_ptr: int, aos_ptr: int, nvar: int | None = None):
        self._check(load().sg_soa_to_aos_dev(self.h, C.c_void_p(soa_ptr), C.c_void_p(aos_ptr), nvar or self.nvar))

    # ---- Newton-Krylov (Solver::solve) and the linear-solver interface
    def solve(self, q, **opts):
        """Pseudo-transient Newton-Krylov from q; returns (q, report dict,
        history dict).  Raises on NaN; a non-converged run is reported, not
        raised."""
        q = _f64(q, self.shape).copy()
        o = abi.newton_opts(**opts)
        rep = abi.SolveReport()
        rc = load().sg_solve(self.h, _p(q), C.byref(o), C.byref(rep))
        if rc not in (0, -5):
            self._check(rc)
        n = load().sg_solve_history(self.h, None, None, 0)
        l2h = np.zeros((max(n, 1), self.nvar))
        cfl = np.zeros(max(n, 1))
        ur = np.zeros(max(n, 1))
        load().sg_solve_history(self.h, _p(l2h), _p(cfl), n)
        load().sg_solve_ramp_history(self.h, None, _p(ur), n)
        report = {f: getattr(rep, f) for f, _ in abi.SolveReport._fields_}
        report["l2_initial"] = list(rep.l2_initial)[: self.nvar]
        report["l2_final"] = list(rep.l2_final)[: self.nvar]
        report["l2w_final"] = list(rep.l2w_final)[: self.nvar]
        report["code"] = rc
        return q, report, {"l2": l2h[:n], "cfl": cfl[:n], "ur": ur[:n]}

    def step_explicit_dev(self, q_ptr: int, scheme="rk4_jameson", cfl=1.0):
        """One explicit step on a device SoA state in place; returns the
        last stage's residual norms."""
        l2 = np.zeros(self.nvar)
        sc = abi.SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        self._check(load().sg_step_explicit_dev(self.h, C.c_void_p(q_ptr), sc, C.c_double(cfl), _p(l2)))
        return l2

    def solve_explicit(self, q, **opts):
        """Solver::solve with forward Euler or RK4-Jameson from q; returns
        (q, report dict, per-step norms [steps, nvar])."""
        q = _f64(q, self.shape).copy()
        o = abi.explicit_opts(**opts)
        hist = np.zeros((o.iteration_max + 2, self.nvar))
        rep = abi.SolveReport()
        self._check(load().sg_solve_explicit(self.h, _p(q), C.byref(o), C.byref(rep), _p(hist)))
        report = {f: getattr(rep, f) for f, _ in abi.SolveReport._fields_}
        report["l2_initial"] = list(rep.l2_initial)[: self.nvar]
        report["l2_final"] = list(rep.l2_final)[: self.nvar]
        return q, report, hist[: rep.iterations]

    def ls_preallocate(self, restart=30, precond=2, maxit=60, rtol=1e-3):
        self._check(load().sg_ls_preallocate(self.h, restart, precond, maxit, C.c_double(rtol)))

    def ls_set_pc(self, lhs_order=0, pc_dt_factor=1.0, pc_width=0):
        self._check(load().sg_ls_set_pc(self.h, int(lhs_order), C.c_double(pc_dt_factor), int(pc_width)))

    def ls_set_lhs(self, q, dt):
        self._check(load().sg_ls_set_lhs(self.h, _p(_f64(q, self.shape)), _p(_f64(dt, self.shape))))

    def ls_set_rhs(self, r):
        self._check(load().sg_ls_set_rhs(self.h, _p(_f64(r, self.shape))))

    def ls_solve(self):
        dq = np.empty(self.shape)
        it = C.c_int()
        rr = C.c_double()
        rc = load().sg_ls_solve(self.h, _p(dq), C.byref(it), C.byref(rr))
        if rc not in (0, -5):
            self._check(rc)
        return dq, it.value, rr.value

    def ls_precond_apply(self, v, transpose=False):
        z = np.empty(self.shape)
        self._check(load().sg_ls_precond_apply(self.h, _p(_f64(v, self.shape)), _p(z), int(bool(transpose))))
        return z

    # ---- adjoint
    def adjoint_solve(self, q, rhs, restart=40, maxit=400, rtol=1e-10, precond=4, pc_cfl=1e3, lhs_order=0,
                       pc_width=0):
        q = _f64(q, self.shape)
        rhs = _f64(rhs, self.shape)
        psi = np.empty(self.shape)
        o = abi.AdjointOpts(restart=restart, maxit=maxit, rtol=rtol, precond=precond, lhs_order=lhs_order,
                            pc_width=pc_width, pc_cfl=pc_cfl)
        rep = abi.AdjointReport()
        rc = load().sg_adjoint_solve(self.h, _p(q), _p(rhs), _p(psi), C.byref(o), C.byref(rep))
        if rc not in (0, -5):
            self._check(rc)
        return psi, {"iterations": rep.iterations, "relres": rep.relres, "wall_s": rep.wall_s,
                     "reorthogonalised": rep.reorthogonalised, "converged": rc == 0}

    def objective_cf(self, q, cf_target, grad=True):
        q = _f64(q, self.shape)
        t = _f64(cf_target)
        F = C.c_double()
        g = np.empty(self.shape) if grad else None
        self._check(load().sg_objective_cf(self.h, _p(q), _p(t), C.byref(F), _p(g)))
        return F.value, g

    def adjoint_gradient(self, q, cf_target, restart=40, maxit=400, rtol=1e-10, precond=4, pc_cfl=1e3, lhs_order=0,
                          pc_width=0):
        q = _f64(q, self.shape)
        t = _f64(cf_target)
        gb = np.empty((self.nic, self.njc))
        F = C.c_double()
        o = abi.AdjointOpts(restart=restart, maxit=maxit, rtol=rtol, precond=precond, lhs_order=lhs_order,
                            pc_width=pc_width, pc_cfl=pc_cfl)
        rep = abi.AdjointReport()
        rc = load().sg_adjoint_gradient(self.h, _p(q), _p(t), _p(gb), C.byref(F), C.byref(o), C.byref(rep))
        if rc not in (0, -5):
            self._check(rc)
        return gb, F.value, {"iterations": rep.iterations, "relres": rep.relres, "wall_s": rep.wall_s,
                             "reorthogonalised": rep.reorthogonalised, "converged": rc == 0}

    def invert(self, q, cf_target, beta=None, iterations=50, memory=10, reg=0.0, step0=0.2, beta_min=0.0,
               beta_max=5.0, max_backtracks=6, newton=None, adjoint=None):
        """Adjoint-driven L-BFGS field inversion of beta from the converged
        state q at the initial beta (default 1); returns (q, beta, report,
        J history)."""
        q = _f64(q, self.shape).copy()
        b = np.ones((self.nic, self.njc)) if beta is None else _f64(beta, (self.nic, self.njc)).copy()
        adj = dict(restart=60, maxit=2000, rtol=1e-10, precond=4, lhs_order=0, pc_cfl=300.0)
        adj.update(adjoint or {})
        o = abi.InversionOpts(iterations=iterations, memory=memory, reg=reg, step0=step0, beta_min=beta_min,
                              beta_max=beta_max, max_backtracks=max_backtracks, reserved=0,
                              newton=abi.newton_opts(**(newton or {})), adjoint=abi.AdjointOpts(**adj))
        rep = abi.InversionReport()
        hist = np.full(iterations + 1, np.nan)
        self._check(load().sg_invert(self.h, _p(q), _p(_f64(cf_target)), _p(b), C.byref(o), C.byref(rep),
                                     _p(hist)))
        report = {f: getattr(rep, f) for f, _ in abi.InversionReport._fields_}
        return q, b, report, hist[: rep.iterations + 1]

    # ---- beta closures (no counterpart in the reference)
    CLOSURE_FIELD, CLOSURE_ANALYTIC, CLOSURE_MLP = 0, 1, 2

    def set_closure(self, kind, coef=None, params=None):
        """kind 0: beta field (sg_set_beta); 1: beta = 1 + c6 tanh(c[:5].f + c5);
        2: beta = 1 + MLP(f) with params (SG_MLP_NPARAM float32)."""
        c = None if coef is None else _f64(coef, (7,))
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p is not None and p.size != Mlp.NPARAM:
            raise ValueError("MLP parameter vector must have %d entries" % Mlp.NPARAM)
        self._check(load().sg_set_closure(self.h, int(kind), _p(c), _p(p)))

    def set_closure_mlp(self, net):
        self._check(load().sg_set_closure_mlp(self.h, net.h))

    def closure_eval(self, q):
        """Local features [nic][njc][5] and the closure's beta [nic][njc] at q."""
        q = _f64(q, self.shape)
        f = np.empty((self.nic, self.njc, 5))
        b = np.empty((self.nic, self.njc))
        self._check(load().sg_closure_eval(self.h, _p(q), _p(f), _p(b)))
        return f, b


class Mlp:
    """Closure network 5 -> 64 -> 64 -> 1 (tanh), trained on the GPU
    (sg_mlp_* entry points).  Parameters: float32 vector
    W1[64][5] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3."""
    NPARAM = 4609
    SLICES = {"W1": (0, (64, 5)), "b1": (320, (64,)), "W2": (384, (64, 64)), "b2": (4480, (64,)),
              "W3": (4544, (64,)), "b3": (4608, ())}

    def __init__(self, max_rows=65536, seed=0):
        lib = load()
        lib.sg_mlp_last_error.restype = C.c_char_p
        lib.sg_mlp_last_error.argtypes = [C.c_void_p]
        lib.sg_mlp_destroy.argtypes = [C.c_void_p]
        lib.sg_mlp_destroy.restype = None
        h = C.c_void_p()
        rc = lib.sg_mlp_create(C.byref(h), int(max_rows), C.c_ulonglong(seed))
        if rc != 0:
            raise SgError(rc, "sg_mlp_create failed")
        self.h = h
        self.max_rows = max_rows

    def close(self):
        if getattr(self, "h", None):
            load().sg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise SgError(rc, load().sg_mlp_last_error(self.h).decode())

    @property
    def params(self):
        p = np.empty(self.NPARAM, dtype=np.float32)
        self._check(load().sg_mlp_get_params(self.h, _p(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.NPARAM:
            raise ValueError("expected %d parameters" % self.NPARAM)
        self._check(load().sg_mlp_set_params(self.h, _p(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(load().sg_mlp_forward(self.h, x.shape[0], _p(x), _p(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        loss = C.c_float()
        g = np.empty(self.NPARAM, dtype=np.float32)
        self._check(load().sg_mlp_loss_grad(self.h, x.shape[0], _p(x), _p(t), C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, t, epochs=1, batch=65536, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        hist = np.zeros(max(epochs, 1), dtype=np.float32)
        self._check(load().sg_mlp_train(self.h, x.shape[0], _p(x), _p(t), int(epochs), int(batch), C.c_float(lr),
                                        C.c_ulonglong(seed), _p(hist)))
        return hist[:epochs]
