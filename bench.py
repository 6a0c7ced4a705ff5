"""Benchmark of the fused FP64 residual on one MI355X (plus the adjoint
gradient and the learned-closure solve).

Headline (value): turbulent flat plate with Spalart-Allmaras, 545 x 385
vertices (544 x 384 cells, 5 unknowns per cell).  One step = one residual
evaluation of the whole grid including its L2 norms (the fused residual kernel
whose last block finalises the norms) on a seeded perturbed freestream state
already resident in HBM.  Steps rotate over independent state/residual buffer
sets totalling > 512 MB so that the 256 MB Infinity Cache cannot hold the
working set.  Timed with HIP events on the library's own stream.

roofline: the residual kernel alone (launches without the norm), algorithmic
bytes of the SA residual (q 5 + normals 4 + volume 1 + wall distance 1 +
beta 1 + R 5 = 17 doubles = 136 B per cell) over its average duration;
traffic is the calibrated HBM bytes per launch from the committed counter
profile of the same grid.  cpu_baseline times the CPU oracle (a C++
restatement of the reference's residual with the reference's OpenMP loops) on
the box's host cores on the same grid and state.

adjoint_gradient: wall time of one adjoint beta-gradient of the wall-Cf
objective on the 705 x 321 channel bump (beta = 1 + 8 seeded blobs, target Cf
from a second seeded beta*), after converging the primal states.

ml_closure_vs_inversion: on the 1409 x 641 bump (h 0.065, w 0.25, Re 5e6),
the solve with the closure network (trained on the inverted beta fields of the
24 training bumps) fused into the SA source against the plain-SA solve and
against a full adjoint field inversion on the same grid: wall-Cf L2 error
of the closure solve against the inversion, and the inversion's time (its
per-iteration time measured here) over the closure solve's.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 residual Mcells/s; adjoint-gradient s; ML-closure vs plain-SA solve time"
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--ni", type=int, default=545)
    ap.add_argument("--nj", type=int, default=385)
    ap.add_argument("--min-bytes", type=float, default=600e6, help="rotating working set")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient leg")
    ap.add_argument("--no-closure", action="store_true", help="skip the ML-closure vs plain-SA leg")
    ap.add_argument("--adjoint-grid", default="705x321")
    ap.add_argument("--closure-grid", default="1409x641")
    ap.add_argument("--inversion-iterations", type=int, default=2,
                    help="L-BFGS iterations of the field inversion timed in the closure leg")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="MB: also time a plain device copy of this size (HBM counter calibration)")
    ap.add_argument("--no-extra-grids", action="store_true",
                    help="skip the 1409x641 roofline line and the 97x65 laminar evaluations/s")
    return ap.parse_args()


def _grid(s):
    ni, nj = s.lower().split("x")
    return int(ni), int(nj)


def newton_leg(ni, nj):
    """The SA flat plate's steady solve: Newton-GMRES(30) from freestream
    with the fused residual and matrix-free J.v, chunked line ILU(0) of the
    first-order Jacobian, to a flux-balance norm max_k ||V R_k||_2 < 1e-12.
    Reports the reference's norm max_k ||R_k||_2 (R = flux balance / V) at
    the end too: with 1e-6 wall spacing its FP64 roundoff floor lies far
    above 1e-10, the flux-balance norm's does not."""
    from structured_amd import cases, fiml, lib
    case = cases.sa_plate(ni, nj)
    eq = lib.Equation(case)
    t0 = time.perf_counter()
    q, r, hist = eq.solve(case.freestream_state(), **fiml.NEWTON)
    t = time.perf_counter() - t0
    eq.close()
    return {"grid": "%dx%d" % (ni, nj), "solve_s": t, "newton": r["iterations"], "converged": bool(r["converged"]),
            "linear_iterations": r["linear_iterations"], "l2w_final_max": max(r["l2w_final"]),
            "l2_final_max": max(r["l2_final"]), "l2_initial_max": max(r["l2_initial"]),
            "tolerance_l2w": fiml.NEWTON["tolerance"]}


def adjoint_cpu_baseline(case, q, gmres_iterations):
    """The adjoint gradient's CPU baseline on a bounded sample: the CPU
    oracle's Jacobian-block extraction (5 x 5 coloured dual sweeps of the
    restated reference residual) timed whole, plus 3 transposed Krylov
    iterations (block J^T psi, block-Jacobi apply, CGS2 against 60 basis
    vectors: the GPU run's restart 120 averages 60) timed and scaled to the
    GPU run's iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as of
    threads = os.cpu_count() or 1
    t_blk, t_it = of.adjoint_sample(case, q, threads=threads, iters=3, basis=60)
    return {"value": t_blk + gmres_iterations * t_it, "unit": "s", "cores": threads, "kind": "port",
            "sample": "Jacobian blocks %.2f s (measured) + %d Krylov iterations x %.4f s (3 measured) on %d threads"
                      % (t_blk, gmres_iterations, t_it, threads),
            "jacobian_blocks_s": t_blk, "krylov_iteration_s": t_it}


def adjoint_leg(grid, cpu=True):
    """Adjoint beta-gradient of the wall-Cf objective on the channel bump:
    beta = 1 + 8 seeded blobs, target Cf from a second seeded beta*.  The
    primal states are converged first (untimed for the metric); the metric is
    the wall time of sg_adjoint_gradient (objective, dF/dq, Jacobian blocks,
    preconditioner, transposed GMRES, dR/dbeta)."""
    from structured_amd import cases, fiml, lib
    ni, nj = _grid(grid)
    case = cases.sa_bump(ni, nj)
    eq = lib.Equation(case)
    t0 = time.perf_counter()
    q1, r1, _ = eq.solve(case.freestream_state(), **fiml.NEWTON)
    t_cold = time.perf_counter() - t0
    eq.set_beta(cases.blob_beta(case, seed=2))
    qs, r2, _ = eq.solve(q1, **fiml.WARM)
    _, _, cf_target, _ = eq.wall(qs)
    eq.set_beta(cases.blob_beta(case, seed=1))
    t0 = time.perf_counter()
    q, r3, _ = eq.solve(q1, **fiml.WARM)
    t_warm = time.perf_counter() - t0
    adj = dict(fiml.ADJOINT)
    t0 = time.perf_counter()
    g, F, rep = eq.adjoint_gradient(q, cf_target, **adj)
    t_adj = time.perf_counter() - t0
    out = {"grid": grid, "cells": case.nic * case.njc, "adjoint_gradient_s": t_adj,
           "gmres_iterations": rep["iterations"], "relres": rep["relres"], "rtol": adj["rtol"],
           "converged": rep["converged"], "F": F, "grad_max": float(np.abs(g).max()),
           "primal_cold_solve_s": t_cold, "primal_cold_newton": r1["iterations"],
           "primal_warm_solve_s": t_warm, "primal_converged": [bool(r["converged"]) for r in (r1, r2, r3)],
           "inversion_iteration_s_estimate": t_warm + t_adj}
    out["preconditioner_apply"] = ilu_apply_leg(eq, q, adj)
    eq.close()
    if cpu:
        out["cpu_baseline"] = adjoint_cpu_baseline(case, q, rep["iterations"])
    return out


def ilu_apply_leg(eq, q, adj, applies=50):
    """Per-apply time of the adjoint's preconditioner (chunked line ILU(0)
    of the first-order Jacobian at the adjoint's pc CFL, chunk width 32):
    M^-1 and M^-T on device vectors, back to back on the library's stream,
    timed with HIP events on that stream, and the factorisation (blocks by
    coloured J.v sweeps, gather, factor) timed by the host around set_lhs."""
    import ctypes as C
    import torch
    from structured_amd import lib
    so = lib.load()
    dev = torch.device("cuda:0")
    dt = eq.calc_dt(q, adj["pc_cfl"])
    eq.ls_preallocate(restart=10, precond=adj["precond"], maxit=10, rtol=1e-6)
    eq.ls_set_pc(adj["lhs_order"], 1.0, 32)
    eq.ls_set_lhs(q, dt)
    t0 = time.perf_counter()
    eq.ls_set_lhs(q, dt)
    t_setup = time.perf_counter() - t0
    n = eq.nvar * eq.plane
    vs = [torch.randn(n, dtype=torch.float64, device=dev) for _ in range(2)]
    zs = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(2)]
    stream = torch.cuda.ExternalStream(eq.stream(), device=dev)
    out = {"precond": adj["precond"], "lhs_order": adj["lhs_order"], "pc_cfl": adj["pc_cfl"], "width": 32,
           "set_lhs_ms": 1e3 * t_setup, "applies": applies}
    for name, tr in (("apply_inverse_us", 0), ("apply_inverse_transpose_us", 1)):
        def run(k_):
            for k in range(k_):
                rc = so.sg_ls_precond_apply_dev_async(eq.h, C.c_void_p(vs[k % 2].data_ptr()),
                                                      C.c_void_p(zs[k % 2].data_ptr()), tr)
                assert rc == 0, so.sg_last_error(eq.h)
        run(5)
        eq.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run(applies)
        e1.record(stream)
        eq.synchronize()
        out[name] = 1e3 * e0.elapsed_time(e1) / applies
    del vs, zs
    torch.cuda.empty_cache()
    return out


DATA = os.path.join(ROOT, "structured_amd", "data")


def closure_leg(grid, inversion_iterations):
    """The learned closure in the solve loop against plain SA and against a
    full adjoint field inversion, all on the unseen bump (h 0.065, w 0.25,
    Re 5e6) on the same grid.

    - The network (5-64-64-1 tanh) is the one trained on the GPU by this
      library's kernels on the inverted beta fields of the 24 training bumps
      (385x193, 50 L-BFGS iterations each): structured_amd/data/closure_net.npz.
    - The inversion reference is structured_amd/data/reference_<grid>.npz: the
      wall Cf of a 50-iteration sg_invert run on this grid against the
      truth's Cf (beta* = 1 + 0.3 tanh(c . f + c0)), and its wall time.
    - Timed here: the plain-SA solve and the closure solve (both from
      freestream) and `inversion_iterations` L-BFGS iterations of the same
      inversion from the plain-SA state, whose per-iteration time gives the
      inversion's cost on this GPU."""
    from structured_amd import cases, fiml, lib
    ni, nj = _grid(grid)
    net_path = os.path.join(DATA, "closure_net.npz")
    ref_path = os.path.join(DATA, "reference_%s.npz" % grid)
    case = cases.sa_bump(ni, nj, height=0.065, width=0.25, re=5e6)
    eq = lib.Equation(case)
    t0 = time.perf_counter()
    q_sa, r_sa, _ = eq.solve(case.freestream_state(), **fiml.NEWTON)
    t_sa = time.perf_counter() - t0
    _, _, cf_sa, _ = eq.wall(q_sa)
    out = {"grid": grid, "cells": case.nic * case.njc, "plain_sa_solve_s": t_sa,
           "plain_sa_newton": r_sa["iterations"], "plain_sa_converged": bool(r_sa["converged"])}
    if not (os.path.exists(net_path) and os.path.exists(ref_path)):
        out["missing"] = [p for p in (net_path, ref_path) if not os.path.exists(p)]
        eq.close()
        return out
    params = np.load(net_path)["params"].astype(np.float32)
    ref = np.load(ref_path)
    cf_inv, cf_true = ref["cf_inv"], ref["cf_true"]
    eq.set_closure(2, params=params)
    o = dict(fiml.NEWTON, tolerance=fiml.MLP_TOL, stall_window=fiml.MLP_STALL)
    t0 = time.perf_counter()
    q_ml, r_ml, _ = eq.solve(case.freestream_state(), **o)
    t_ml = time.perf_counter() - t0
    _, _, cf_ml, _ = eq.wall(q_ml)
    eq.set_closure(0)
    # the inversion's cost on this GPU: L-BFGS iterations from the plain-SA state
    t0 = time.perf_counter()
    _, _, irep, _ = eq.invert(q_sa, cf_true, iterations=inversion_iterations, newton=fiml.RESOLVE,
                              adjoint=fiml.ADJOINT)
    t_it = (time.perf_counter() - t0) / max(irep["iterations"], 1)
    eq.close()
    n_ref = int(ref["iterations_done"])
    out.update({
        "ml_closure_solve_s": t_ml, "ml_closure_newton": r_ml["iterations"],
        "ml_closure_converged": bool(fiml.closure_solve_ok(r_ml)), "ml_over_sa_time": t_ml / t_sa,
        "cf_l2_err": fiml.cf_error(cf_ml, cf_inv),
        "cf_l2_err_sa_vs_inversion": fiml.cf_error(cf_sa, cf_inv),
        "cf_err_vs_truth": {"ml": fiml.cf_error(cf_ml, cf_true), "inversion": fiml.cf_error(cf_inv, cf_true),
                            "plain_sa": fiml.cf_error(cf_sa, cf_true)},
        "inversion_iterations_timed": int(irep["iterations"]), "inversion_iteration_s": t_it,
        "inversion_reference_iterations": n_ref, "inversion_reference_s": float(ref["inversion_s"]),
        "time_ratio": t_it * n_ref / t_ml,
        "time_ratio_reference_run": float(ref["inversion_s"]) / t_ml,
    })
    return out


def time_residual(eq, q0, steps, min_bytes, with_norm, rotate=True):
    """Average per-launch time (ms) of back-to-back residual launches on the
    library's stream, timed with HIP events on that stream.  rotate: cycle
    over independent state/residual sets of at least min_bytes in all (cold,
    beyond the 256 MB Infinity Cache); else one set (hot)."""
    import torch
    dev = torch.device("cuda:0")
    nv, plane = eq.nvar, eq.plane
    set_bytes = 2 * nv * plane * 8
    nsets = max(2, int(np.ceil(min_bytes / set_bytes))) if rotate else 1
    aos = torch.from_numpy(q0.reshape(-1)).to(dev)
    qs, rs = [], []
    for _ in range(nsets):
        qd = torch.zeros(nv * plane, dtype=torch.float64, device=dev)
        eq.aos_to_soa_dev(aos.data_ptr(), qd.data_ptr())
        qs.append(qd)
        rs.append(torch.zeros(nv * plane, dtype=torch.float64, device=dev))
    l2 = torch.zeros(nv, dtype=torch.float64, device=dev)
    eq.synchronize()
    stream = torch.cuda.ExternalStream(eq.stream(), device=dev)

    def run(n):
        for k in range(n):
            s = k % nsets
            eq.residual_dev_async(qs[s].data_ptr(), rs[s].data_ptr(), l2.data_ptr() if with_norm else None)

    run(min(steps, 20))
    eq.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    run(steps)
    e1.record(stream)
    eq.synchronize()
    ms = e0.elapsed_time(e1) / steps
    l2h = l2.cpu().numpy()
    del qs, rs
    torch.cuda.empty_cache()
    return ms, nsets * set_bytes, l2h


def extra_grids(steps):
    """The 1409 x 641 bump (h 0.065, w 0.25, Re 5e6; the largest grid of the
    workloads) residual roofline, cold and hot, and the 97 x 65 laminar
    plate's residual evaluations per second (whole evaluations: kernel and
    norms)."""
    from structured_amd import cases, lib
    out = {}
    case = cases.sa_bump(1409, 641, height=0.065, width=0.25, re=5e6)
    eq = lib.Equation(case)
    q0 = case.perturbed_state(seed=3)
    cells = case.nic * case.njc
    n = max(20, steps // 4)
    ms_cold, nbytes, _ = time_residual(eq, q0, n, 1.2e9, with_norm=False)
    ms_hot, _, _ = time_residual(eq, q0, n, 0, with_norm=False, rotate=False)
    ms_step, _, _ = time_residual(eq, q0, n, 1.2e9, with_norm=True)
    eq.close()
    ach = 136 * cells / (ms_cold * 1e-3)
    out["bump_1409x641"] = {"cells": cells, "kernel_ms_cold": ms_cold, "kernel_ms_hot": ms_hot,
                            "step_ms_cold_with_norm": ms_step, "rotating_bytes": nbytes,
                            "mcells_s_step": cells / (ms_step * 1e-3) / 1e6,
                            "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9,
                                         "unit": "GB/s", "frac": ach / HBM_PEAK, "bytes_per_cell": 136}}
    case = cases.laminar_plate(97, 65)
    eq = lib.Equation(case)
    q0 = case.perturbed_state(seed=3)
    cells = case.nic * case.njc
    ms_step, nbytes, _ = time_residual(eq, q0, steps, 600e6, with_norm=True)
    ms_hot, _, _ = time_residual(eq, q0, steps, 0, with_norm=True, rotate=False)
    eq.close()
    out["laminar_97x65"] = {"cells": cells, "residual_evals_per_s_cold": 1e3 / ms_step,
                            "residual_evals_per_s_hot": 1e3 / ms_hot, "step_ms_cold": ms_step, "step_ms_hot": ms_hot,
                            "mcells_s_cold": cells / (ms_step * 1e-3) / 1e6, "bytes_per_cell": 104}
    return out


def measured_traffic(ni, nj):
    """HBM bytes per launch of the residual kernel from the committed
    counter profile (FETCH_SIZE / WRITE_SIZE passes scaled by a copy of known
    size), when it was taken on this grid; else None."""
    path = os.path.join(ROOT, "profiles", "residual_sa_%dx%d_traffic.md" % (ni, nj))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].split()[0])
    return None


def cpu_baseline(case, q, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as of
    threads = os.cpu_count() or 1
    t1 = of.time_residual(case, q, warm=1, reps=1, threads=threads)
    reps = max(3, min(50, int(budget_s / max(t1, 1e-6))))
    t = of.time_residual(case, q, warm=1, reps=reps, threads=threads)
    cells = case.nic * case.njc
    t_single = of.time_residual(case, q, warm=0, reps=1, threads=1)
    return {"value": cells / t / 1e6, "unit": "Mcells/s", "cores": threads, "kind": "port",
            "sample": "median of %d residual evaluations of the same %dx%d SA grid and state (after 1 warm-up), "
                      "OpenMP over %d host threads; 1-thread: %.4f Mcells/s (the primitive, reconstruction, "
                      "convective-flux and assembly loops are threaded; the boundary conditions, face gradients "
                      "and viscous fluxes run serially as in the reference, so threads add little)"
                      % (reps, case.ni, case.nj, threads, cells / t_single / 1e6),
            "single_thread_mcells_s": cells / t_single / 1e6}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from structured_amd import cases, lib

    torch.cuda.set_device(0)
    case = cases.sa_plate(a.ni, a.nj)
    cells = case.nic * case.njc
    eq = lib.Equation(case)
    nv, plane = eq.nvar, eq.plane
    q0 = case.perturbed_state(seed=3)
    set_bytes = 2 * nv * plane * 8
    nsets = max(2, int(np.ceil(a.min_bytes / set_bytes)))
    dev = torch.device("cuda:0")
    aos = torch.from_numpy(q0.reshape(-1)).to(dev)
    qs, rs = [], []
    for s in range(nsets):
        qd = torch.zeros(nv * plane, dtype=torch.float64, device=dev)
        eq.aos_to_soa_dev(aos.data_ptr(), qd.data_ptr())
        qs.append(qd)
        rs.append(torch.zeros(nv * plane, dtype=torch.float64, device=dev))
    l2 = torch.zeros(nv, dtype=torch.float64, device=dev)  # sg_residual_dev_async writes nvar norms
    eq.synchronize()
    stream = torch.cuda.ExternalStream(eq.stream(), device=dev)

    def run(n, with_norm=True):
        for k in range(n):
            s = k % nsets
            eq.residual_dev_async(qs[s].data_ptr(), rs[s].data_ptr(), l2.data_ptr() if with_norm else None)

    run(a.warmup)
    eq.synchronize()
    torch.cuda.synchronize()
    # timed region: whole steps (residual + norm)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    run(a.steps)
    e1.record(stream)
    eq.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / a.steps
    # residual kernel alone, back to back (the roofline's kernel duration)
    k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    k0.record(stream)
    run(a.steps, with_norm=False)
    k1.record(stream)
    eq.synchronize()
    ms_kernel = k0.elapsed_time(k1) / a.steps
    l2h = l2.cpu().numpy()[:nv]
    if not np.all(np.isfinite(l2h)):
        raise SystemExit("non-finite residual norm")
    bytes_cell = 17 * 8
    achieved = bytes_cell * cells / (ms_kernel * 1e-3)
    # hot: the same launches on one buffer set (working set in the caches)
    h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    h0.record(stream)
    for _ in range(a.steps):
        eq.residual_dev_async(qs[0].data_ptr(), rs[0].data_ptr(), l2.data_ptr())
    h1.record(stream)
    eq.synchronize()
    ms_step_hot = h0.elapsed_time(h1) / a.steps
    # tangent (J.v) kernel, back to back without synchronising, rotating over
    # the state sets (the direction and output are one extra pair)
    vd = torch.randn(nv * plane, dtype=torch.float64, device=dev)
    jv_out = torch.zeros_like(vd)
    for _ in range(5):
        eq.jv_dev_async(qs[0].data_ptr(), vd.data_ptr(), jv_out.data_ptr())
    eq.synchronize()
    j0, j1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nj = max(10, a.steps // 10)
    j0.record(stream)
    for k in range(nj):
        eq.jv_dev_async(qs[k % nsets].data_ptr(), vd.data_ptr(), rs[k % nsets].data_ptr())
    j1.record(stream)
    eq.synchronize()
    ms_jv = j0.elapsed_time(j1) / nj

    out = {
        "metric": METRIC,
        "value": cells * a.steps / (ms_step * a.steps * 1e-3) / 1e6,
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "SA flat-plate FP64 residual (fused kernel + L2 norm)", "grid": "%dx%d" % (a.ni, a.nj),
                   "cells": cells, "unknowns_per_cell": nv, "rotating_sets": nsets,
                   "rotating_bytes": nsets * set_bytes},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": measured_traffic(a.ni, a.nj),
                     "kernel": "residual_kernel<double,SA>",
                     "kernel_ms": ms_kernel, "bytes_per_cell": bytes_cell},
        "hot": {"ms_per_step": ms_step_hot, "mcells_s": cells / (ms_step_hot * 1e-3) / 1e6},
        "jv": {"ms": ms_jv, "mcells_s": cells / (ms_jv * 1e-3) / 1e6,
               "achieved_GBs_at_176B_cell": 176 * cells / (ms_jv * 1e-3) / 1e9,
               "frac": 176 * cells / (ms_jv * 1e-3) / HBM_PEAK},
        "wall_s": wall,
        "l2": [float(x) for x in l2h],
    }
    if a.copy_calibration > 0:
        n = int(a.copy_calibration * 1e6 / 8)
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(10):
            dst.copy_(src)
        c1.record()
        torch.cuda.synchronize()
        ms = c0.elapsed_time(c1) / 10
        out["copy_calibration"] = {"bytes_read": n * 8, "bytes_written": n * 8, "ms": ms,
                                   "GBs": 2 * n * 8 / (ms * 1e-3) / 1e9}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, q0, a.cpu_seconds)
    eq.close()
    del qs, rs
    torch.cuda.empty_cache()
    if not a.no_extra_grids:
        out["grids"] = extra_grids(a.steps)
        out["newton_sa_plate"] = newton_leg(a.ni, a.nj)
    if not a.no_adjoint:
        out["adjoint_gradient"] = adjoint_leg(a.adjoint_grid, cpu=not a.no_cpu)
    if not a.no_closure:
        out["ml_closure_vs_inversion"] = closure_leg(a.closure_grid, a.inversion_iterations)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
