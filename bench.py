"""Benchmark of the fused FP64 residual on one MI355X.

Workload: turbulent flat plate with Spalart-Allmaras, 545 x 385 vertices
(544 x 384 cells, 5 unknowns per cell), the residual configuration named by
the project's headline metric.  One step = one residual evaluation of the
whole grid (fused residual kernel with the per-block norm partial sums, then
the norm finalisation) on a seeded perturbed freestream state already
resident in HBM.  Steps rotate over independent state/residual buffer sets
totalling > 512 MB so that the 256 MB Infinity Cache cannot hold the working
set.

Prints one JSON line.  The roofline entry uses the algorithmic bytes of the
SA residual (q 5 + metrics 5 + wall distance 1 + beta 1 + R 5 = 17 doubles =
136 B per cell) over the residual kernel's average duration, timed with HIP
events on the library's own stream.  cpu_baseline times the CPU oracle (a
C++ restatement of the reference's residual with the reference's OpenMP
loops) on the box's host cores on the same grid and state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

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
    return ap.parse_args()


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
                      "OpenMP over %d host threads; 1-thread: %.4f Mcells/s" % (reps, case.ni, case.nj, threads,
                                                                                 cells / t_single / 1e6),
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
    l2 = torch.zeros(8, dtype=torch.float64, device=dev)
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
    # tangent (J.v) throughput on the same grid, one buffer set
    vd = torch.randn(nv * plane, dtype=torch.float64, device=dev)
    jv_out = torch.zeros_like(vd)
    for _ in range(5):
        eq.jv_dev(qs[0].data_ptr(), vd.data_ptr(), jv_out.data_ptr())
    j0, j1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nj = max(10, a.steps // 10)
    j0.record(stream)
    lib_ = lib.load()
    import ctypes as C
    for _ in range(nj):
        lib_.sg_jv_dev(eq.h, C.c_void_p(qs[0].data_ptr()), C.c_void_p(vd.data_ptr()), C.c_void_p(jv_out.data_ptr()))
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
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "residual_kernel<double,SA>",
                     "kernel_ms": ms_kernel, "bytes_per_cell": bytes_cell},
        "jv": {"ms": ms_jv, "mcells_s": cells / (ms_jv * 1e-3) / 1e6,
               "achieved_GBs_at_176B_cell": 176 * cells / (ms_jv * 1e-3) / 1e9},
        "wall_s": wall,
        "l2": [float(x) for x in l2h],
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, q0, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
