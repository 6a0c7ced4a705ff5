"""Field inversion and machine learning of the SA production multiplier.

1. Synthetic truth: beta* = 1 + amp tanh(c . f + c0) on the local features
   f (the analytic closure, FP64); its converged wall Cf is the target.
2. Field inversion: adjoint-driven L-BFGS from beta = 1 (sg_invert).
3. Training data: (features, beta_inv - 1) at the inverted solution.
4. The closure network, trained on the GPU (sg_mlp_train).
5. Prediction: a forward solve with the network fused into the SA source.

Everything runs through the C ABI (structured_amd.lib); this module only
sequences the calls and keeps the timings.
"""
from __future__ import annotations

import time

import numpy as np

from . import cases, lib

# Newton-Krylov settings of the SA solves: pseudo-transient continuation with
# switched evolution relaxation on the flux-balance norm; GMRES(30)
# preconditioned by the chunked line ILU(0) of the first-order Jacobian
# (lhs_order 1) on the 3x3-box pattern (precond 7: that Jacobian has no +-2
# blocks, so only fill is dropped), which stays effective up to CFL 1e10 on
# 1e-6 wall spacing.
NEWTON = dict(cfl=5.0, ser=1, ser_exponent=1.5, cfl_max=1e10, gmres_rtol=1e-4, precond=7, lhs_order=1,
              pc_width=32, pc_cfl_max=0.0, gmres_restart=30, gmres_maxit=150, iteration_max=200, norm_weighted=1,
              tolerance=1e-12)
# warm starts (a converged state of a nearby problem): the initial residual is
# small, so SER would keep the CFL near its start; start high instead
WARM = dict(NEWTON, cfl=1e5)
# re-solves inside the inversion: a change of beta late in an inversion can
# leave ||V R|| below 1e-12, so that a re-solve to that tolerance would not
# take a single Newton step and the objective would not see the change (the
# 1409x641 inversion stopped so after 7 iterations); these run to the
# roundoff floor instead (stall_window: no halving over 3 steps, within 1000x
# of the tolerance)
RESOLVE = dict(WARM, iteration_max=150, tolerance=1e-15, stall_window=3)
# the BF16 closure network makes beta a piecewise-constant function of the
# state at its rounding points, which sets a floor on the flux-balance norm
# (measured ~1.5e-11 on the 65x49 plate against 2e-12 for plain SA)
MLP_TOL = 1e-10
MLP_STALL = 8
# the adjoint's transposed GMRES: restart 60 stalled near relres 1e-1 on two
# of the 24 training bumps (the narrowest, w 0.12 and 0.15); 150 converged
# there in 1,335 and 1,944 iterations, and 120 is used throughout.  On the
# 1409x641 test bump the first gradient takes 6,472 iterations to 1e-8, so
# the cap is 10,000 (the 385x193 training inversions ran with 4,000, which 4
# of their ~800 adjoint solves reached)
ADJOINT = dict(restart=120, maxit=10000, rtol=1e-8, precond=7, lhs_order=1, pc_cfl=1e4)


def truth_coef(seed: int = 0, amp: float = 0.3) -> np.ndarray:
    """Seeded coefficients of beta* = 1 + amp tanh(c . f + c0)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2.0, 2.0, 5)
    c0 = rng.uniform(-0.5, 0.5)
    return np.r_[c, c0, amp]


def bump_case(ni, nj, height, width, re=3e6, dy1=1e-6):
    return cases.sa_bump(ni, nj, height=height, width=width, dy1=dy1, re=re)


def timed_solve(eq, q0, **opts):
    t = time.perf_counter()
    q, rep, hist = eq.solve(q0, **opts)
    return q, rep, time.perf_counter() - t


def closure_solve_ok(rep, floor=1e-9):
    """Converged, or stalled at the closure's floor below `floor`."""
    return bool(rep["converged"]) or (bool(rep["stalled"]) and max(rep["l2w_final"]) < floor)


def cf_error(cf, ref):
    """Relative L2 error of a wall Cf distribution."""
    return float(np.linalg.norm(cf - ref) / max(np.linalg.norm(ref), 1e-300))


def run_case(case, coef, invert_iterations=50, newton=None, verbose=False, resolve=None):
    """Truth, baseline and inversion on one case; returns a dict with the
    states, beta fields, wall Cf and timings."""
    newton = dict(NEWTON, **(newton or {}))
    eq = lib.Equation(case)
    out = {"case": case.name}
    q_sa, rep, t = timed_solve(eq, case.freestream_state(), **newton)
    out.update(q_sa=q_sa, sa_report=rep, sa_s=t)
    if verbose:
        print("%s: plain SA %.2fs (%d Newton steps, converged %d)" % (case.name, t, rep["iterations"],
                                                                     rep["converged"]), flush=True)
    eq.set_closure(1, coef=coef)
    q_true, rep, t = timed_solve(eq, q_sa, **dict(newton, cfl=WARM["cfl"]))
    out.update(q_true=q_true, true_report=rep, true_s=t)
    feats_true, beta_true = eq.closure_eval(q_true)
    _, _, cf_true, _ = eq.wall(q_true)
    eq.set_closure(0)
    _, _, cf_sa, _ = eq.wall(q_sa)
    out.update(beta_true=beta_true, features_true=feats_true, cf_true=cf_true, cf_sa=cf_sa)
    if invert_iterations <= 0:
        eq.close()
        return out
    t0 = time.perf_counter()
    q_inv, beta_inv, irep, jh = eq.invert(q_sa, cf_true, iterations=invert_iterations,
                                          newton=dict(resolve or RESOLVE), adjoint=ADJOINT)
    out.update(q_inv=q_inv, beta_inv=beta_inv, inv_report=irep, inv_hist=jh, inv_s=time.perf_counter() - t0)
    _, _, cf_inv, _ = eq.wall(q_inv)
    feats, _ = eq.closure_eval(q_inv)
    out.update(cf_inv=cf_inv, features=feats)
    if verbose:
        irep = out["inv_report"]
        print("%s: SA %.2fs (%d it), truth %.2fs, inversion %.2fs (%d it, J %.3e -> %.3e); "
              "Cf err SA %.3e inv %.3e" % (case.name, out["sa_s"], out["sa_report"]["iterations"], out["true_s"],
                                           out["inv_s"], irep["iterations"], irep["J0"], irep["J"],
                                           cf_error(cf_sa, cf_true), cf_error(cf_inv, cf_true)), flush=True)
    eq.close()
    return out


def training_set(results, wall_band=None):
    """Stack (features, beta_inv - 1) over the inverted cases.  wall_band:
    keep only cells with f1 (log wall-distance Reynolds number) above it."""
    X, T = [], []
    for r in results:
        f = r["features"].reshape(-1, 5)
        t = (r["beta_inv"] - 1.0).reshape(-1)
        if wall_band is not None:
            m = f[:, 1] > wall_band
            f, t = f[m], t[m]
        X.append(f)
        T.append(t)
    return np.concatenate(X).astype(np.float32), np.concatenate(T).astype(np.float32)


def train(X, T, epochs=200, batch=65536, lr=1e-3, seed=0):
    net = lib.Mlp(max_rows=min(batch, len(X)), seed=seed)
    t = time.perf_counter()
    hist = net.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return net, hist, time.perf_counter() - t


def predict(case, net, q0=None, newton=None):
    """Forward solve with the network fused into the SA source; returns
    (q, report, seconds, wall Cf)."""
    o = dict(NEWTON, tolerance=MLP_TOL, stall_window=MLP_STALL)
    if q0 is not None:
        o["cfl"] = WARM["cfl"]
    o.update(newton or {})
    eq = lib.Equation(case)
    eq.set_closure_mlp(net)
    q, rep, t = timed_solve(eq, case.freestream_state() if q0 is None else q0, **o)
    _, _, cf, _ = eq.wall(q)
    eq.close()
    return q, rep, t, cf


# ---------------------------------------------------------------- experiment
def _bumps(n, seed):
    rng = np.random.default_rng(seed)
    return [(float(h), float(w)) for h, w in zip(rng.uniform(0.02, 0.08, n), rng.uniform(0.1, 0.3, n))]


def stage_invert(args):
    """Invert a slice of the seeded training bumps; saves features and
    inverted beta (lower half of the grid, float32) per case."""
    import json
    import os
    os.makedirs(args.out, exist_ok=True)
    coef = truth_coef(seed=args.truth_seed)
    ni, nj = (int(v) for v in args.train_grid.split("x"))
    bumps = _bumps(args.ncases, args.case_seed)
    lo, hi = (int(v) for v in args.cases.split(":"))
    for k in range(lo, min(hi, len(bumps))):
        h, w = bumps[k]
        case = bump_case(ni, nj, h, w, re=3e6)
        r = run_case(case, coef, invert_iterations=args.iterations, verbose=True, resolve=dict(RESOLVE, verbose=1))
        half = case.njc // 2
        np.savez_compressed(os.path.join(args.out, "case_%02d.npz" % k),
                            features=r["features"][:, :half].astype(np.float32),
                            beta_inv=r["beta_inv"][:, :half].astype(np.float32),
                            beta_true=r["beta_true"][:, :half].astype(np.float32),
                            features_true=r["features_true"][:, :half].astype(np.float32))
        rep = r["inv_report"]
        summary = dict(case=k, height=h, width=w, sa_s=r["sa_s"], true_s=r["true_s"], inv_s=r["inv_s"],
                       inv_iterations=rep["iterations"], J0=rep["J0"], J=rep["J"],
                       evaluations=rep["evaluations"], failed_solves=rep["failed_solves"],
                       adjoint_unconverged=rep["adjoint_unconverged"], adjoint_relres_max=rep["adjoint_relres_max"],
                       adjoint_iterations=rep["adjoint_iterations"], steepest_restarts=rep["steepest_restarts"],
                       cf_err_sa=cf_error(r["cf_sa"], r["cf_true"]), cf_err_inv=cf_error(r["cf_inv"], r["cf_true"]),
                       sa_converged=bool(r["sa_report"]["converged"]),
                       true_converged=bool(r["true_report"]["converged"]))
        with open(os.path.join(args.out, "case_%02d.json" % k), "w") as f:
            json.dump(summary, f)
        print("case", json.dumps(summary), flush=True)


def stage_test(args):
    """Train on every saved case, then on the unseen bump compare plain SA,
    the learned closure and the full inversion against the truth."""
    import glob
    import json
    import os
    coef = truth_coef(seed=args.truth_seed)
    files = sorted(glob.glob(os.path.join(args.data, "case_*.npz")))
    X = np.concatenate([np.load(f)["features"].reshape(-1, 5) for f in files]).astype(np.float32)
    T = np.concatenate([np.load(f)["beta_inv"].reshape(-1) - 1.0 for f in files]).astype(np.float32)
    print("training set: %d cases, %d samples" % (len(files), len(X)), flush=True)
    net, hist, t_train = train(X, T, epochs=args.epochs, batch=65536, lr=1e-3, seed=args.net_seed)
    print("trained %d epochs in %.2fs: loss %.3e -> %.3e" % (args.epochs, t_train, hist[0], hist[-1]), flush=True)
    ni, nj = (int(v) for v in args.test_grid.split("x"))
    case = bump_case(ni, nj, 0.065, 0.25, re=5e6)
    # the inversion's primal re-solves report every Newton step (progress)
    res = run_case(case, coef, invert_iterations=args.iterations, verbose=True, resolve=dict(RESOLVE, verbose=1))
    q_ml, rep_ml, t_ml, cf_ml = predict(case, net)
    out = dict(train_cases=len(files), train_samples=int(len(X)), train_s=t_train, train_loss=[float(hist[0]),
                                                                                              float(hist[-1])],
               test_grid=args.test_grid, sa_s=res["sa_s"], truth_s=res["true_s"], inversion_s=res["inv_s"],
               inversion_iterations=res["inv_report"]["iterations"], ml_solve_s=t_ml,
               ml_solve_ok=closure_solve_ok(rep_ml), ml_newton=rep_ml["iterations"],
               cf_err_sa_vs_truth=cf_error(res["cf_sa"], res["cf_true"]),
               cf_err_inv_vs_truth=cf_error(res["cf_inv"], res["cf_true"]),
               cf_err_ml_vs_truth=cf_error(cf_ml, res["cf_true"]),
               cf_err_ml_vs_inversion=cf_error(cf_ml, res["cf_inv"]),
               inversion_over_ml_time=res["inv_s"] / t_ml)
    print("result", json.dumps(out), flush=True)
    if args.out:
        os.makedirs(args.out, exist_ok=True)
        with open(os.path.join(args.out, "test.json"), "w") as f:
            json.dump(out, f)


def stage_train(args):
    """Train the closure network on every saved case (features, beta_inv -
    1, lower half of each grid) and save its parameters with the loss
    history."""
    import glob
    import json
    import os
    files = sorted(glob.glob(os.path.join(args.data, "case_*.npz")))
    X = np.concatenate([np.load(f)["features"].reshape(-1, 5) for f in files]).astype(np.float32)
    T = np.concatenate([np.load(f)["beta_inv"].reshape(-1) - 1.0 for f in files]).astype(np.float32)
    print("training set: %d cases, %d samples" % (len(files), len(X)), flush=True)
    net, hist, t_train = train(X, T, epochs=args.epochs, batch=65536, lr=1e-3, seed=args.net_seed)
    print("trained %d epochs in %.2fs: loss %.3e -> %.3e" % (args.epochs, t_train, hist[0], hist[-1]), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.net_out)), exist_ok=True)
    np.savez(args.net_out, params=net.params, loss=np.asarray(hist, dtype=np.float32))
    meta = dict(cases=len(files), samples=int(len(X)), epochs=args.epochs, batch=65536, lr=1e-3, seed=args.net_seed,
                train_s=t_train, loss_first=float(hist[0]), loss_last=float(hist[-1]))
    with open(os.path.splitext(args.net_out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print("net", json.dumps(meta), flush=True)


def stage_reference(args):
    """The unseen bump (h 0.065, w 0.25, Re 5e6) on the test grid: plain SA,
    the truth (analytic beta*) and `iterations` L-BFGS iterations of the
    field inversion, resumed from a saved beta field when --resume is given
    (the L-BFGS memory starts empty again).  Saves the wall Cf of all three,
    the inverted beta and the timings."""
    import json
    import os
    coef = truth_coef(seed=args.truth_seed)
    ni, nj = (int(v) for v in args.test_grid.split("x"))
    case = bump_case(ni, nj, 0.065, 0.25, re=5e6)
    eq = lib.Equation(case)
    q_sa, r_sa, t_sa = timed_solve(eq, case.freestream_state(), **NEWTON)
    print("plain SA %.1fs (%d steps)" % (t_sa, r_sa["iterations"]), flush=True)
    eq.set_closure(1, coef=coef)
    q_true, r_true, t_true = timed_solve(eq, q_sa, **WARM)
    _, _, cf_true, _ = eq.wall(q_true)
    eq.set_closure(0)
    _, _, cf_sa, _ = eq.wall(q_sa)
    done = 0
    wall_before = 0.0
    beta0 = None
    q0 = q_sa
    if args.resume:
        z = np.load(args.resume)
        beta0, done, wall_before = z["beta_inv"], int(z["iterations_done"]), float(z["inversion_s"])
        eq.set_beta(beta0)
        q0, _, _ = timed_solve(eq, q_sa, **WARM)
        print("resumed from %s after %d iterations" % (args.resume, done), flush=True)
    # the inversion runs in chunks of --chunk iterations, saving after each,
    # so that a run cut short keeps what it did (the L-BFGS memory restarts
    # at every chunk, as it does on --resume)
    os.makedirs(args.out, exist_ok=True)
    chunk = args.chunk if args.chunk > 0 else args.iterations
    q_cur, beta_cur = q0, beta0
    t_inv, its, evals, failed, adj_unconv = 0.0, 0, 0, 0, 0
    J0, J, jhs = None, None, []
    while its < args.iterations:
        n = min(chunk, args.iterations - its)
        t0 = time.perf_counter()
        q_cur, beta_cur, irep, jh = eq.invert(q_cur, cf_true, beta=beta_cur, iterations=n,
                                              newton=dict(RESOLVE, verbose=1), adjoint=ADJOINT)
        t_inv += time.perf_counter() - t0
        J0 = irep["J0"] if J0 is None else J0
        J = irep["J"]
        its += irep["iterations"]
        evals += irep["evaluations"]
        failed += irep["failed_solves"]
        adj_unconv += irep["adjoint_unconverged"]
        jhs.append(np.asarray(jh)[:irep["iterations"] + 1] if not jhs else np.asarray(jh)[1:irep["iterations"] + 1])
        _, _, cf_inv, _ = eq.wall(q_cur)
        total = done + its
        np.savez_compressed(os.path.join(args.out, "reference_%s.npz" % args.test_grid), beta_inv=beta_cur,
                            cf_inv=cf_inv, cf_true=cf_true, cf_sa=cf_sa, J_hist=np.concatenate(jhs),
                            iterations_done=total, inversion_s=wall_before + t_inv)
        out = dict(test_grid=args.test_grid, sa_s=t_sa, sa_newton=r_sa["iterations"], truth_s=t_true,
                   iterations_this_run=its, iterations_done=total, inversion_s_this_run=t_inv,
                   inversion_s_total=wall_before + t_inv, J0=J0, J=J, evaluations=evals, failed_solves=failed,
                   adjoint_unconverged=adj_unconv, chunk=chunk, cf_err_sa_vs_truth=cf_error(cf_sa, cf_true),
                   cf_err_inv_vs_truth=cf_error(cf_inv, cf_true))
        with open(os.path.join(args.out, "reference_%s.json" % args.test_grid), "w") as f:
            json.dump(out, f, indent=1)
        print("reference", json.dumps(out), flush=True)
        if irep["iterations"] < n:  # the line search found no further descent
            break
    eq.close()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="field inversion and machine learning of the SA beta closure")
    ap.add_argument("stage", choices=["invert", "test", "train", "reference"])
    ap.add_argument("--ncases", type=int, default=24)
    ap.add_argument("--cases", default="0:24", help="slice of the training bumps to invert")
    ap.add_argument("--train-grid", default="385x193")
    ap.add_argument("--test-grid", default="1409x641")
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--truth-seed", type=int, default=1)
    ap.add_argument("--case-seed", type=int, default=7)
    ap.add_argument("--net-seed", type=int, default=0)
    ap.add_argument("--data", default="fiml_data")
    ap.add_argument("--out", default="fiml_data")
    ap.add_argument("--net-out", default="fiml_data/closure_net.npz")
    ap.add_argument("--resume", default="", help="reference stage: saved reference_*.npz to continue from")
    ap.add_argument("--chunk", type=int, default=10, help="reference stage: iterations per saved chunk (0: all)")
    args = ap.parse_args(argv)
    {"invert": stage_invert, "test": stage_test, "train": stage_train, "reference": stage_reference}[args.stage](args)


if __name__ == "__main__":
    main()
