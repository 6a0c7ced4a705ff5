"""Field inversion and the learned closure end to end on small channel bumps.

Truth: beta* = 1 + 0.3 tanh(c . f + c0) (the FP64 analytic closure with
seeded coefficients) gives the target wall Cf.  The adjoint-driven L-BFGS
inversion from beta = 1 must reduce the Cf mismatch; the closure network,
trained on the GPU and fused into the SA source, must predict the wall Cf of
an unseen bump better than plain SA (what the network gains over the method
without it)."""
import numpy as np
import pytest

from structured_amd import cases, fiml, lib

pytestmark = pytest.mark.gpu

GRID = dict(ni=81, nj=41, dy1=1e-5)


def bump(h, w):
    return cases.sa_bump(GRID["ni"], GRID["nj"], height=h, width=w, dy1=GRID["dy1"], re=3e6)


def test_inversion_reduces_mismatch():
    coef = fiml.truth_coef(seed=1)
    res = fiml.run_case(bump(0.05, 0.2), coef, invert_iterations=6, verbose=True)
    for k in ("sa_report", "true_report"):
        assert res[k]["converged"], (k, res[k]["iterations"], res[k]["l2w_final"])
    rep = res["inv_report"]
    print("inversion", {k: rep[k] for k in ("iterations", "evaluations", "failed_solves", "J0", "J", "wall_s",
                                            "primal_s", "adjoint_s")}, flush=True)
    print("J history", res["inv_hist"], flush=True)
    assert rep["iterations"] >= 3
    assert rep["J"] < 0.2 * rep["J0"]
    assert np.all(np.diff(res["inv_hist"]) <= 0)
    e_sa = fiml.cf_error(res["cf_sa"], res["cf_true"])
    e_inv = fiml.cf_error(res["cf_inv"], res["cf_true"])
    print("Cf error vs truth: plain SA %.3e, inversion %.3e" % (e_sa, e_inv), flush=True)
    assert e_inv < 0.5 * e_sa


def test_learned_closure_beats_plain_sa_on_unseen_bump():
    coef = fiml.truth_coef(seed=1)
    train_cases = [bump(0.05, 0.2), bump(0.06, 0.25)]
    results = [fiml.run_case(c, coef, invert_iterations=5, verbose=True) for c in train_cases]
    for r in results:
        assert r["sa_report"]["converged"] and r["true_report"]["converged"], r["case"]
    test_case = bump(0.045, 0.2)
    truth = fiml.run_case(test_case, coef, invert_iterations=0, verbose=True)
    e_sa = fiml.cf_error(truth["cf_sa"], truth["cf_true"])

    # (a) network trained on the truth beta at the truth solutions: the fused
    # closure must recover most of the truth's Cf correction
    X = np.concatenate([r["features_true"].reshape(-1, 5) for r in results]).astype(np.float32)
    T = np.concatenate([(r["beta_true"] - 1).reshape(-1) for r in results]).astype(np.float32)
    net, hist, t_train = fiml.train(X, T, epochs=400, batch=8192, lr=3e-3, seed=3)
    print("train (truth beta): loss %.3e -> %.3e in %.2fs" % (hist[0], hist[-1], t_train), flush=True)
    assert hist[-1] < 0.05 * hist[0]
    q, rep, t, cf = fiml.predict(test_case, net, q0=truth["q_sa"])
    assert fiml.closure_solve_ok(rep), (rep["iterations"], rep["l2w_final"])
    e_mlp = fiml.cf_error(cf, truth["cf_true"])
    print("unseen bump Cf error vs truth: plain SA %.3e, MLP(truth-trained) %.3e, solve %.2fs" % (e_sa, e_mlp, t),
          flush=True)
    assert e_mlp < 0.3 * e_sa

    # (b) network trained on the inverted beta fields (the field-inversion /
    # machine-learning loop): must still beat plain SA on the unseen bump
    X2, T2 = fiml.training_set(results)
    net2, hist2, _ = fiml.train(X2, T2, epochs=400, batch=8192, lr=3e-3, seed=4)
    q2, rep2, t2, cf2 = fiml.predict(test_case, net2, q0=truth["q_sa"])
    assert fiml.closure_solve_ok(rep2), (rep2["iterations"], rep2["l2w_final"])
    e_fiml = fiml.cf_error(cf2, truth["cf_true"])
    print("unseen bump Cf error vs truth: MLP(inversion-trained) %.3e, solve %.2fs" % (e_fiml, t2), flush=True)
    assert e_fiml < e_sa
