"""The data files the bench's closure leg reads (structured_amd/data/):

- closure_net.npz: the 5-64-64-1 network trained on the GPU on the inverted
  beta fields of the 24 training bumps; its parameters have the library's
  layout and size, are finite, the loss history fell by two orders, and the
  network's beta - 1 on features across their range stays in the band the
  truth closure spans (0.3 tanh(.) -> |beta - 1| <= 0.3, with margin);
- reference_<grid>.npz (written by `fiml reference`, beta field dropped): the
  inversion's wall Cf is closer to the truth's than plain SA's, and its
  iteration count and wall time are recorded.
Loaded with numpy's default (no pickles)."""
import os

import numpy as np
import pytest

import mlp_model as mm

DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "structured_amd", "data")


def test_closure_network_file():
    z = np.load(os.path.join(DATA, "closure_net.npz"))
    p, loss = z["params"], z["loss"]
    assert p.dtype == np.float32 and p.shape == (mm.NPARAM,)
    assert np.all(np.isfinite(p))
    assert loss[-1] < 1e-2 * loss[0]
    f = np.random.default_rng(0).uniform(0.0, 1.0, size=(4096, 5)).astype(np.float32)
    f[:, 3] = 2.0 * f[:, 3] - 1.0  # f3 = tanh(.) spans [-1, 1]
    b = mm.forward(p, f, bf=False)[0].reshape(-1)
    assert np.all(np.isfinite(b)) and np.abs(b).max() < 0.6, np.abs(b).max()


@pytest.mark.parametrize("grid", ["1409x641"])
def test_reference_inversion_file(grid):
    path = os.path.join(DATA, "reference_%s.npz" % grid)
    if not os.path.exists(path):
        pytest.skip("reference inversion not recorded for %s" % grid)
    z = np.load(path)
    cf_inv, cf_true, cf_sa = z["cf_inv"], z["cf_true"], z["cf_sa"]
    assert cf_inv.shape == cf_true.shape == cf_sa.shape and cf_inv.ndim == 1
    err = lambda a: np.linalg.norm(a - cf_true) / np.linalg.norm(cf_true)
    assert err(cf_inv) < 0.1 * err(cf_sa), (err(cf_inv), err(cf_sa))
    assert int(z["iterations_done"]) >= 1 and float(z["inversion_s"]) > 0.0
    jh = z["J_hist"]
    assert np.all(np.diff(jh[np.isfinite(jh)]) <= 0.0)
