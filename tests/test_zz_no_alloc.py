"""sg_residual_dev_async_noalloc (static device scratch, no allocation)
against sg_residual_dev_async (context scratch): the same residual and
norms, bit for bit, on two calls in a row with the second grid smaller, then
the first grid again; a grid beyond the scratch's tile limit is refused."""
import ctypes as C

import numpy as np
import pytest

from structured_amd import cases, lib

pytestmark = pytest.mark.gpu


def _device_state(eq, q, torch):
    dev = torch.device("cuda:0")
    aos = torch.from_numpy(np.ascontiguousarray(q).reshape(-1)).to(dev)
    qd = torch.zeros(eq.nvar * eq.plane, dtype=torch.float64, device=dev)
    eq.aos_to_soa_dev(aos.data_ptr(), qd.data_ptr())
    eq.synchronize()
    return qd


def _run(eq, fn, qd, torch):
    rd = torch.full_like(qd, -1.0)
    l2 = torch.full((eq.nvar + 2,), -7.0, dtype=torch.float64, device=qd.device)
    rc = fn(eq.h, C.c_void_p(qd.data_ptr()), C.c_void_p(rd.data_ptr()), C.c_void_p(l2.data_ptr()))
    assert rc == 0, lib.load().sg_last_error(eq.h)
    eq.synchronize()
    return rd.cpu().numpy(), l2.cpu().numpy()


def test_noalloc_residual_matches_context_scratch():
    torch = pytest.importorskip("torch")
    so = lib.load()
    big = cases.sa_plate(161, 97, dy1=1e-4)
    small = cases.laminar_plate(33, 21)
    eq_big, eq_small = lib.Equation(big), lib.Equation(small)
    qb = _device_state(eq_big, big.perturbed_state(seed=1), torch)
    qs = _device_state(eq_small, small.perturbed_state(seed=2), torch)
    first = None
    for eq, qd in ((eq_big, qb), (eq_small, qs), (eq_big, qb)):
        r_ref, l2_ref = _run(eq, so.sg_residual_dev_async, qd, torch)
        r, l2 = _run(eq, so.sg_residual_dev_async_noalloc, qd, torch)
        nv = eq.nvar
        print(eq.case.name, "l2", l2[:nv])
        assert np.array_equal(r, r_ref)
        assert np.array_equal(l2[:nv], l2_ref[:nv]) and np.all(np.isfinite(l2[:nv]))
        assert np.all(l2[nv:] == -7.0)  # exactly nvar norms written
        if eq is eq_big:
            if first is None:
                first = (r, l2)
            else:
                assert np.array_equal(first[0], r) and np.array_equal(first[1], l2)
    # without a norm buffer nothing of the scratch is used
    rd = torch.full_like(qs, -1.0)  # padding lanes as _run leaves them
    assert so.sg_residual_dev_async_noalloc(eq_small.h, C.c_void_p(qs.data_ptr()), C.c_void_p(rd.data_ptr()),
                                            None) == 0
    eq_small.synchronize()
    assert np.array_equal(rd.cpu().numpy(), _run(eq_small, so.sg_residual_dev_async, qs, torch)[0])
    eq_big.close()
    eq_small.close()


def test_noalloc_rejects_grid_beyond_scratch():
    torch = pytest.importorskip("torch")
    so = lib.load()
    ni, nj = 1985, 1800  # 64 x 257 = 16448 tiles of 31 x 7 cells > 16384
    xv = np.repeat(np.linspace(0.0, 2.0, ni)[:, None], nj, axis=1)
    yv = np.repeat(np.linspace(0.0, 1.0, nj)[None, :], ni, axis=0)
    fs = cases.laminar_plate(33, 21).fs
    bcs = [dict(type="freestream", face=f, start=0, end=-1) for f in ("bottom", "top", "left", "right")]
    case = cases.Case("uniform_%dx%d" % (ni, nj), xv, yv, fs, bcs)
    eq = lib.Equation(case)
    qd = torch.zeros(eq.nvar * eq.plane, dtype=torch.float64, device="cuda:0")
    l2 = torch.zeros(eq.nvar, dtype=torch.float64, device="cuda:0")
    rc = so.sg_residual_dev_async_noalloc(eq.h, C.c_void_p(qd.data_ptr()), C.c_void_p(qd.data_ptr()),
                                          C.c_void_p(l2.data_ptr()))
    assert rc == -1 and b"too large" in so.sg_last_error(eq.h)
    eq.close()
