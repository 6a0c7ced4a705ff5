"""C-ABI surface checks that need no GPU: the library loads, exports every
function declared in include/structured_gpu.h, and rejects malformed
boundary lists / shapes before touching the device."""
import ctypes as C
import os

import numpy as np
import pytest

from structured_amd import abi, cases, lib


def test_library_exports_every_declared_symbol():
    so = lib.load()
    names = lib.exported_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def _create(case, bcs=None, flux=0):
    so = lib.load()
    xv = np.ascontiguousarray(case.xv)
    yv = np.ascontiguousarray(case.yv)
    fs = abi.Freestream(**case.fs)
    b = abi.make_bcs(case.bcs if bcs is None else bcs)
    o = abi.Opts(order=case.order, turbulence=case.turbulence, tail=case.tail, flux=flux, dpdx=0.0, dpdy=0.0)
    h = C.c_void_p()
    rc = so.sg_create(case.ni, case.nj, xv.ctypes.data_as(C.c_void_p), yv.ctypes.data_as(C.c_void_p), C.byref(fs),
                      C.cast(b, C.c_void_p), len(case.bcs if bcs is None else bcs), C.byref(o), C.byref(h))
    msg = so.sg_last_error(None).decode()
    if rc == 0:
        so.sg_destroy(h)
    return rc, msg


def test_bc_read_before_write_rejected():
    """The right outflow copies the bottom-row ghost at its corner; listing it
    before the bottom wall makes the reference read stale memory."""
    case = cases.laminar_plate(21, 13)
    bad = [case.bcs[4]] + case.bcs[:4]
    rc, msg = _create(case, bad)
    assert rc == -2 and "before any boundary writes it" in msg


def test_bc_missing_corner_rejected_when_viscous():
    case = cases.laminar_plate(21, 13)
    bcs = [dict(type="slipwall", face="bottom", start=1, end=5),
           dict(type="wall", face="bottom", start=6, end=-2),
           dict(type="freestream", face="top", start=1, end=-2),
           dict(type="freestream", face="left", start=1, end=-2),
           dict(type="freestream", face="right", start=1, end=-2)]
    rc, msg = _create(case, bcs)
    assert rc == -2 and "corner" in msg


def test_slipwall_on_corner_rejected():
    case = cases.laminar_plate(21, 13)
    bcs = [dict(type="slipwall", face="bottom", start=0, end=5)] + case.bcs[1:]
    rc, msg = _create(case, bcs)
    assert rc == -2


def test_unknown_flux_rejected():
    """solver.flux other than roe / ausm (the reference logs "Flux not
    found.", src/model/eulerequation.cpp:126-127) is refused at creation."""
    rc, msg = _create(cases.laminar_plate(21, 13), flux=7)
    assert rc == -1 and "flux" in msg


def test_bad_shape_rejected():
    case = cases.laminar_plate(21, 13)
    so = lib.load()
    h = C.c_void_p()
    fs = abi.Freestream(**case.fs)
    o = abi.Opts(order=3)
    xv = np.zeros(4)
    rc = so.sg_create(2, 2, xv.ctypes.data_as(C.c_void_p), xv.ctypes.data_as(C.c_void_p), C.byref(fs), None, 0,
                      C.byref(o), C.byref(h))
    assert rc == -1
