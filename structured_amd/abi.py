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


class Opts(C.Structure):
    _fields_ = [("order", C.c_int), ("turbulence", C.c_int), ("tail", C.c_int), ("reserved", C.c_int),
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
