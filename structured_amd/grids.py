"""Structured grids and on-disk formats.

Vertex arrays are ``xv[ni][nj]`` / ``yv[ni][nj]`` (nj fastest), as the
reference's ``Mesh::xv`` (src/utils/mesh.cpp:344-389).  The writers/readers
follow the reference loaders: ``p3d`` (src/utils/mesh.cpp:146-169: block
count, ni nj, then every x with i fastest, then every y) and ``simple``
(:133-144: x y pairs, i fastest).
"""
from __future__ import annotations

import numpy as np


def geometric_ratio(dy1: float, ncell: int, height: float = 1.0) -> float:
    """Ratio r with dy1 * (r**ncell - 1) / (r - 1) == height (bisection)."""
    if dy1 * ncell >= height:
        raise ValueError("first spacing too large for a stretched distribution")
    lo, hi = 1.0 + 1e-14, 2.0
    while dy1 * (hi ** ncell - 1.0) / (hi - 1.0) < height:
        hi *= 1.5
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if dy1 * (mid ** ncell - 1.0) / (mid - 1.0) < height:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def stretched(dy1: float, ncell: int, height: float = 1.0) -> np.ndarray:
    """ncell+1 points on [0, height] with first spacing dy1, geometric growth."""
    r = geometric_ratio(dy1, ncell, height)
    d = dy1 * r ** np.arange(ncell)
    y = np.concatenate([[0.0], np.cumsum(d)])
    y *= height / y[-1]
    return y


def flat_plate(ni: int, nj: int, dy1: float, x0: float = -0.33, x1: float = 2.0, ymax: float = 1.0,
               frac_upstream: float = 0.25):
    """Flat-plate grid: uniform cells on [x0, 0] and on [0, x1] with a vertex
    at the leading edge x = 0; geometric stretching in y from dy1.

    Returns (xv, yv, n_up) where n_up is the number of cells upstream of the
    plate (x < 0)."""
    ncx = ni - 1
    n_up = max(1, int(round(ncx * frac_upstream)))
    xs = np.concatenate([np.linspace(x0, 0.0, n_up + 1)[:-1], np.linspace(0.0, x1, ncx - n_up + 1)])
    ys = stretched(dy1, nj - 1, ymax)
    xv = np.repeat(xs[:, None], nj, axis=1).copy()
    yv = np.repeat(ys[None, :], ni, axis=0).copy()
    return xv, yv, n_up


def bump_channel(ni: int, nj: int, dy1: float, height: float, width: float, x0: float = -1.5, x1: float = 1.5,
                 ytop: float = 1.0):
    """Channel with lower wall y = height * exp(-(x/width)^2) on [x0, x1] and a
    flat top at ytop; each i-line is stretched from the wall with first
    spacing dy1 scaled to the local channel height."""
    xs = np.linspace(x0, x1, ni)
    yw = height * np.exp(-(xs / width) ** 2)
    s = stretched(dy1 / (ytop - 0.0), nj - 1, 1.0)
    xv = np.repeat(xs[:, None], nj, axis=1).copy()
    yv = yw[:, None] + (ytop - yw[:, None]) * s[None, :]
    return xv, yv


def write_p3d(path: str, xv: np.ndarray, yv: np.ndarray) -> None:
    ni, nj = xv.shape
    with open(path, "w") as f:
        f.write("1\n%d %d\n" % (ni, nj))
        for arr in (xv, yv):
            for v in arr.T.reshape(-1):
                f.write("%.20e\n" % v)


def read_p3d(path: str, ni: int, nj: int):
    tok = open(path).read().split()
    if not tok:
        raise ValueError("empty grid file")
    nblock, ni_, nj_ = int(tok[0]), int(tok[1]), int(tok[2])
    if nblock != 1 or ni_ != ni or nj_ != nj:
        raise ValueError("grid header %s %s %s does not match ni=%d nj=%d" % (nblock, ni_, nj_, ni, nj))
    vals = np.array(tok[3:3 + 2 * ni * nj], dtype=np.float64)
    if vals.size != 2 * ni * nj:
        raise ValueError("grid file truncated")
    xv = vals[:ni * nj].reshape(nj, ni).T.copy()
    yv = vals[ni * nj:].reshape(nj, ni).T.copy()
    return xv, yv


def read_simple(path: str, ni: int, nj: int):
    vals = np.array(open(path).read().split(), dtype=np.float64)
    if vals.size < 2 * ni * nj:
        raise ValueError("grid file truncated")
    pairs = vals[:2 * ni * nj].reshape(nj, ni, 2)
    return pairs[:, :, 0].T.copy(), pairs[:, :, 1].T.copy()
