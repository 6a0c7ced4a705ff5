"""Flow cases: grids, freestream, ordered boundary lists and options.

Non-dimensionalisation follows the reference's examples: rho_inf = 1,
p_inf = T_inf = 1/1.4 (so R = p/(rho T) = 1 and the freestream sound speed is
1), u_inf = Mach number, plate/bump length 1, mu_inf = u_inf / Re.

Boundary lists are applied in order, exactly like the reference's
BoundaryContainer (src/model/bc.cpp:430-433).  The lists below are ordered so
that every ghost cell another boundary condition reads (the corners read by
the right outflow) is written earlier in the same pass.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import grids


@dataclass
class Case:
    name: str
    xv: np.ndarray
    yv: np.ndarray
    fs: dict
    bcs: list
    order: int = 2
    turbulence: int = 0
    tail: int = 1
    dpdx: float = 0.0
    dpdy: float = 0.0
    flux: str = "roe"  # solver.flux: "roe" or "ausm"
    meta: dict = field(default_factory=dict)

    @property
    def ni(self):
        return self.xv.shape[0]

    @property
    def nj(self):
        return self.xv.shape[1]

    @property
    def nic(self):
        return self.ni - 1

    @property
    def njc(self):
        return self.nj - 1

    @property
    def nvar(self):
        return 5 if self.turbulence else 4

    def freestream_state(self) -> np.ndarray:
        """EulerEquation::initialize (src/model/eulerequation.cpp:262-277), AoS."""
        f = self.fs
        q = np.empty((self.nic, self.njc, self.nvar))
        q[..., 0] = f["rho_inf"]
        q[..., 1] = f["rho_inf"] * f["u_inf"]
        q[..., 2] = f["rho_inf"] * f["v_inf"]
        q[..., 3] = f["p_inf"] / (1.4 - 1.0) + 0.5 * f["rho_inf"] * (f["u_inf"] ** 2 + f["v_inf"] ** 2)
        if self.turbulence:
            q[..., 4] = f["rho_inf"] * f.get("nut_inf", 0.0)
        return q

    def perturbed_state(self, seed: int = 0, amp: float = 0.01) -> np.ndarray:
        """Freestream plus a seeded smooth perturbation of relative size amp
        (several Fourier modes per variable), away from the |.| kinks of the
        uniform state."""
        rng = np.random.default_rng(seed)
        q = self.freestream_state()
        xc = 0.25 * (self.xv[:-1, :-1] + self.xv[1:, :-1] + self.xv[:-1, 1:] + self.xv[1:, 1:])
        yc = 0.25 * (self.yv[:-1, :-1] + self.yv[1:, :-1] + self.yv[:-1, 1:] + self.yv[1:, 1:])
        lx = max(np.ptp(xc), 1e-12)
        ly = max(np.ptp(yc), 1e-12)
        scale = np.array([1.0, 0.2, 0.2, 2.0, 1e-3])[: self.nvar]
        base = np.abs(q).max(axis=(0, 1))
        for k in range(self.nvar):
            s = np.zeros_like(xc)
            for _ in range(4):
                kx, ky = rng.uniform(0.5, 3.0, size=2)
                ph = rng.uniform(0, 2 * np.pi, size=2)
                s += rng.uniform(-1, 1) * np.sin(2 * np.pi * kx * xc / lx + ph[0]) * np.cos(2 * np.pi * ky * yc / ly + ph[1])
            mag = max(base[k], scale[k])
            q[..., k] += amp * mag * s
        return q


def _fs(mach: float, re: float, nut_ratio: float = 0.0, pr: float = 0.72, aoa_deg: float = 0.0):
    mu = mach / re
    return dict(rho_inf=1.0, u_inf=mach, v_inf=0.0, p_inf=1.0 / 1.4, T_inf=1.0 / 1.4, mu_inf=mu, pr_inf=pr,
                aoa_deg=aoa_deg, nut_inf=nut_ratio * mu)


def plate_bcs(n_up: int):
    """Bottom: symmetry (slipwall) for the n_up cells with x < 0, adiabatic
    no-slip wall from the leading edge on; top freestream; left freestream;
    right outflow (after the bottom and top rows it copies at the corners)."""
    return [
        dict(type="slipwall", face="bottom", start=1, end=n_up),
        dict(type="wall", face="bottom", start=n_up + 1, end=-2, u=0.0, v=0.0),
        dict(type="freestream", face="top", start=0, end=-1),
        dict(type="freestream", face="left", start=0, end=-1),
        dict(type="outflow", face="right", start=0, end=-1),
    ]


def laminar_plate(ni: int = 97, nj: int = 65, dy1: float = 1e-4, mach: float = 0.2, re: float = 1e5) -> Case:
    """Laminar flat plate (97x65, M=0.2, Re_L=1e5, Pr=0.72, Roe + MUSCL)."""
    xv, yv, n_up = grids.flat_plate(ni, nj, dy1)
    return Case("laminar_plate_%dx%d" % (ni, nj), xv, yv, _fs(mach, re), plate_bcs(n_up), order=2,
                turbulence=0, tail=1, meta=dict(n_up=n_up))


def sa_plate(ni: int = 545, nj: int = 385, dy1: float = 1e-6, mach: float = 0.2, re: float = 5e6,
             nut_ratio: float = 3.0) -> Case:
    """Turbulent flat plate with SA (545x385, dy1=1e-6, Re_L=5e6, nu~/nu=3)."""
    xv, yv, n_up = grids.flat_plate(ni, nj, dy1)
    return Case("sa_plate_%dx%d" % (ni, nj), xv, yv, _fs(mach, re, nut_ratio), plate_bcs(n_up), order=2,
                turbulence=1, tail=1, meta=dict(n_up=n_up))


def bump_bcs():
    """Lower wall no-slip over the whole channel, top symmetry (slipwall),
    left freestream inflow, right outflow."""
    return [
        dict(type="wall", face="bottom", start=1, end=-2, u=0.0, v=0.0),
        dict(type="slipwall", face="top", start=1, end=-2),
        dict(type="freestream", face="left", start=0, end=-1),
        dict(type="outflow", face="right", start=0, end=-1),
    ]


def sa_bump(ni: int = 705, nj: int = 321, height: float = 0.05, width: float = 0.2, dy1: float = 1e-6,
            mach: float = 0.2, re: float = 3e6, nut_ratio: float = 3.0) -> Case:
    """Channel bump with SA (705x321 at height 0.05, Gaussian width 0.2)."""
    xv, yv = grids.bump_channel(ni, nj, dy1, height, width)
    return Case("sa_bump_%dx%d" % (ni, nj), xv, yv, _fs(mach, re, nut_ratio), bump_bcs(), order=2, turbulence=1,
                tail=1, meta=dict(height=height, width=width))


def blob_beta(case: Case, seed: int, nblob: int = 8, amp: float = 0.2) -> np.ndarray:
    """Per-cell beta = 1 + the sum of nblob seeded Gaussian blobs of
    amplitude +-amp and radius 0.05-0.2, centred near the lower wall."""
    rng = np.random.default_rng(seed)
    xc = 0.25 * (case.xv[:-1, :-1] + case.xv[1:, :-1] + case.xv[:-1, 1:] + case.xv[1:, 1:])
    yc = 0.25 * (case.yv[:-1, :-1] + case.yv[1:, :-1] + case.yv[:-1, 1:] + case.yv[1:, 1:])
    b = np.ones_like(xc)
    for _ in range(nblob):
        cx, cy = rng.uniform(-1.0, 1.0), rng.uniform(0.0, 0.1)
        r = rng.uniform(0.05, 0.2)
        b += rng.choice([-amp, amp]) * np.exp(-((xc - cx) ** 2 + (yc - cy) ** 2) / r ** 2)
    return b


def channel_example(path_grid: str | None = None) -> Case:
    """The reference's laminar channel example (examples/laminar/channel,
    3x101 grid, isothermal walls, periodic in x, body force dpdx)."""
    if path_grid is None:
        xv = np.repeat(np.array([0.0, 1e-3, 2e-3])[:, None], 101, axis=1)
        yv = np.repeat((np.arange(101) * 1e-3)[None, :], 3, axis=0)
    else:
        xv, yv = grids.read_p3d(path_grid, 3, 101)
    fs = dict(rho_inf=1.0, u_inf=0.01, v_inf=0.0, p_inf=0.7142857142857143, T_inf=0.7142857142857143, mu_inf=1e-2,
              pr_inf=0.7, aoa_deg=0.0, nut_inf=0.0)
    T = 0.7142857142857143
    bcs = [dict(type="isothermalwall", face="top", start=1, end=-2, u=0.0, v=0.0, T=T),
           dict(type="isothermalwall", face="bottom", start=1, end=-2, u=0.0, v=0.0, T=T),
           dict(type="periodic", face="left", start=0, end=-1)]
    return Case("channel_3x101", xv, yv, fs, bcs, order=2, turbulence=0, tail=1, dpdx=-0.02592)


def naca0012_example(xv: np.ndarray, yv: np.ndarray) -> Case:
    """The reference's laminar NACA 0012 example (examples/laminar/naca0012:
    281x151 C-grid in the "simple" format, Mach 0.5 at 1 degree, mu_inf 1e-4,
    freestream far field, wake cut on the bottom cells 1-40, wall on 41-240,
    tail 41), on the grid given (Mesh::simple_loader's xv, yv)."""
    fs = dict(rho_inf=1.0, u_inf=0.49992384757819563, v_inf=0.0087262032186417558, p_inf=0.7142857142857143,
              T_inf=0.7142857142857143, mu_inf=1e-4, pr_inf=0.7, aoa_deg=1.0, nut_inf=0.0)
    bcs = [dict(type="freestream", face="top", start=0, end=-1),
           dict(type="freestream", face="left", start=0, end=-1),
           dict(type="freestream", face="right", start=0, end=-1),
           dict(type="wake", face="bottom", start=1, end=40),
           dict(type="wall", face="bottom", start=41, end=240, u=0.0, v=0.0)]
    return Case("naca0012_281x151", np.ascontiguousarray(xv), np.ascontiguousarray(yv), fs, bcs, order=2,
                turbulence=0, tail=41)


def to_toml(case: Case, grid_file: str, label: str = "case", solver: dict | None = None) -> str:
    """Reference-format configuration (src/utils/config.cpp:31-87 keys).
    `solver`: extra [solver] keys (scheme, cfl, iteration_max, cfl_ramp, ...)."""
    f = case.fs
    out = ["[geometry]", 'filename = "%s"' % grid_file, "ni = %d" % case.ni, "nj = %d" % case.nj,
           "tail = %d" % case.tail, 'format = "p3d"', "",
           "[freestream]"]
    for k in ("rho_inf", "u_inf", "v_inf", "p_inf", "T_inf", "mu_inf", "pr_inf"):
        out.append("%s = %.17e" % (k, f[k]))
    out.append("aoa = %.17e" % f.get("aoa_deg", 0.0))
    out += ["", "[source]", "dpdx = %.17e" % case.dpdx, "dpdy = %.17e" % case.dpdy, "",
            "[solver]", "order = %d" % case.order, "lhs_order = %d" % case.order, 'flux = "%s"' % case.flux]
    sv = {"iteration_max": 1, **(solver or {})}
    for k, v in sv.items():
        if isinstance(v, bool):
            out.append("%s = %s" % (k, "true" if v else "false"))
        elif isinstance(v, str):
            out.append('%s = "%s"' % (k, v))
        elif isinstance(v, int):
            out.append("%s = %d" % (k, v))
        else:
            out.append("%s = %.17e" % (k, v))
    out += ["", "[io]", 'label = "%s"' % label, "restart = false", "stdout_frequency = 1000000",
            "fileout_frequency = 1000000", ""]
    for b in case.bcs:
        out += ["[[boundary]]", 'name = ""', 'type = "%s"' % b["type"], 'face = "%s"' % b["face"],
                "start = %d" % b.get("start", 0), "end = %d" % b.get("end", 0)]
        for k in ("u", "v", "T"):
            if k in b:
                out.append("%s = %.17e" % (k, b[k]))
        out.append("")
    return "\n".join(out)
