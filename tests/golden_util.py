"""Loads the committed golden fixtures (tests/golden/*.npz + index.json)."""
import json
import os

import numpy as np

from structured_amd import cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    with open(os.path.join(HERE, "index.json")) as f:
        return sorted(json.load(f).keys())


def load(name):
    with open(os.path.join(HERE, "index.json")) as f:
        meta = json.load(f)[name]
    z = np.load(os.path.join(HERE, name + ".npz"))
    case = cases.Case(name, z["xv"], z["yv"], meta["fs"], meta["bcs"], order=meta["order"],
                      turbulence=meta["turbulence"], tail=meta["tail"], dpdx=meta["dpdx"], dpdy=meta["dpdy"],
                      flux=meta.get("flux", "roe"))
    return case, {k: z[k] for k in z.files}, meta


def explicit_names():
    with open(os.path.join(HERE, "explicit.json")) as f:
        return sorted(json.load(f).keys())


def load_explicit(name):
    """(case, q0, solver keys, the reference's final q) of an explicit run."""
    with open(os.path.join(HERE, "explicit.json")) as f:
        meta = json.load(f)[name]
    case, z, _ = load(meta["base"])
    q = np.load(os.path.join(HERE, name + ".npz"))["q"]
    return case, z["q"], meta["solver"], q


def load_naca():
    """(case, seeded state, fixture arrays) of the reference's NACA 0012
    example fixture; R_sub / dt_sub hold every sub-th cell."""
    z = np.load(os.path.join(HERE, "naca0012_281x151.npz"))
    case = cases.naca0012_example(z["xv"], z["yv"])
    q = case.perturbed_state(seed=int(z["seed"]))
    return case, q, {k: z[k] for k in z.files}
