"""Loader of the direct-Ewald golden fixture (tests/golden/direct_ewald_n1000.npz)."""
from __future__ import annotations

import os

import numpy as np

from conductorrigid_amd.system import make_system

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "direct_ewald_n1000.npz")


def load():
    """(system, fixture dict); the system is rebuilt from its seed and checked
    against the stored positions."""
    f = dict(np.load(PATH, allow_pickle=False))
    nb, bs, seed = (int(v) for v in f["system"])
    s = make_system(nb, bs, float(f["phi"]), seed=seed, charges="uniform")
    if not np.array_equal(s.pos, f["pos"]) or not np.array_equal(s.body_charge, f["body_charge"]):
        raise RuntimeError("seeded system differs from the fixture's positions")
    return s, f
