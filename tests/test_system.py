"""CPU: the seeded cluster generator (SURVEY's construction: beads of radius 1, each new bead
touching a bead of its own body at centre distance exactly 2, no two beads closer than 2,
cubic box L = (N (4 pi / 3) / phi)^(1/3)) at the batched workload's densest setting,
volume fraction 0.35 with mixed 8/20/40-bead bodies, which needs the dense growth."""
import math

import numpy as np
from scipy.spatial import cKDTree

from conductorrigid_amd.system import make_batch, make_system


def _check_packing(s, phi):
    N = s.N
    L = float(s.L[0])
    assert abs(L - (N * 4.0 * math.pi / 3.0 / phi) ** (1.0 / 3.0)) < 1e-9 * L
    x = s.pos[:, :3] + 0.5 * L
    tree = cKDTree(x, boxsize=L)
    close = tree.query_pairs(2.0 - 1e-7, output_type="ndarray")
    assert close.size == 0, "beads overlap"
    # every bead of a body touches another bead of the same body (one connected cluster)
    for r in range(s.N_rigid):
        n = int(s.rigid_size[r])
        mem = s.rigid_members[r * s.Nmax:r * s.Nmax + n].astype(np.int64)
        d = x[mem][:, None, :] - x[mem][None, :, :]
        d -= L * np.rint(d / L)
        touch = np.abs(np.linalg.norm(d, axis=2) - 2.0) < 1e-7
        seen = {0}
        frontier = [0]
        while frontier:
            i = frontier.pop()
            for j in np.nonzero(touch[i])[0]:
                if j not in seen:
                    seen.add(int(j))
                    frontier.append(int(j))
        assert len(seen) == n, f"body {r} is not one touching cluster"


def test_dense_growth_reaches_phi_035():
    s = make_system(100, (8, 20, 40), 0.35, seed=1, charges="uniform", dense_growth=True)
    s.validate()
    _check_packing(s, 0.35)
    assert set(np.unique(s.rigid_size).tolist()) <= {8, 20, 40}
    assert abs(s.body_charge.sum()) < 1e-9


def test_batch_boxes_span_the_volume_fractions():
    items = make_batch(6, seed=7, n_range=(300, 1200), phi_range=(0.33, 0.35))
    for s, p in items:
        s.validate()
        _check_packing(s, p["phi"])
        assert 0.33 <= p["phi"] <= 0.35
