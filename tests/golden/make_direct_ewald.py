"""Writes tests/golden/direct_ewald_n1000.npz: the FP64 direct-Ewald solution (bead
charges, body potentials, bead forces) of a seeded 1000-bead rigid-cluster system at
the N=1M workload's method parameters (xi = 0.6, real-space cutoff 4, volume
fraction 0.3, E0 = (0, 0, 1)).  Direct Ewald: closed-form shell-shell real space over
periodic images within the cutoff and an explicit k-sum to kmax = 5.4
(exp(-kmax^2 / 4 xi^2) = 1.6e-9), no mesh, no window, no table.

Run from the repository root:  python tests/golden/make_direct_ewald.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from conductorrigid_amd.system import make_system  # noqa: E402

SYSTEM = dict(n_bodies=50, body_sizes=20, phi=0.3, seed=21, charges="uniform")
XI, RC, KMAX, TOL = 0.6, 4.0, 5.4, 1e-11
E0 = np.array([0.0, 0.0, 1.0])


def main():
    O.build()
    s = make_system(SYSTEM["n_bodies"], SYSTEM["body_sizes"], SYSTEM["phi"], seed=SYSTEM["seed"],
                    charges=SYSTEM["charges"])
    o = O.Oracle(s, XI, 1e-6, rc=RC, mesh=(27, 27, 27), P=8)
    q, Phi, F, iters = o.direct_solve(E0, RC, KMAX, tol=TOL)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "direct_ewald_n1000.npz")
    np.savez_compressed(out, pos=s.pos, body_charge=s.body_charge, q=q, Phi=Phi, F=F[:, :3],
                        xi=XI, rc=RC, kmax=KMAX, tol=TOL, E0=E0, iters=iters,
                        system=np.array([SYSTEM["n_bodies"], SYSTEM["body_sizes"], SYSTEM["seed"]]),
                        phi=SYSTEM["phi"])
    print(f"wrote {out}: N={s.N} iters={iters}")


if __name__ == "__main__":
    main()
