"""GPU: state that carries across steps, and run-to-run reproducibility.

* Checkpoint: save_state / load_state keep the bead charges and the body potentials
  (the reference loses the potentials across a restart, cppmodule/ConductorRigid.h:67); a
  solve from the restored state starts at the converged answer.
* FP32 bulk is bit-reproducible: the fixed-point spread sums integers, the contraction,
  pair sum, preconditioner and Krylov reductions have fixed orders, so two evaluations --
  on one context or on two -- give the same bits."""
import copy
import os

import numpy as np
import pytest

from helpers import n1k_system, rel, small_system

pytestmark = pytest.mark.gpu

SMALL = dict(xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7)
N1K = dict(xi=0.929, errortol=1e-6, rc=4.0, mesh=(64, 64, 64), P=8)


def _cr(s, **kw):
    from conductorrigid_amd.compute import ConductorRigid

    base = dict(field=(0.0, 0.0, 1.0), precision="fp64", gmres_tol=1e-12, restart=30, max_iter=600,
                preconditioner="block")
    base.update(SMALL)
    base.update(kw)
    return ConductorRigid(system=s, **base)


def test_checkpoint_round_trip(tmp_path, require_gpu):
    s = small_system(seed=51)
    cr = _cr(copy.deepcopy(s), field=(0.3, -0.2, 1.1))
    cr.computeForces(0, x0="zero")
    q0, P0, F0 = cr.get_charges(), cr.get_potentials(), cr.get_forces()
    path = os.path.join(tmp_path, "state.npz")
    cr.save_state(path)
    cr2 = _cr(copy.deepcopy(s))  # default field: the checkpoint restores it
    cr2.load_state(path)
    assert np.array_equal(cr2.get_charges(), q0) and np.array_equal(cr2.get_potentials(), P0)
    st = cr2.computeForces(0, x0="previous")
    print(f"[state] restored start: {st['iterations']} iterations", flush=True)
    assert st["iterations"] <= 1 and st["fallback"] == 0
    assert rel(cr2.get_charges(), q0) < 1e-10 and rel(cr2.get_forces()[:, :3], F0[:, :3]) < 1e-10
    # a checkpoint of another system size is refused
    other = _cr(small_system(seed=52, isolated=3))
    with pytest.raises(ValueError):
        other.load_state(path)


@pytest.mark.parametrize("sysname", ["small", "n1k"])
def test_fp32_bulk_is_bit_reproducible(sysname, require_gpu):
    s, prm = (small_system(seed=53), SMALL) if sysname == "small" else (n1k_system(), N1K)
    a = _cr(copy.deepcopy(s), precision="fp32", gmres_tol=1e-6, **prm)
    runs = []
    for _ in range(2):
        st = a.computeForces(0, x0="zero")
        runs.append((a.get_charges(), a.get_potentials(), a.get_forces(), st["iterations"]))
    b = _cr(copy.deepcopy(s), precision="fp32", gmres_tol=1e-6, **prm)
    st = b.computeForces(0, x0="zero")
    runs.append((b.get_charges(), b.get_potentials(), b.get_forces(), st["iterations"]))
    for r in runs[1:]:
        assert r[3] == runs[0][3]
        for x, y in zip(r[:3], runs[0][:3]):
            assert np.array_equal(x.view(np.int64), y.view(np.int64))
