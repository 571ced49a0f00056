"""GPU: the headline workload itself (N = 1,000,000 beads in 50,000 rigid bodies of 20,
volume fraction 0.3, mesh 256^3, P = 8, rc = 4, xi = 0.6) through the C ABI.

The headline runs the 256-point register FFT passes, the fixed-point spread on 256^3 and
the 1M-bead ELL operator; here they meet the CPU oracle at full size:
  * one FP64 operator application against the oracle's FP64 matvec (1e-12);
  * the FP32-bulk operator against the FP64 one (1e-5);
  * symmetry of the FP64 operator, <A x, y> = <x, A y> (up to rounding);
  * an FP32-bulk solve: converged, FP64 true residual at the tolerance, and its per-body
    constraint rows (sum of the body's charges = its net charge) to the solver tolerance.
"""
import ctypes as C
import sys
import time

import numpy as np
import pytest

import oracle as O
from helpers import rel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[n1m] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def n1m(require_gpu):
    from conductorrigid_amd.system import make_workload

    t0 = time.perf_counter()
    s, w = make_workload("n1m")
    _say(f"system N={s.N} bodies={s.N_isolated} L={s.L[0]:.3f} ({time.perf_counter() - t0:.1f} s)")
    prm = dict(xi=w["xi"], errortol=w["errortol"], rc=w["rc"], mesh=w["mesh"], P=w["P"])
    return s, w, prm


def _case(s, w, prm, precision):
    from helpers import GpuCase

    return GpuCase(s, precision=precision, restart=w["restart"], gmres_tol=w["tol"], max_iter=1000,
                   preconditioner="block", **prm)


def test_n1m_matvec_matches_oracle(n1m):
    s, w, prm = n1m
    g64 = _case(s, w, prm, "fp64")
    nl = g64.cr.dsys.nlist_host()
    _say(f"pairs {int(nl[1].sum())}")
    t0 = time.perf_counter()
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"], nlist=nl)
    x = np.random.default_rng(41).normal(size=g64.n)
    yo = o.matvec(x)
    _say(f"oracle matvec {time.perf_counter() - t0:.1f} s")
    y64 = g64.matvec(x)
    v = rel(y64, yo)
    _say(f"fp64 matvec vs oracle {v:.3e}")
    assert v < 1e-12
    # symmetry of the FP64 operator
    z = np.random.default_rng(42).normal(size=g64.n)
    az = g64.matvec(z)
    a, b = float(np.dot(y64, z)), float(np.dot(x, az))
    sym = abs(a - b) / (np.linalg.norm(y64) * np.linalg.norm(z))
    _say(f"fp64 symmetry {sym:.3e}")
    assert sym < 1e-12
    del g64
    g32 = _case(s, w, prm, "fp32")
    y32 = g32.matvec(x)
    v32 = rel(y32, yo)
    _say(f"fp32 matvec vs oracle {v32:.3e}")
    assert v32 < 1e-5


def test_n1m_fp32_solve_constraints(n1m):
    s, w, prm = n1m
    g32 = _case(s, w, prm, "fp32")
    q, Phi, F, st = g32.solve("zero")
    _say(f"fp32 solve iters {st['iterations']} conv {st['converged']} true_relres {st['true_relres']:.3e}")
    assert st["converged"] == 1 and st["fallback"] == 0
    assert st["true_relres"] < 2.0 * w["tol"]
    sums = np.bincount(s.iso_membership(), weights=q, minlength=s.N_isolated)
    dev = float(np.max(np.abs(sums - s.body_charge)))
    _say(f"constraint rows max |sum q - Q| {dev:.3e}")
    # the rows are part of the residual GMRES drives below tol * ||b||
    assert dev < 1e-3
    assert np.all(np.isfinite(F))
