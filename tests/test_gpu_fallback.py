"""GPU: plain GMRES from x0 = 0 as the fallback and judge of a warm start, and the
single-precision boundary (cr_compute_force_f32).

A solve started from the caller's charges is redone from zero when it meets a
non-finite residual or stops at the iteration limit; the answer is then the zero-start
answer, with stats["fallback"] = 1.  The reference accepts whatever its solver returns
(cppmodule/ConductorRigid.cu:1142-1147)."""
import copy
import ctypes as C

import numpy as np
import pytest

from helpers import rel, small_system

pytestmark = pytest.mark.gpu

SMALL = dict(xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7)


def _cr(s, **kw):
    from conductorrigid_amd.compute import ConductorRigid

    base = dict(field=(0.0, 0.0, 1.0), precision="fp64", gmres_tol=1e-12, restart=30, max_iter=600,
                preconditioner="block")
    base.update(SMALL)
    base.update(kw)
    return ConductorRigid(system=s, **base)


def _zero_answer(s, **kw):
    cr = _cr(copy.deepcopy(s), **kw)
    st = cr.computeForces(0, x0="zero")
    assert st["fallback"] == 0
    return cr.get_charges(), cr.get_potentials(), cr.get_forces()[:, :3], st


def _check_same(cr, ref, tol=1e-10):
    q0, P0, F0, _ = ref
    vq, vP, vF = rel(cr.get_charges(), q0), rel(cr.get_potentials(), P0), rel(cr.get_forces()[:, :3], F0)
    print(f"[fallback] vs zero start: q {vq:.1e} Phi {vP:.1e} F {vF:.1e}", flush=True)
    assert vq < tol and vP < tol and vF < tol


@pytest.mark.parametrize("mode", ["previous", "charges"])
def test_nan_start_falls_back_to_zero(mode, require_gpu):
    s = small_system(seed=31)
    ref = _zero_answer(s)
    cr = _cr(copy.deepcopy(s))
    cr.charges.fill_(float("nan"))
    cr.potentials.fill_(float("nan"))
    st = cr.computeForces(0, x0=mode)
    print(f"[fallback] NaN {mode}: iters {st['iterations']} fallback {st['fallback']}", flush=True)
    assert st["fallback"] == 1 and st["converged"] == 1
    _check_same(cr, ref)


def test_iteration_limit_falls_back_to_zero(require_gpu):
    """A garbage start with a one-iteration limit: not converged, so the solve is redone
    from zero under the same limit and equals the zero start's one-iteration answer."""
    s = small_system(seed=32)
    ref = _zero_answer(s, max_iter=1)
    cr = _cr(copy.deepcopy(s), max_iter=1)
    rng = np.random.default_rng(5)
    g = cr.torch.from_numpy(rng.normal(size=s.N) * 50.0).to(cr.charges.device)
    cr.charges.copy_(g)
    cr.potentials.fill_(3.0)
    st = cr.computeForces(0, x0="previous")
    print(f"[fallback] garbage start, max_iter 1: iters {st['iterations']} fallback {st['fallback']}", flush=True)
    assert st["fallback"] == 1
    _check_same(cr, ref, tol=1e-10)


def test_untrained_network_start_gives_zero_start_answer(require_gpu):
    """The start of an untrained (random) network: the converged answer is the zero
    start's, whether or not the fallback was needed."""
    import conductorrigid_amd.warmstart as WS

    s = small_system(seed=33)
    ref = _zero_answer(s)
    cr = _cr(copy.deepcopy(s))
    mlp = WS.WarmStartMLP(device=0, seed=123)
    mlp.warm_start(cr)
    st = cr.computeForces(0, x0="charges")
    print(f"[fallback] untrained network start: iters {st['iterations']} fallback {st['fallback']}", flush=True)
    assert st["converged"] == 1
    _check_same(cr, ref)


def test_good_start_does_not_fall_back(require_gpu):
    s = small_system(seed=34)
    cr = _cr(copy.deepcopy(s))
    cr.computeForces(0, x0="zero")
    st = cr.computeForces(0, x0="previous")
    assert st["fallback"] == 0 and st["converged"] == 1 and st["iterations"] <= 1


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_single_precision_boundary(precision, require_gpu):
    """cr_compute_force_f32 with float arrays (as a SINGLE_PRECISION HOOMD build hands them
    over) against the double entry on the same float-rounded inputs: equal to float
    rounding of the outputs."""
    from conductorrigid_amd import _lib

    s = small_system(seed=35)
    s.pos = s.pos.astype(np.float32).astype(np.float64)
    s.rigid_com = s.rigid_com.astype(np.float32).astype(np.float64)
    s.body_charge = s.body_charge.astype(np.float32).astype(np.float64)
    tight = 1e-12 if precision == "fp64" else 1e-7
    cr = _cr(copy.deepcopy(s), precision=precision, gmres_tol=tight)
    cr.computeForces(0, x0="zero")
    q64, P64, F64 = cr.get_charges(), cr.get_potentials(), cr.get_forces()
    torch = cr.torch
    dev = cr.charges.device
    d = cr.dsys
    pos32 = d.pos.float().contiguous()
    com32 = d.rigid_com.float().contiguous()
    bq32 = d.body_charge.float().contiguous()
    q32 = torch.zeros(s.N, dtype=torch.float32, device=dev)
    phi32 = torch.zeros(max(1, s.N_isolated), dtype=torch.float32, device=dev)
    f32 = torch.zeros(s.N, 4, dtype=torch.float32, device=dev)
    sv = _lib.CrSystemF32(s.N, pos32.data_ptr(), d.tag.data_ptr(), d.body_of_tag.data_ptr(), bq32.data_ptr(),
                          s.N_rigid, s.N_isolated, max(1, s.Nmax), d.rigid_membership.data_ptr(),
                          d.rigid_members.data_ptr(), d.rigid_size.data_ptr(), com32.data_ptr(), d.nlist.data_ptr(),
                          d.head_list.data_ptr(), d.n_neigh.data_ptr(), (C.c_double * 3)(*[float(v) for v in s.L]))
    st = _lib.CrSolveStats()
    _lib.check(cr.lib.cr_compute_force_f32(cr.ctx, C.byref(sv), q32.data_ptr(), phi32.data_ptr(), f32.data_ptr(),
                                           1, C.byref(st), cr.stream()))
    torch.cuda.synchronize()
    vq = rel(q32.double().cpu().numpy(), q64)
    vP = rel(phi32.double().cpu().numpy()[:s.N_isolated], P64)
    vF = rel(f32.double().cpu().numpy()[:, :3], F64[:, :3])
    print(f"[f32 boundary] {precision}: q {vq:.1e} Phi {vP:.1e} F {vF:.1e} iters {st.iterations}", flush=True)
    lim = 1e-6 if precision == "fp64" else 1e-5
    assert st.converged == 1 and vq < lim and vP < lim and vF < lim
