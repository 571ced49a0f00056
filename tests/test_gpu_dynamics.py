"""The consumers around the force compute: the Brownian rigid-body step (cr_bd_step)
and box changes picked up by computeForces."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel, small_system

pytestmark = pytest.mark.gpu

SMALL = dict(xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7)


def _cr(s, **kw):
    from conductorrigid_amd.compute import ConductorRigid

    return ConductorRigid(system=s, field=(0.0, 0.0, 1.0), precision="fp64", gmres_tol=1e-12, restart=30,
                          max_iter=600, preconditioner="block", **{**SMALL, **kw})


def _mimg(d, L):
    return d - L * np.round(d / L)


def _rotate(v, th):
    """Rodrigues rotation of rows v by the rotation vector th."""
    a = np.linalg.norm(th)
    if a == 0.0:
        return v.copy()
    k = th / a
    return v * np.cos(a) + np.cross(k, v) * np.sin(a) + np.outer(v @ k, k) * (1 - np.cos(a))


def test_bd_step_moves_rigid_bodies_rigidly(require_gpu):
    """Without noise and contact repulsion a step is deterministic: each rigid body
    translates by dt F / gamma_t and rotates about its centre by dt T / gamma_r (F, T the
    sums of its bead forces and their torques about the centre); free beads move by
    dt f / gamma_t."""
    from conductorrigid_amd import _lib

    s = small_system(seed=13)
    cr = _cr(s)
    cr.computeForces(0, x0="zero")
    torch = cr.torch
    F = cr.get_forces()[:, :3]
    pos0 = cr.dsys.pos.cpu().numpy()[:, :3].copy()
    com0 = cr.dsys.rigid_com.cpu().numpy()[:, :3].copy()
    L = s.L
    dt, gt, gr = 1e-3, 1.0, 20.0
    zr = torch.zeros(max(1, 6 * s.N_rigid), dtype=torch.float64, device="cuda")
    zb = torch.zeros(max(1, 3 * s.N), dtype=torch.float64, device="cuda")
    sv = cr.dsys.view()
    _lib.check(cr.lib.cr_bd_step(cr.ctx, C.byref(sv), cr.dsys.pos.data_ptr(), cr.dsys.rigid_com.data_ptr(),
                                 cr.forces.data_ptr(), zr.data_ptr(), zb.data_ptr(), dt, 1.0, gt, gr, 0.0,
                                 cr.stream()))
    torch.cuda.synchronize()
    pos1 = cr.dsys.pos.cpu().numpy()[:, :3]
    com1 = cr.dsys.rigid_com.cpu().numpy()[:, :3]
    worst = 0.0
    for r in range(s.N_rigid):
        mem = s.rigid_members[r * s.Nmax:r * s.Nmax + int(s.rigid_size[r])].astype(np.int64)
        v = _mimg(pos0[mem] - com0[r], L)
        Fb = F[mem].sum(0)
        Tb = np.cross(v, F[mem]).sum(0)
        R1 = com0[r] + dt * Fb / gt
        want = _rotate(v, dt * Tb / gr) + R1
        worst = max(worst, np.abs(_mimg(pos1[mem] - want, L)).max(), np.abs(_mimg(com1[r] - R1, L)).max())
        # rigid: pairwise distances unchanged (offsets from the centre are already unwrapped;
        # bead pairs of a body may be more than L/2 apart in this small box)
        d0 = np.linalg.norm(v[:, None] - v[None], axis=2)
        v1 = _mimg(pos1[mem] - com1[r], L)
        d1 = np.linalg.norm(v1[:, None] - v1[None], axis=2)
        assert np.abs(d1 - d0).max() < 1e-12
    free = np.nonzero(s.rigid_membership == 0xFFFFFFFF)[0]
    if free.size:
        worst = max(worst, np.abs(_mimg(pos1[free] - (pos0[free] + dt * F[free] / gt), L)).max())
    print(f"[bd] max deviation from the rigid-body update: {worst:.2e}", flush=True)
    assert worst < 1e-12


def test_box_change_is_picked_up(require_gpu):
    """computeForces after an affine rescale of the box gives what a context built on the
    rescaled system gives (mesh kept, spacing, widths, k-scale and neighbour list rebuilt)."""
    import copy

    from conductorrigid_amd.compute import update_positions

    s = small_system(seed=17)
    cr = _cr(s)
    cr.computeForces(0, x0="zero")
    f = 1.04
    s2 = copy.deepcopy(s)
    s2.L = s.L * f
    s2.pos = s.pos.copy()
    s2.pos[:, :3] *= f
    s2.rigid_com = s.rigid_com.copy()
    s2.rigid_com[:, :3] *= f
    cr.system.L = s2.L.copy()
    update_positions(cr, s2.pos, s2.rigid_com)
    cr.computeForces(0, x0="zero")
    ref = _cr(copy.deepcopy(s2))
    ref.computeForces(0, x0="zero")
    vq = rel(cr.get_charges(), ref.get_charges())
    vF = rel(cr.get_forces()[:, :3], ref.get_forces()[:, :3])
    print(f"[box] rescaled box: charges {vq:.1e}, forces {vF:.1e}", flush=True)
    assert vq < 1e-9 and vF < 1e-9
    # against the CPU oracle built on the rescaled system (spacing, widths, k-scale follow the box)
    import oracle as O

    o = O.Oracle(s2, SMALL["xi"], SMALL["errortol"], rc=SMALL["rc"], mesh=SMALL["mesh"], P=SMALL["P"],
                 nlist=cr.dsys.nlist_host())
    qo, Po, Fo, so = o.compute_forces([0.0, 0.0, 1.0], restart=30, maxiter=600, tol=1e-12)
    vq, vP, vF = rel(cr.get_charges(), qo), rel(cr.get_potentials(), Po), rel(cr.get_forces()[:, :3], Fo[:, :3])
    print(f"[box] rescaled box vs oracle: charges {vq:.1e}, potentials {vP:.1e}, forces {vF:.1e}", flush=True)
    assert vq < 1e-10 and vP < 1e-10 and vF < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_set_box_regenerates_on_device(precision, require_gpu):
    """cr_set_box on a stretched (non-cubic) and then a cubic box: the context's k-space
    step and matvec equal the oracle's for that box (tables regenerated on the device)."""
    import copy

    import oracle as O
    from helpers import GpuCase

    s = small_system(seed=19)
    g = GpuCase(s, precision=precision, **SMALL)
    tol = 1e-12 if precision == "fp64" else 2e-6
    for f in ((1.03, 0.98, 1.01), (0.97, 0.97, 0.97)):
        s2 = copy.deepcopy(s)
        s2.L = s.L * np.asarray(f)
        s2.pos = s.pos.copy()
        s2.pos[:, :3] *= np.asarray(f)
        s2.rigid_com = s.rigid_com.copy()
        s2.rigid_com[:, :3] *= np.asarray(f)
        _lib_set_box(g, s2.L)
        from conductorrigid_amd.compute import update_positions

        g.cr.system.L = s2.L.copy()
        update_positions(g.cr, s2.pos, s2.rigid_com)
        o = O.Oracle(s2, SMALL["xi"], SMALL["errortol"], rc=SMALL["rc"], mesh=SMALL["mesh"], P=SMALL["P"],
                     nlist=g.cr.dsys.nlist_host())
        grid = np.random.default_rng(3).normal(size=SMALL["mesh"])
        vr = rel(g.recip(grid), o.recip_grid(grid))
        x = np.random.default_rng(4).normal(size=g.n)
        vm = rel(g.matvec(x), o.matvec(x))
        print(f"[box] {precision} box {f}: recip {vr:.1e} matvec {vm:.1e}", flush=True)
        assert vr < tol and vm < max(tol, 1e-12 if precision == "fp64" else 1e-5)


def _lib_set_box(g, L):
    from conductorrigid_amd import _lib

    _lib.check(g.lib.cr_set_box(g.ctx, (C.c_double * 3)(*[float(v) for v in L]), g.st()))
    g.torch.cuda.synchronize()
