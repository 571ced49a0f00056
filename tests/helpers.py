"""Shared helpers for the parity tests: systems, GPU context wrappers."""
from __future__ import annotations

import ctypes as C

import numpy as np

from conductorrigid_amd import _lib
from conductorrigid_amd.system import make_system


def rel(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return float(d / n) if n > 0 else float(d)


def small_system(seed=7, isolated=5, shuffle=True):
    """A few hundred beads: bodies of 8/20 beads plus isolated beads, tags permuted."""
    s = make_system(12, [8, 20], 0.2, seed=seed, charges="uniform", n_isolated_beads=isolated,
                    spatial_sort=shuffle)
    return s


def n1k_system():
    return make_system(50, 20, 0.2, seed=1, charges="alternating")


class GpuCase:
    """A HIP context + device system for one (system, parameter set)."""

    def __init__(self, sysdef, xi, errortol, rc=0.0, mesh=None, P=0, precision="fp64", restart=0,
                 max_iter=0, gmres_tol=0.0, skin=0.4, preconditioner="diagonal"):
        from conductorrigid_amd.compute import ConductorRigid

        self.cr = ConductorRigid(system=sysdef, xi=xi, errortol=errortol, rc=rc, mesh=mesh, P=P,
                                 precision=precision, restart=restart, max_iter=max_iter, gmres_tol=gmres_tol,
                                 skin=skin, field=(0.0, 0.0, 1.0), preconditioner=preconditioner)
        self.torch = self.cr.torch
        self.lib = self.cr.lib
        self.ctx = self.cr.ctx
        self.s = sysdef
        self.n = sysdef.N + sysdef.N_isolated
        self.G = int(np.prod(self.cr.params["mesh"]))
        self.mesh = tuple(self.cr.params["mesh"])

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")

    def zeros(self, *shape):
        return self.torch.zeros(*shape, dtype=self.torch.float64, device="cuda")

    def st(self):
        return self.cr.stream()

    def sv(self):
        return self.cr.dsys.view()

    def call(self, name, *args):
        _lib.check(getattr(self.lib, name)(self.ctx, *args))
        self.torch.cuda.synchronize()

    def spread(self, q):
        g = self.zeros(self.G)
        sv = self.sv()
        self.call("cr_spread", C.byref(sv), self.dev(q).data_ptr(), g.data_ptr(), self.st())
        return g.cpu().numpy().reshape(self.mesh)

    def recip(self, grid):
        gi = self.dev(np.ravel(grid))
        go = self.zeros(self.G)
        self.call("cr_recip_grid", gi.data_ptr(), go.data_ptr(), self.st())
        return go.cpu().numpy().reshape(self.mesh)

    def fft(self, z, sign):
        zi = self.dev(np.stack([z.real, z.imag], axis=-1).ravel())
        zo = self.zeros(2 * self.G)
        self.call("cr_fft_c2c", zi.data_ptr(), zo.data_ptr(), int(sign), self.st())
        a = zo.cpu().numpy().reshape(self.mesh + (2,))
        return a[..., 0] + 1j * a[..., 1]

    def contract_after_recip(self, grid_in):
        """grid -> recip(grid) kept in the context -> contract."""
        gi = self.dev(np.ravel(grid_in))
        go = self.zeros(self.G)
        self.call("cr_recip_grid", gi.data_ptr(), go.data_ptr(), self.st())
        phi = self.zeros(self.s.N)
        sv = self.sv()
        self.call("cr_contract", C.byref(sv), phi.data_ptr(), self.st())
        return go.cpu().numpy().reshape(self.mesh), phi.cpu().numpy()

    def contract_force_after_recip(self, grid_in, q):
        gi = self.dev(np.ravel(grid_in))
        self.call("cr_recip_grid", gi.data_ptr(), 0, self.st())
        F = self.zeros(self.s.N, 4)
        sv = self.sv()
        self.call("cr_contract_force", C.byref(sv), self.dev(q).data_ptr(), F.data_ptr(), self.st())
        return F.cpu().numpy()

    def real_space(self, x, phi_recip):
        y = self.dev(np.concatenate([phi_recip, np.zeros(self.s.N_isolated)]))
        sv = self.sv()
        self.call("cr_real_space", C.byref(sv), self.dev(x).data_ptr(), y.data_ptr(), self.st())
        return y.cpu().numpy()[:self.s.N]

    def real_space_force(self, q, F0):
        F = self.dev(F0)
        sv = self.sv()
        self.call("cr_real_space_force", C.byref(sv), self.dev(q).data_ptr(), F.data_ptr(), self.st())
        return F.cpu().numpy()

    def matvec(self, x):
        y = self.zeros(self.n)
        sv = self.sv()
        self.call("cr_matvec", C.byref(sv), self.dev(x).data_ptr(), y.data_ptr(), self.st())
        return y.cpu().numpy()

    def precondition(self, x):
        y = self.zeros(self.n)
        sv = self.sv()
        self.call("cr_precondition", C.byref(sv), self.dev(x).data_ptr(), y.data_ptr(), self.st())
        return y.cpu().numpy()

    def objective(self):
        b = self.zeros(self.n)
        sv = self.sv()
        self.call("cr_objective", C.byref(sv), b.data_ptr(), self.st())
        return b.cpu().numpy()

    def solve(self, x0="zero"):
        stats = self.cr.computeForces(0, x0=x0)
        self.torch.cuda.synchronize()
        return self.cr.get_charges(), self.cr.get_potentials(), self.cr.get_forces(), stats
