"""ctypes wrapper of the CPU FP64 oracle (oracle/cr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcr_oracle.so")

dptr = C.POINTER(C.c_double)
uptr = C.POINTER(C.c_uint)


class Params(C.Structure):
    _fields_ = [("L", C.c_double * 3), ("xi", C.c_double), ("errortol", C.c_double),
                ("rc", C.c_double), ("kcut", C.c_double), ("N", C.c_int * 3), ("P", C.c_int),
                ("h", C.c_double * 3), ("eta", C.c_double * 3), ("N_table", C.c_int),
                ("drtable", C.c_double), ("selfterm", C.c_double), ("prefac", C.c_double),
                ("xiterm", C.c_double), ("quadW", C.c_double)]

    @property
    def G(self) -> int:
        return self.N[0] * self.N[1] * self.N[2]

    def as_dict(self) -> dict:
        return {k: (list(getattr(self, k)) if isinstance(getattr(self, k), C.Array) else getattr(self, k))
                for k, _ in self._fields_}


class System(C.Structure):
    _fields_ = [("N", C.c_int), ("pos", dptr), ("tag", uptr), ("body_of_tag", uptr),
                ("body_charge", dptr), ("N_rigid", C.c_int), ("N_isolated", C.c_int),
                ("Nmax", C.c_int), ("rigid_membership", uptr), ("rigid_members", uptr),
                ("rigid_size", uptr), ("rigid_com", dptr), ("nlist", uptr), ("head_list", uptr),
                ("n_neigh", uptr)]


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.cro_set_params.argtypes = [C.POINTER(Params), dptr, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_int), C.c_int]
        L.cro_smooth_list.argtypes = [C.POINTER(C.c_int), C.c_int]
        L.cro_smooth_list.restype = C.c_int
        L.cro_tables.argtypes = [C.POINTER(Params), dptr, dptr]
        L.cro_kscale.argtypes = [C.POINTER(Params), dptr]
        L.cro_phiq_exact.argtypes = [C.c_double, C.c_double]
        L.cro_phiq_exact.restype = C.c_double
        L.cro_gradphiq_exact.argtypes = [C.c_double, C.c_double]
        L.cro_gradphiq_exact.restype = C.c_double
        L.cro_table_lookup.argtypes = [C.POINTER(Params), dptr, C.c_double]
        L.cro_table_lookup.restype = C.c_double
        L.cro_spread.argtypes = [C.POINTER(Params), C.c_int, dptr, dptr, dptr]
        L.cro_fft3d.argtypes = [C.POINTER(C.c_int), dptr, dptr, C.c_int]
        L.cro_dft3d_brute.argtypes = [C.POINTER(C.c_int), dptr, dptr, dptr, dptr, C.c_int]
        L.cro_recip_grid.argtypes = [C.POINTER(Params), dptr, dptr]
        L.cro_contract.argtypes = [C.POINTER(Params), C.c_int, dptr, dptr, dptr]
        L.cro_contract_force.argtypes = [C.POINTER(Params), C.c_int, dptr, dptr, dptr, dptr]
        L.cro_real_space.argtypes = [C.POINTER(Params), C.POINTER(System), dptr, uptr, dptr, dptr, dptr]
        L.cro_real_space_force.argtypes = [C.POINTER(Params), C.POINTER(System), dptr, dptr, dptr, dptr]
        L.cro_sum_charges.argtypes = [C.POINTER(System), uptr, dptr, dptr]
        L.cro_objective.argtypes = [C.POINTER(System), dptr, dptr, dptr]
        L.cro_precondition.argtypes = [C.POINTER(System), uptr, dptr, dptr]
        L.cro_create.argtypes = [C.POINTER(Params), C.POINTER(System)]
        L.cro_create.restype = C.c_void_p
        L.cro_destroy.argtypes = [C.c_void_p]
        L.cro_matvec.argtypes = [C.c_void_p, dptr, dptr]
        L.cro_apply_precond.argtypes = [C.c_void_p, dptr, dptr]
        L.cro_gmres.argtypes = [C.c_void_p, dptr, dptr, C.c_int, C.c_int, C.c_double,
                                C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.cro_gmres.restype = C.c_int
        L.cro_compute_forces.argtypes = [C.c_void_p, dptr, dptr, dptr, dptr, C.c_int, C.c_int,
                                         C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.cro_compute_forces.restype = C.c_int
        L.cro_direct_matvec.argtypes = [C.POINTER(Params), C.POINTER(System), C.c_double, C.c_double,
                                        dptr, dptr]
        L.cro_direct_force.argtypes = [C.POINTER(Params), C.POINTER(System), C.c_double, C.c_double,
                                       dptr, dptr, dptr]
        L.cro_build_nlist.argtypes = [C.c_int, dptr, dptr, C.c_double, uptr, uptr, uptr, C.c_long]
        L.cro_build_nlist.restype = C.c_long
        L.cro_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(dptr)


def _u(a: np.ndarray):
    return a.ctypes.data_as(uptr)


def smooth_list() -> list[int]:
    buf = (C.c_int * 512)()
    n = lib().cro_smooth_list(buf, 512)
    return list(buf[:n])


def make_params(L, xi, errortol, rc=0.0, mesh=None, P=0) -> Params:
    p = Params()
    Lb = np.ascontiguousarray(L, dtype=np.float64)
    m = (C.c_int * 3)(*(mesh if mesh is not None else (0, 0, 0)))
    lib().cro_set_params(C.byref(p), _d(Lb), float(xi), float(errortol), float(rc), m, int(P))
    return p


def build_nlist(pos: np.ndarray, L, rlist: float):
    """Full neighbour list in HOOMD layout (head_list, n_neigh, nlist)."""
    N = pos.shape[0]
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    Lb = np.ascontiguousarray(L, dtype=np.float64)
    head = np.zeros(N, dtype=np.uint32)
    nn = np.zeros(N, dtype=np.uint32)
    cap = max(1, N * 64)
    while True:
        nl = np.zeros(cap, dtype=np.uint32)
        tot = lib().cro_build_nlist(N, _d(pos), _d(Lb), float(rlist), _u(head), _u(nn), _u(nl), cap)
        if tot <= cap:
            return head, nn, nl[:max(tot, 1)].copy()
        cap = int(tot)


class Oracle:
    """Oracle bound to one system and one parameter set (owns the C context)."""

    def __init__(self, sysdef, xi, errortol, rc=0.0, mesh=None, P=0, nlist=None, skin=0.4):
        self.sysdef = sysdef
        self.p = make_params(sysdef.L, xi, errortol, rc, mesh, P)
        s = sysdef
        self._keep = dict(
            pos=np.ascontiguousarray(s.pos, dtype=np.float64),
            tag=np.ascontiguousarray(s.tag, dtype=np.uint32),
            body_of_tag=np.ascontiguousarray(s.body_of_tag, dtype=np.uint32),
            body_charge=np.ascontiguousarray(s.body_charge, dtype=np.float64),
            rigid_membership=np.ascontiguousarray(s.rigid_membership, dtype=np.uint32),
            rigid_members=np.ascontiguousarray(s.rigid_members if s.rigid_members.size else np.zeros(1, np.uint32), dtype=np.uint32),
            rigid_size=np.ascontiguousarray(s.rigid_size if s.rigid_size.size else np.zeros(1, np.uint32), dtype=np.uint32),
            rigid_com=np.ascontiguousarray(s.rigid_com if s.rigid_com.size else np.zeros((1, 4)), dtype=np.float64),
        )
        if nlist is None:
            nlist = build_nlist(s.pos, s.L, self.p.rc + skin)
        head, nn, nl = nlist
        self._keep.update(head=np.ascontiguousarray(head, dtype=np.uint32),
                          nn=np.ascontiguousarray(nn, dtype=np.uint32),
                          nl=np.ascontiguousarray(nl, dtype=np.uint32))
        k = self._keep
        self.sys = System(s.N, _d(k["pos"]), _u(k["tag"]), _u(k["body_of_tag"]), _d(k["body_charge"]),
                          s.N_rigid, s.N_isolated, s.Nmax, _u(k["rigid_membership"]), _u(k["rigid_members"]),
                          _u(k["rigid_size"]), _d(k["rigid_com"]), _u(k["nl"]), _u(k["head"]), _u(k["nn"]))
        self.ctx = lib().cro_create(C.byref(self.p), C.byref(self.sys))
        self.n = s.N + s.N_isolated

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                lib().cro_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    @property
    def nlist(self):
        return self._keep["head"], self._keep["nn"], self._keep["nl"]

    # --- tables and scale ---
    def tables(self):
        nt = self.p.N_table + 1
        a = np.zeros(2 * nt)
        b = np.zeros(2 * nt)
        lib().cro_tables(C.byref(self.p), _d(a), _d(b))
        return a.reshape(nt, 2), b.reshape(nt, 2)

    def kscale(self):
        S = np.zeros(self.p.G)
        lib().cro_kscale(C.byref(self.p), _d(S))
        return S.reshape(tuple(self.p.N))

    # --- pieces ---
    def spread(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        g = np.zeros(self.p.G)
        lib().cro_spread(C.byref(self.p), self.sysdef.N, _d(self._keep["pos"]), _d(q), _d(g))
        return g.reshape(tuple(self.p.N))

    def recip_grid(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel().copy()
        S = np.ascontiguousarray(self.kscale().ravel())
        lib().cro_recip_grid(C.byref(self.p), _d(S), _d(g))
        return g.reshape(tuple(self.p.N))

    def contract(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
        out = np.zeros(self.sysdef.N)
        lib().cro_contract(C.byref(self.p), self.sysdef.N, _d(self._keep["pos"]), _d(g), _d(out))
        return out

    def contract_force(self, q, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros((self.sysdef.N, 4))
        lib().cro_contract_force(C.byref(self.p), self.sysdef.N, _d(self._keep["pos"]), _d(q), _d(g), _d(out))
        return out

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().cro_matvec(self.ctx, _d(x), _d(y))
        return y

    def precondition(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().cro_apply_precond(self.ctx, _d(x), _d(y))
        return y

    def objective(self, E0):
        b = np.zeros(self.n)
        E = np.ascontiguousarray(E0, dtype=np.float64)
        Lb = np.ascontiguousarray(self.sysdef.L, dtype=np.float64)
        lib().cro_objective(C.byref(self.sys), _d(Lb), _d(E), _d(b))
        return b

    def gmres(self, b, x0=None, restart=10, maxiter=100, tol=1e-6):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it = C.c_int(0)
        res = C.c_double(0)
        conv = lib().cro_gmres(self.ctx, _d(x), _d(b), int(restart), int(maxiter), float(tol),
                               C.byref(it), C.byref(res))
        return x, dict(converged=bool(conv), iters=it.value, resid=res.value)

    def compute_forces(self, E0, q0=None, Phi0=None, restart=10, maxiter=100, tol=1e-6):
        s = self.sysdef
        q = np.zeros(s.N) if q0 is None else np.array(q0, dtype=np.float64)
        Phi = np.zeros(s.N_isolated) if Phi0 is None else np.array(Phi0, dtype=np.float64)
        F = np.zeros((s.N, 4))
        E = np.ascontiguousarray(E0, dtype=np.float64)
        it = C.c_int(0)
        res = C.c_double(0)
        conv = lib().cro_compute_forces(self.ctx, _d(E), _d(q), _d(Phi), _d(F), int(restart), int(maxiter),
                                        float(tol), C.byref(it), C.byref(res))
        return q, Phi, F, dict(converged=bool(conv), iters=it.value, resid=res.value)

    def direct_matvec(self, q, rc_real, kmax):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros(self.sysdef.N)
        lib().cro_direct_matvec(C.byref(self.p), C.byref(self.sys), float(rc_real), float(kmax), _d(q), _d(out))
        return out

    def direct_force(self, q, E0, rc_real, kmax):
        q = np.ascontiguousarray(q, dtype=np.float64)
        E = np.ascontiguousarray(E0, dtype=np.float64)
        out = np.zeros((self.sysdef.N, 4))
        lib().cro_direct_force(C.byref(self.p), C.byref(self.sys), float(rc_real), float(kmax), _d(q), _d(E), _d(out))
        return out

    def direct_solve(self, E0, rc_real, kmax, tol=1e-12, restart=60, maxiter=2000):
        """Direct-Ewald force evaluation: the saddle-point system of the reference
        (cppmodule/ConductorRigid.cu:862-1020) with M from direct_matvec (closed-form
        real space over images within rc_real, explicit k-sum to kmax) solved by
        GMRES to `tol` with the reference's diagonal preconditioner, then the direct
        forces.  Returns (q, Phi, F, iterations)."""
        import scipy.sparse.linalg as spla

        N, B = self.sysdef.N, self.sysdef.N_isolated
        iso = self.sysdef.iso_membership().astype(np.int64)

        def A(x):
            q, Phi = x[:N], x[N:]
            top = -self.direct_matvec(q, rc_real, kmax) + Phi[iso]
            return np.concatenate([top, np.bincount(iso, weights=q, minlength=B)])

        n = N + B
        Aop = spla.LinearOperator((n, n), matvec=A, dtype=np.float64)
        Mop = spla.LinearOperator((n, n), matvec=self.precondition, dtype=np.float64)
        b = self.objective(E0)
        count = [0]
        x, info = spla.gmres(Aop, b, M=Mop, rtol=tol, atol=0.0, restart=restart, maxiter=maxiter,
                             callback=lambda _r: count.__setitem__(0, count[0] + 1), callback_type="pr_norm")
        if info != 0:
            raise RuntimeError(f"direct-Ewald GMRES did not converge (info={info})")
        q, Phi = x[:N], x[N:]
        return q, Phi, self.direct_force(q, E0, rc_real, kmax), count[0]


def fft3d(re, im, sign):
    n = (C.c_int * 3)(*re.shape)
    r = np.ascontiguousarray(re, dtype=np.float64).copy()
    i = np.ascontiguousarray(im, dtype=np.float64).copy()
    lib().cro_fft3d(n, _d(r), _d(i), int(sign))
    return r, i


def dft3d_brute(re, im, sign):
    n = (C.c_int * 3)(*re.shape)
    r = np.ascontiguousarray(re, dtype=np.float64)
    i = np.ascontiguousarray(im, dtype=np.float64)
    o_r = np.zeros_like(r)
    o_i = np.zeros_like(i)
    lib().cro_dft3d_brute(n, _d(r), _d(i), _d(o_r), _d(o_i), int(sign))
    return o_r, o_i


def phiq_exact(r, xi):
    return lib().cro_phiq_exact(float(r), float(xi))


def gradphiq_exact(r, xi):
    return lib().cro_gradphiq_exact(float(r), float(xi))
