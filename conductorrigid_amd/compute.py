"""Python plugin API of the ConductorRigid force compute, mirroring the reference's
``compute.py`` (pymodule/compute.py:16-64) on top of the MI355X HIP library.

Reference API kept: ``ConductorRigid(group, body, charge, field=[0,0,0], xi=0.5,
errortol=1e-3, fileprefix="", period=0)`` with ``update_field``,
``update_parameters`` and ``update_coeffs``.  HOOMD is not present, so the system
(positions, box, rigid data) is handed over as a :class:`SystemDefinition` through the
``system`` keyword, and ``compute(timestep)`` plays HOOMD's
``ForceCompute::compute`` (cppmodule/ConductorRigid.cc:403-557).  Keyword-only
extensions expose what the reference derives from ``errortol`` alone (mesh, P, rc,
GMRES restart / iteration limit / tolerance), the precision mode and the GMRES
initial guess.

Device memory is held in torch tensors (plumbing only); all compute runs in the HIP
library through its C ABI.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os

import numpy as np

from . import _lib
from .system import NOT_MEMBER, SystemDefinition

_PRECISIONS = {"fp64": _lib.CR_PRECISION_FP64, "fp32": _lib.CR_PRECISION_FP32_BULK}
_PRECONDITIONERS = {"diagonal": _lib.CR_PRECOND_DIAGONAL, "block": _lib.CR_PRECOND_NEARFIELD_BLOCK}
# x0: "previous" = the reference's warm start from the last solution; "zero"; "charges" =
# from the charges currently in ``self.charges`` (e.g. a learned prediction), projected
# onto the body charges, with fitted body potentials
_X0_MODES = {"previous": _lib.CR_X0_PREVIOUS, "zero": _lib.CR_X0_ZERO, "charges": _lib.CR_X0_CHARGES}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("ConductorRigid requires a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class DeviceSystem:
    """Device-resident copy of a :class:`SystemDefinition` plus its neighbour list,
    exposed as the C ABI's ``cr_system`` view."""

    def __init__(self, sysdef: SystemDefinition, ctx, rlist: float, device: int = 0, stream=None):
        torch = _torch()
        self.torch = torch
        self.sysdef = sysdef
        self.dev = torch.device("cuda", device)
        d = self.dev

        def u32(a):
            a = np.ascontiguousarray(a, dtype=np.uint32)
            if a.size == 0:
                a = np.zeros(1, dtype=np.uint32)
            return torch.from_numpy(a.view(np.int32)).to(d)

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size == 0:
                a = np.zeros(1, dtype=np.float64)
            return torch.from_numpy(a).to(d)

        s = sysdef
        self.pos = f64(s.pos)
        self.tag = u32(s.tag)
        self.body_of_tag = u32(s.body_of_tag)
        self.body_charge = f64(s.body_charge)
        self.rigid_membership = u32(s.rigid_membership)
        self.rigid_members = u32(s.rigid_members)
        self.rigid_size = u32(s.rigid_size)
        self.rigid_com = f64(s.rigid_com)
        self.rlist = float(rlist)
        self.ctx = ctx
        self.head_list = torch.zeros(max(1, s.N), dtype=torch.int32, device=d)
        self.n_neigh = torch.zeros(max(1, s.N), dtype=torch.int32, device=d)
        self.nlist = torch.zeros(max(1, s.N * 48), dtype=torch.int32, device=d)
        self.build_nlist(stream)

    def build_nlist(self, stream=None) -> int:
        """Neighbour list (full, r < rc + skin) built on the GPU by the library."""
        lib = _lib.load()
        torch = self.torch
        s = self.sysdef
        L = (C.c_double * 3)(*[float(x) for x in s.L])
        total = C.c_long(0)
        st = stream if stream is not None else torch.cuda.current_stream(self.dev).cuda_stream
        while True:
            _lib.check(lib.cr_build_nlist(self.ctx, s.N, self.pos.data_ptr(), L, self.rlist,
                                          self.head_list.data_ptr(), self.n_neigh.data_ptr(),
                                          self.nlist.data_ptr(), self.nlist.numel(), C.byref(total), st))
            if total.value <= self.nlist.numel():
                break
            self.nlist = torch.zeros(int(total.value * 1.2) + 1, dtype=torch.int32, device=self.dev)
        self.n_pairs = int(total.value)
        return self.n_pairs

    def view(self) -> _lib.CrSystem:
        s = self.sysdef
        return _lib.CrSystem(s.N, self.pos.data_ptr(), self.tag.data_ptr(), self.body_of_tag.data_ptr(),
                             self.body_charge.data_ptr(), s.N_rigid, s.N_isolated, max(1, s.Nmax),
                             self.rigid_membership.data_ptr(), self.rigid_members.data_ptr(),
                             self.rigid_size.data_ptr(), self.rigid_com.data_ptr(), self.nlist.data_ptr(),
                             self.head_list.data_ptr(), self.n_neigh.data_ptr(),
                             (C.c_double * 3)(*[float(x) for x in s.L]))

    def nlist_host(self):
        return (self.head_list.cpu().numpy().view(np.uint32), self.n_neigh.cpu().numpy().view(np.uint32),
                self.nlist[:max(1, self.n_pairs)].cpu().numpy().view(np.uint32))


class ConductorRigid:
    """Computes the forces on charged rigid-body conductors in an external field
    (reference class ``ConductorRigid``, pymodule/compute.py:16-64)."""

    def __init__(self, group=None, body=None, charge=None, field=(0.0, 0.0, 0.0), xi=0.5, errortol=1e-3,
                 fileprefix="", period=0, *, system: SystemDefinition, mesh=None, P=0, rc=0.0, restart=0,
                 max_iter=0, gmres_tol=0.0, precision="fp64", x0="previous", skin=0.4, device=0, t0=0,
                 preconditioner="diagonal"):
        if group not in (None, "all"):
            raise ValueError("only the group of all particles is supported")
        torch = _torch()
        self.torch = torch
        self.device = device
        self.lib = _lib.load()
        if body is not None:
            system.body_of_tag = np.asarray(body, dtype=np.uint32)
        if charge is not None:
            system.body_charge = np.asarray(charge, dtype=np.float64)
        system.validate()
        self.system = system
        # the reference's compute.py sets the neighbour-list cutoff from (xi, errortol)
        # (pymodule/compute.py:27); an explicit rc override takes precedence here
        self.rcut = float(rc) if rc > 0 else math.sqrt(-math.log(errortol)) / xi
        self.field = [float(f) for f in field]
        self.xi = float(xi)
        self.errortol = float(errortol)
        self.fileprefix = fileprefix
        self.period = int(period)
        self.t0 = int(t0)
        self.enabled = True
        self._opts = dict(mesh=mesh, P=P, rc=rc, restart=restart, max_iter=max_iter, gmres_tol=gmres_tol,
                          precision=precision, preconditioner=preconditioner)
        self.x0_mode = _X0_MODES[x0]
        ctx = C.c_void_p()
        _lib.check(self.lib.cr_create(C.byref(ctx), int(device)))
        self.ctx = ctx
        self.skin = float(skin)
        self.SetParams()
        self.dsys = DeviceSystem(system, self.ctx, self.rcut + self.skin, device)
        dev = self.dsys.dev
        N = system.N
        self.charges = torch.zeros(max(1, N), dtype=torch.float64, device=dev)        # HOOMD particle charges
        self.potentials = torch.zeros(max(1, system.N_isolated), dtype=torch.float64, device=dev)
        self.forces = torch.zeros((max(1, N), 4), dtype=torch.float64, device=dev)
        self.last_stats = None
        self.UpdateField(self.field)

    # --- reference C++ method names (cppmodule/ConductorRigid.cc:624-633) ---
    def SetParams(self) -> None:
        o = self._opts
        mesh = o["mesh"] or (0, 0, 0)
        p = _lib.CrParams((C.c_double * 3)(*[float(x) for x in self.system.L]), self.xi, self.errortol,
                          float(o["rc"]), (C.c_int * 3)(*[int(m) for m in mesh]), int(o["P"]),
                          _PRECISIONS[o["precision"]], int(o["restart"]), int(o["max_iter"]), float(o["gmres_tol"]),
                          _PRECONDITIONERS[o["preconditioner"]])
        _lib.check(self.lib.cr_set_params(self.ctx, C.byref(p), self.system.N, self.system.N_isolated))
        r = _lib.CrResolved()
        _lib.check(self.lib.cr_get_resolved(self.ctx, C.byref(r)))
        self.params = r.as_dict()
        self._box = tuple(float(x) for x in self.system.L)

    def UpdateField(self, field) -> None:
        self.field = [float(f) for f in field]
        _lib.check(self.lib.cr_update_field(self.ctx, (C.c_double * 3)(*self.field)))

    def UpdateParams(self, field, fileprefix, period, t0=0) -> None:
        self.UpdateField(field)
        self.fileprefix = fileprefix
        self.period = int(period)
        self.t0 = int(t0)

    # --- reference Python API (pymodule/compute.py:54-64) ---
    def update_field(self, field) -> None:
        self.UpdateField(field)

    def update_parameters(self, field, fileprefix, period, timestep=0) -> None:
        self.UpdateParams(field, fileprefix, period, timestep)
        self.SetParams()

    def update_coeffs(self) -> None:
        pass

    # --- force evaluation (ConductorRigid::computeForces, cppmodule/ConductorRigid.cc:403-557) ---
    def stream(self):
        return self.torch.cuda.current_stream(self.dsys.dev).cuda_stream

    def computeForces(self, timestep: int = 0, x0: str | None = None, stats: bool = True) -> dict | None:
        mode = self.x0_mode if x0 is None else _X0_MODES[x0]
        if tuple(float(x) for x in self.system.L) != self._box:
            # the box changed (e.g. NPT): the neighbour list is rebuilt here (HOOMD's side);
            # the library takes up the new box itself from the view's box lengths
            # (mesh spacing, Gaussian widths, k-scale: cr_set_box)
            self._box = tuple(float(x) for x in self.system.L)
            self.dsys.build_nlist()
        st = _lib.CrSolveStats()
        sv = self.dsys.view()
        code = self.lib.cr_compute_force(self.ctx, C.byref(sv), self.charges.data_ptr(), self.potentials.data_ptr(),
                                         self.forces.data_ptr(), mode, C.byref(st) if stats else None, self.stream())
        _lib.check(code)
        self.last_stats = st.as_dict() if stats else None
        if self.period > 0 and (int(timestep) - self.t0) % self.period == 0:
            self.OutputData(int(timestep))
        return self.last_stats

    compute = computeForces

    def autotune(self) -> dict:
        """Time the binned spread + contraction for every admissible tile shape on the
        current configuration and keep the fastest (C ABI ``cr_autotune``)."""
        t = _lib.CrTuning()
        sv = self.dsys.view()
        _lib.check(self.lib.cr_autotune(self.ctx, C.byref(sv), C.byref(t), self.stream()))
        return {"tile": list(t.tile), "lds_bytes": t.lds_bytes, "spread_contract_ms": t.spread_contract_ms,
                "candidates": t.candidates}

    def set_residual_check(self, on: bool) -> None:
        """Recompute the true residual (one extra FP64 operator application) whenever
        stats are returned (default on; cr_set_residual_check).  Off, ``true_resid`` and
        ``true_relres`` read -1."""
        _lib.check(self.lib.cr_set_residual_check(self.ctx, int(bool(on))))

    def set_tiles(self, tile, lds_bytes: int = 64 * 1024) -> None:
        _lib.check(self.lib.cr_set_tiles(self.ctx, (C.c_int * 3)(*[int(v) for v in tile]), int(lds_bytes)))

    def OutputData(self, timestep: int) -> str:
        """Text dump in tag order (cppmodule/ConductorRigid.cc:560-622); returns the file name."""
        s = self.system
        fn = output_filename(self.fileprefix, timestep)
        write_output(fn, s.tag, s.pos, self.charges.cpu().numpy()[:s.N], self.forces.cpu().numpy()[:s.N])
        return fn

    # --- checkpoint ---
    def save_state(self, path: str) -> None:
        """Write the solver state that carries across steps (bead charges and body
        potentials, the previous step's x0), the external field and the parameters to an
        .npz file.  The reference keeps the charges in HOOMD's particle data but loses the
        body potentials (cppmodule/ConductorRigid.h:67) across a restart; both are kept."""
        meta = {"N": int(self.system.N), "N_isolated": int(self.system.N_isolated), "xi": float(self.xi),
                "errortol": float(self.errortol), "params": {k: v for k, v in self.params.items()
                                                            if isinstance(v, (int, float, list))}}
        np.savez(path, charges=self.get_charges(), potentials=self.get_potentials(),
                 field=np.asarray(self.field, dtype=np.float64), L=np.asarray(self.system.L, dtype=np.float64),
                 meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))

    def load_state(self, path: str) -> None:
        """Restore a state written by save_state (same system size and method parameters);
        the next computeForces(x0="previous") starts from it."""
        with np.load(path, allow_pickle=False) as f:
            meta = json.loads(bytes(f["meta"]).decode())
            q, phi, field = f["charges"], f["potentials"], f["field"]
        if meta["N"] != self.system.N or meta["N_isolated"] != self.system.N_isolated:
            raise ValueError("checkpoint is for a system of another size")
        for k in ("mesh", "P", "rc"):
            if meta["params"].get(k) != self.params.get(k):
                raise ValueError(f"checkpoint parameter {k} differs: {meta['params'].get(k)} vs {self.params.get(k)}")
        t = self.torch
        self.charges[:len(q)].copy_(t.from_numpy(np.ascontiguousarray(q, dtype=np.float64)))
        if len(phi):
            self.potentials[:len(phi)].copy_(t.from_numpy(np.ascontiguousarray(phi, dtype=np.float64)))
        self.update_field(tuple(float(v) for v in field))

    # --- numpy views ---
    def get_charges(self) -> np.ndarray:
        return self.charges.cpu().numpy()[:self.system.N].copy()

    def get_potentials(self) -> np.ndarray:
        return self.potentials.cpu().numpy()[:self.system.N_isolated].copy()

    def get_forces(self) -> np.ndarray:
        return self.forces.cpu().numpy()[:self.system.N].copy()

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                self.lib.cr_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass


def update_positions(cr: ConductorRigid, pos: np.ndarray, rigid_com: np.ndarray | None = None) -> None:
    """Replace bead positions (and centres of mass) and rebuild the neighbour list."""
    s = cr.system
    s.pos = np.ascontiguousarray(pos, dtype=np.float64)
    cr.dsys.pos.copy_(cr.torch.from_numpy(s.pos))
    if rigid_com is not None:
        s.rigid_com = np.ascontiguousarray(rigid_com, dtype=np.float64)
        cr.dsys.rigid_com.copy_(cr.torch.from_numpy(s.rigid_com))
    cr.dsys.build_nlist()




def output_filename(prefix: str, timestep: int) -> str:
    """``prefix.%010d.txt`` (cppmodule/ConductorRigid.cc:563-567)."""
    return f"{prefix}.{int(timestep):010d}.txt"


def write_output(filename: str, tag, pos, q, F) -> None:
    """OutputData's text format (cppmodule/ConductorRigid.cc:560-622): sections "Bead
    Position", "Bead Charge", "Bead Force", each in particle-tag order (global index
    rtag[t] for tag t), numbers as an ostream with setprecision(10) writes them (printf
    %.10g), vector components separated and followed by two spaces.  pos and F are
    indexed by global index, tag[i] is the tag of global index i."""
    tag = np.asarray(tag, dtype=np.int64)
    n = tag.size
    rtag = np.empty(n, dtype=np.int64)
    rtag[tag] = np.arange(n)
    pos = np.asarray(pos, dtype=np.float64)[rtag]
    q = np.asarray(q, dtype=np.float64)[rtag]
    F = np.asarray(F, dtype=np.float64)[rtag]
    with open(filename, "w") as f:
        f.write("Bead Position\n")
        for p in pos:
            f.write(f"{p[0]:.10g}  {p[1]:.10g}  {p[2]:.10g}  \n")
        f.write("Bead Charge\n")
        for v in q:
            f.write(f"{v:.10g}\n")
        f.write("Bead Force\n")
        for v in F:
            f.write(f"{v[0]:.10g}  {v[1]:.10g}  {v[2]:.10g}  \n")


class BatchConductorRigid:
    """Independent periodic boxes solved together (C ABI ``cr_batch_compute_force``).

    Each box is a :class:`ConductorRigid` with its own system, parameters and tiling; a
    call gives every box the charges, potentials and forces its own ``computeForces``
    would give, while all boxes share each spread / contraction launch and each
    lockstep Arnoldi step (replayed from a captured HIP graph when ``use_graph``).  No
    reference counterpart: the reference handles one box per plugin instance."""

    def __init__(self, boxes: list, device: int = 0, use_graph: bool = True):
        if not boxes:
            raise ValueError("a batch needs at least one box")
        self.boxes = list(boxes)
        self.lib = _lib.load()
        self.use_graph = bool(use_graph)
        self.torch = self.boxes[0].torch
        h = C.c_void_p()
        _lib.check(self.lib.cr_batch_create(C.byref(h), int(device)))
        self.handle = h
        self.last_stats = None

    def autotune(self) -> list[dict]:
        return [b.autotune() for b in self.boxes]

    def computeForces(self, x0: str = "previous", stats: bool = True) -> list[dict] | None:
        n = len(self.boxes)
        ctxs = (C.c_void_p * n)(*[b.ctx.value for b in self.boxes])
        views = [b.dsys.view() for b in self.boxes]
        systems = (_lib.CrSystem * n)(*views)
        q = (C.c_void_p * n)(*[b.charges.data_ptr() for b in self.boxes])
        phi = (C.c_void_p * n)(*[b.potentials.data_ptr() for b in self.boxes])
        force = (C.c_void_p * n)(*[b.forces.data_ptr() for b in self.boxes])
        st = (_lib.CrSolveStats * n)()
        code = self.lib.cr_batch_compute_force(self.handle, n, ctxs, systems, q, phi, force, _X0_MODES[x0],
                                               int(self.use_graph), st if stats else None, self.boxes[0].stream())
        _lib.check(code)
        self.last_stats = [st[i].as_dict() for i in range(n)] if stats else None
        return self.last_stats

    compute = computeForces

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.lib.cr_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def batch_boxes(items, precision: str = "fp32", preconditioner: str = "block", restart: int = 30,
                max_iter: int = 300, gmres_tol: float = 0.0, field=(0.0, 0.0, 1.0), device: int = 0) -> list:
    """One :class:`ConductorRigid` per (system, params) item of ``system.make_batch``."""
    out = []
    for sysdef, p in items:
        out.append(ConductorRigid(system=sysdef, xi=p["xi"], errortol=p["errortol"], rc=p["rc"], mesh=p["mesh"],
                                  P=p["P"], precision=precision, restart=restart, max_iter=max_iter,
                                  gmres_tol=gmres_tol, field=field, device=device, preconditioner=preconditioner))
    return out


__all__ = ["ConductorRigid", "BatchConductorRigid", "DeviceSystem", "batch_boxes", "update_positions", "NOT_MEMBER",
           "output_filename", "write_output"]
