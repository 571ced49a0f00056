"""Learned warm start for the force compute's GMRES solve.

A per-bead MLP (C ABI ``cr_mlp_*``, kernels in csrc/k_mlp.hip) predicts each bead's
induced charge relative to its body's uniform share Q_b / n_b (in units of 4 pi) from
local geometry.  The
prediction enters the solve as x0 (``ConductorRigid.computeForces(x0="charges")``): the
library shifts it per body onto the body's net charge and fits the body potentials, then
GMRES runs to the same tolerance as from a zero guess.  No reference counterpart: the
reference warm-starts from the previous step's solution only
(cppmodule/ConductorRigid.cu:1125-1127).

Training data are (features, solved charges) pairs the solver itself produces along a
Brownian-dynamics trajectory (:func:`run_trajectory`), and training runs on the GPU with
the library's own forward/backward kernels and Adam (:func:`fit`).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


class WarmStartMLP:
    """The network's parameters and training state on one GPU."""

    IN = _lib.CR_MLP_IN
    PARAMS = _lib.CR_MLP_PARAMS

    def __init__(self, device: int = 0, seed: int = 0, params: np.ndarray | None = None):
        self.lib = _lib.load()
        self.torch = _torch()
        self.device = device
        h = C.c_void_p()
        _lib.check(self.lib.cr_mlp_create(C.byref(h), int(device), C.c_ulonglong(seed)))
        self.h = h
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.cr_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # --- parameters ---
    def get_params(self) -> np.ndarray:
        theta = np.zeros(self.PARAMS, dtype=np.float32)
        _lib.check(self.lib.cr_mlp_get_params(self.h, theta.ctypes.data))
        return theta

    def set_params(self, theta) -> None:
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.shape != (self.PARAMS,):
            raise ValueError(f"expected {self.PARAMS} parameters")
        _lib.check(self.lib.cr_mlp_set_params(self.h, theta.ctypes.data))

    def save(self, path: str) -> None:
        np.savez_compressed(path, theta=self.get_params())

    @classmethod
    def load(cls, path: str, device: int = 0) -> "WarmStartMLP":
        with np.load(path, allow_pickle=False) as f:
            return cls(device=device, params=f["theta"])

    # --- data ---
    def features(self, cr) -> "object":
        """bf16 feature rows [N, IN] (int16 storage) of ``cr``'s current configuration."""
        t = self.torch
        X = t.empty((max(1, cr.system.N), self.IN), dtype=t.int16, device=cr.dsys.dev)
        sv = cr.dsys.view()
        _lib.check(self.lib.cr_mlp_features(cr.ctx, C.byref(sv), X.data_ptr(), cr.stream()))
        return X

    def targets(self, cr) -> "object":
        """Training targets (q_i - Q_b / n_b) / 4 pi from ``cr``'s current (solved) charges."""
        t = self.torch
        y = t.empty(max(1, cr.system.N), dtype=t.float32, device=cr.dsys.dev)
        sv = cr.dsys.view()
        _lib.check(self.lib.cr_mlp_targets(cr.ctx, C.byref(sv), cr.charges.data_ptr(), y.data_ptr(), cr.stream()))
        return y

    # --- network ---
    def predict(self, X, stream=None):
        t = self.torch
        M = int(X.shape[0])
        y = t.empty(max(1, M), dtype=t.float32, device=X.device)
        st = stream if stream is not None else t.cuda.current_stream(X.device).cuda_stream
        _lib.check(self.lib.cr_mlp_predict(self.h, M, X.data_ptr(), y.data_ptr(), st))
        return y[:M]

    def gradient(self, X, target):
        t = self.torch
        M = int(X.shape[0])
        g = t.empty(self.PARAMS, dtype=t.float32, device=X.device)
        y = t.empty(M, dtype=t.float32, device=X.device)
        st = t.cuda.current_stream(X.device).cuda_stream
        _lib.check(self.lib.cr_mlp_gradient(self.h, M, X.data_ptr(), target.data_ptr(), y.data_ptr(), g.data_ptr(),
                                            st))
        return g, y

    def train_step(self, X, target, lr: float):
        t = self.torch
        st = t.cuda.current_stream(X.device).cuda_stream
        _lib.check(self.lib.cr_mlp_train_step(self.h, int(X.shape[0]), X.data_ptr(), target.data_ptr(),
                                              C.c_float(lr), None, st))

    # --- warm start ---
    def warm_start(self, cr) -> None:
        """Write the predicted charges into ``cr.charges``; follow with
        ``cr.computeForces(x0="charges")``."""
        X = self.features(cr)
        y = self.predict(X, cr.stream())
        sv = cr.dsys.view()
        _lib.check(self.lib.cr_mlp_charges(cr.ctx, C.byref(sv), y.data_ptr(), cr.charges.data_ptr(), cr.stream()))

    def warm_start_from_previous(self, cr, E_prev) -> None:
        """``cr.charges`` holds the previous step's solution for field ``E_prev``; add the
        network's predicted response to the field change at the current geometry,
        4 pi (y(E0) - y(E_prev)) = 4 pi (E0 - E_prev) . a, and leave it there for
        ``cr.computeForces(x0="charges")``."""
        X = self.features(cr)
        y_now = self.predict(X, cr.stream())
        set_gate(X, E_prev)
        y_prev = self.predict(X, cr.stream())
        n = cr.system.N
        cr.charges[:n].add_((y_now - y_prev).double() * (4.0 * np.pi))


def set_gate(X, E0) -> None:
    """Overwrite the field gate columns of feature rows X (int16 storage of bf16) with
    E0 as bf16 pairs hi + lo, as the feature kernel writes them."""
    e = np.asarray(E0, dtype=np.float32).reshape(3)
    u = e.view(np.uint32).astype(np.uint64)
    hi = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    lo_f = (e - (hi.astype(np.uint32) << 16).view(np.float32)).astype(np.float32)
    v = lo_f.view(np.uint32).astype(np.uint64)
    lo = ((v + 0x7FFF + ((v >> 16) & 1)) >> 16).astype(np.uint16)
    g = _lib.CR_MLP_TRUNK
    for d in range(3):
        X[:, g + d] = int(hi.view(np.int16)[d])
        X[:, g + 3 + d] = int(lo.view(np.int16)[d])


def bd_step(cr, dt: float, kT: float = 1.0, gamma_t: float = 1.0, gamma_r: float = 20.0, k_rep: float = 100.0,
            generator=None) -> None:
    """One Brownian rigid-body step of ``cr``'s system from its current forces (C ABI
    ``cr_bd_step``), then the neighbour list is rebuilt.  gamma_r / gamma_t = 20 is the
    Stokes ratio 4 a^2 / 3 for a body of radius a ~ 4; k_rep is the stiffness of the
    contact repulsion between bodies.  Host copies of the positions are refreshed lazily
    by :func:`sync_host`."""
    t = _torch()
    s = cr.system
    dev = cr.dsys.dev
    nr = t.randn(max(1, 6 * s.N_rigid), dtype=t.float64, device=dev, generator=generator)
    free = bool(np.any(s.rigid_membership == 0xFFFFFFFF))
    nb = t.randn(max(1, 3 * s.N), dtype=t.float64, device=dev, generator=generator) if free else None
    sv = cr.dsys.view()
    _lib.check(cr.lib.cr_bd_step(cr.ctx, C.byref(sv), cr.dsys.pos.data_ptr(), cr.dsys.rigid_com.data_ptr(),
                                 cr.forces.data_ptr(), nr.data_ptr(), nb.data_ptr() if nb is not None else None,
                                 float(dt), float(kT), float(gamma_t), float(gamma_r), float(k_rep), cr.stream()))
    cr.dsys.build_nlist()


def sync_host(cr) -> None:
    """Copy the device positions and centres of mass back into ``cr.system``."""
    s = cr.system
    s.pos = cr.dsys.pos.cpu().numpy().copy()
    if s.N_rigid:
        s.rigid_com = cr.dsys.rigid_com.cpu().numpy().copy()


def run_trajectory(cr, steps: int, dt: float, field_range=(0.5, 2.0), collect=None, mlp: WarmStartMLP | None = None,
                   seed: int = 0, x0: str = "zero", log=None):
    """Brownian-dynamics trajectory with one force evaluation per step.  Before each
    step |E0| is drawn uniformly from ``field_range`` (direction z).  For steps in
    ``collect`` the (features, targets) pair of the solved state is kept (requires
    ``mlp`` for the feature/target transforms).  Returns per-step records
    (iterations, solve ms) and the collected pairs."""
    t = _torch()
    rng = np.random.default_rng(seed)
    gen = t.Generator(device=cr.dsys.dev)
    gen.manual_seed(seed)
    collect = set(collect or ())
    recs, pairs = [], []
    for k in range(steps):
        E = float(rng.uniform(*field_range))
        cr.update_field((0.0, 0.0, E))
        t.cuda.synchronize()
        t0 = time.perf_counter()
        st = cr.computeForces(k, x0=x0)
        t.cuda.synchronize()
        recs.append(dict(step=k, E0=E, iterations=st["iterations"], ms=1e3 * (time.perf_counter() - t0),
                         converged=st["converged"]))
        if k in collect:
            pairs.append((mlp.features(cr), mlp.targets(cr)))
        if log and k % 100 == 0:
            log(f"step {k}: |E0|={E:.2f} iters={st['iterations']}")
        bd_step(cr, dt, generator=gen)
    sync_host(cr)
    return recs, pairs


def fit(mlp: WarmStartMLP, pairs, epochs: int = 4, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
        lr_final: float | None = None, log=None):
    """Adam over all collected (features, target) rows, shuffled per epoch; the learning
    rate decays geometrically to ``lr_final`` (default lr / 10).  Returns the mean
    training loss per epoch (computed from the forward passes of the steps)."""
    t = _torch()
    X = t.cat([p[0] for p in pairs])
    T = t.cat([p[1] for p in pairs])
    M = int(X.shape[0])
    g = t.Generator(device=X.device)
    g.manual_seed(seed)
    steps_per_epoch = max(1, M // batch)
    total = epochs * steps_per_epoch
    lr_final = lr / 10 if lr_final is None else lr_final
    decay = (lr_final / lr) ** (1.0 / max(1, total - 1))
    hist = []
    step = 0
    for e in range(epochs):
        perm = t.randperm(M, device=X.device, generator=g)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            mlp.train_step(X[idx].contiguous(), T[idx].contiguous(), lr * decay ** step)
            step += 1
        with t.no_grad():
            sub = perm[:min(M, 65536)]
            y = mlp.predict(X[sub].contiguous())
            hist.append(float(((y - T[sub]) ** 2).mean()))
        if log:
            log(f"epoch {e}: loss {hist[-1]:.3e} (target var {float(T.var()):.3e})")
    return hist


__all__ = ["WarmStartMLP", "set_gate", "bd_step", "sync_host", "run_trajectory", "fit"]


def compare_starts(cr, mlp: WarmStartMLP, steps: int, dt: float, field_range=(0.5, 2.0), seed: int = 0,
                   generator=None, log=None):
    """Continue ``cr``'s trajectory for ``steps`` Brownian steps with a new random |E0| per
    step and solve each step four times: from zero, from the previous step's solution,
    from the network's prediction, and from the previous solution plus the predicted
    response to the field change.  Returns per-mode iteration counts and solve times
    (ms, network inference included)."""
    t = _torch()
    rng = np.random.default_rng(seed)
    gen = generator
    if gen is None:
        gen = t.Generator(device=cr.dsys.dev)
        gen.manual_seed(seed)
    modes = ("zero", "previous", "learned", "learned+previous")
    iters = {k: [] for k in modes}
    ms = {k: [] for k in modes}
    E_prev = float(cr.field[2]) if hasattr(cr, "field") else 1.0
    cr.computeForces(0, x0="zero")
    # one untimed pass of each start's preparation, so that first-use costs (kernel code
    # loading, allocator growth) stay out of the per-step times
    q0, P0 = cr.charges.clone(), cr.potentials.clone()
    mlp.warm_start(cr)
    cr.charges.copy_(q0)
    mlp.warm_start_from_previous(cr, (0.0, 0.0, E_prev))
    cr.charges.copy_(q0)
    cr.potentials.copy_(P0)
    t.cuda.synchronize()

    def timed(mode, prep):
        t.cuda.synchronize()
        t0 = time.perf_counter()
        prep()
        st = cr.computeForces(0, x0="zero" if mode == "zero" else ("previous" if mode == "previous" else "charges"))
        t.cuda.synchronize()
        ms[mode].append(1e3 * (time.perf_counter() - t0))
        iters[mode].append(st["iterations"])

    for k in range(steps):
        if log and k % 100 == 0 and k:
            log(f"compare step {k}: mean iters " + ", ".join(f"{m} {np.mean(iters[m]):.1f}" for m in modes))
        bd_step(cr, dt, generator=gen)
        q_prev, P_prev = cr.charges.clone(), cr.potentials.clone()
        E = float(rng.uniform(*field_range))
        cr.update_field((0.0, 0.0, E))
        timed("zero", lambda: None)

        def restore():
            cr.charges.copy_(q_prev)
            cr.potentials.copy_(P_prev)

        timed("previous", restore)
        timed("learned", lambda: mlp.warm_start(cr))

        def corr():
            cr.charges.copy_(q_prev)
            mlp.warm_start_from_previous(cr, (0.0, 0.0, E_prev))

        timed("learned+previous", corr)
        E_prev = E
    return iters, ms


def main(argv=None) -> int:
    """Learned warm start protocol: a Brownian trajectory of the workload with |E0|
    uniform in [0.5, 2] per step; the network is trained on (features, solved charges)
    pairs from the first ``--train`` steps and evaluated on the rest against a zero guess
    and the previous-step start.  Prints one JSON line."""
    import argparse
    import json

    from .compute import ConductorRigid
    from .system import make_workload

    ap = argparse.ArgumentParser(prog="python -m conductorrigid_amd.warmstart")
    ap.add_argument("--workload", default="n100k")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--train", type=int, default=1500)
    ap.add_argument("--collect-every", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--dt", type=float, default=2e-4)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[warmstart] {m}", flush=True)  # noqa: E731
    s, w = make_workload(a.workload)
    cr = ConductorRigid(system=s, field=(0.0, 0.0, 1.0), xi=w["xi"], errortol=w["errortol"], rc=w["rc"],
                        mesh=w["mesh"], P=w["P"], restart=w["restart"], gmres_tol=w["tol"], max_iter=2000,
                        precision=a.precision, preconditioner="block")
    mlp = WarmStartMLP(device=0, seed=a.seed)
    t0 = time.perf_counter()
    recs, pairs = run_trajectory(cr, a.train, a.dt, collect=range(0, a.train, a.collect_every), mlp=mlp,
                                 seed=a.seed, log=log)
    t_traj = time.perf_counter() - t0
    log(f"trajectory {a.train} steps in {t_traj:.1f} s, {len(pairs)} frames collected")
    t0 = time.perf_counter()
    hist = fit(mlp, pairs, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    t_fit = time.perf_counter() - t0
    del pairs
    if a.save:
        mlp.save(a.save)
    iters, ms = compare_starts(cr, mlp, a.steps - a.train, a.dt, seed=a.seed + 1, log=log)
    out = {"workload": a.workload, "beads": s.N, "precision": a.precision, "train_steps": a.train,
           "eval_steps": a.steps - a.train, "frames": len(range(0, a.train, a.collect_every)),
           "train_rows": len(range(0, a.train, a.collect_every)) * s.N, "fit_s": t_fit, "trajectory_s": t_traj,
           "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
           "iters": {k: float(np.mean(v)) for k, v in iters.items()},
           "ms_per_eval": {k: float(np.mean(v)) for k, v in ms.items()}}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
