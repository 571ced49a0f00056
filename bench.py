"""Benchmark: force evaluations per second of the ConductorRigid force compute on one GPU.

One step = one force evaluation (computeForces): positions -> GMRES solve for the
bead charges and body potentials -> bead forces.  The configuration is static and
its neighbour list is built once, outside the timed region.  Workload: N = 1,000,000
beads (50,000 rigid bodies x 20 beads, volume fraction 0.3, seeded), mesh 256^3,
P = 8, rc = 4, xi = 0.6, GMRES restart 30, tolerance 1e-6, E0 = (0,0,1).

Headline value: FP32-bulk mode (spread / FFT / contract / pair sum in float, Krylov
and constraint sums in double), solved from x0 = 0 every step.  The plain FP64 GPU
solve on the same system is timed beside it, with the relative force error of the
FP32-bulk result against it.  roofline: the dominant matvec kernel's algorithmic
bytes per launch over its HIP-event-timed average duration.  cpu_baseline: the CPU
FP64 restatement (oracle/) timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stage: str, N: int, G: int, Gh: int, nn: float, B: int, t: int) -> float:
    """Minimal HBM bytes a stage must move: its inputs once and its outputs once.
    t = bytes per bulk scalar (4 FP32-bulk, 8 FP64); bead record = 4 scalars (x,y,z,q)."""
    rec = 4 * t
    return {
        "mesh_clear": t * G,
        "spread": rec * N + t * G,
        "fft_z_r2c": t * G + 2 * t * Gh,
        "fft_y_fwd": 4 * t * Gh,
        "fft_x_scale": 4 * t * Gh,
        "fft_y_inv": 4 * t * Gh,
        "fft_z_c2r": 2 * t * Gh + t * G,
        "contract": rec * N + t * G + t * N,
        "real_space": (rec + 2 * t + 4 * nn) * N + 16 * B,
        "constraint_rows": 12 * N + 16 * B,
    }[stage]


def cpu_baseline(sysdef, prm, matvecs_per_eval: float, nlist) -> dict:
    """FP64 CPU restatement (oracle): one matvec on the full system, scaled by the
    GPU-measured matvecs per force evaluation (+1 force pass ~ one matvec)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU baseline leg

    O.build()
    o = O.Oracle(sysdef, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"], nlist=nlist)
    x = np.random.default_rng(0).normal(size=o.n)
    o.matvec(x)  # warm (page in)
    t0 = time.perf_counter()
    o.matvec(x)
    dt = time.perf_counter() - t0
    per_eval = dt * (matvecs_per_eval + 1.0)
    return {"value": 1.0 / per_eval, "unit": "force evals/sec", "cores": int(O.lib().cro_num_threads()),
            "kind": "port",
            "sample": f"one FP64 matvec of the CPU restatement on the full N={sysdef.N} system "
                      f"({dt:.2f} s), scaled by {matvecs_per_eval:.0f}+1 matvecs per force evaluation "
                      f"(the FP64 GPU solve's count)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="n1m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        return 2

    import ctypes as C

    import torch

    from conductorrigid_amd import _lib
    from conductorrigid_amd.compute import ConductorRigid
    from conductorrigid_amd.system import make_workload

    t0 = time.perf_counter()
    sysdef, w = make_workload(args.workload)
    log(f"system {args.workload}: N={sysdef.N} bodies={sysdef.N_isolated} L={sysdef.L[0]:.3f} "
        f"({time.perf_counter() - t0:.1f} s)")
    prm = dict(xi=w["xi"], errortol=w["errortol"], rc=w["rc"], mesh=w["mesh"], P=w["P"])
    common = dict(system=sysdef, field=(0.0, 0.0, 1.0), restart=w["restart"], gmres_tol=w["tol"], max_iter=1000,
                  x0="zero", preconditioner="block", **prm)

    def time_steps(cr, steps, warmup):
        for _ in range(warmup):
            cr.computeForces(0)
        torch.cuda.synchronize()
        iters = []
        t = time.perf_counter()
        for _ in range(steps):
            st = cr.computeForces(0)
            iters.append(st["iterations"])
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, iters, st

    cr32 = ConductorRigid(precision="fp32", **common)
    log(f"params {cr32.params['mesh']} P={cr32.params['P']} rc={cr32.params['rc']} pairs={cr32.dsys.n_pairs}")
    dt32, it32, st32 = time_steps(cr32, args.steps, args.warmup)
    F32 = cr32.get_forces()
    log(f"fp32-bulk: {dt32 * 1e3:.2f} ms/eval iters={it32} conv={st32['converged']} true_resid={st32['true_resid']:.3e} tol={st32['tolerance']:.3e}")

    # per-stage kernel timing (HIP events on the library's stream)
    lib = _lib.load()
    x = torch.from_numpy(np.concatenate([cr32.get_charges(), cr32.get_potentials()])).cuda()
    ms = (C.c_double * _lib.CR_NUM_STAGES)()
    sv = cr32.dsys.view()
    _lib.check(lib.cr_profile_matvec(cr32.ctx, C.byref(sv), x.data_ptr(), args.profile_reps, ms,
                                     _lib.CR_NUM_STAGES, cr32.stream()))
    mesh = cr32.params["mesh"]
    G = int(np.prod(mesh))
    Gh = mesh[0] * mesh[1] * (mesh[2] // 2 + 1)
    nn = cr32.dsys.n_pairs / sysdef.N
    stages = {}
    for k, name in enumerate(_lib.STAGE_NAMES):
        b = algorithmic_bytes(name, sysdef.N, G, Gh, nn, sysdef.N_isolated, 4)
        stages[name] = {"ms": ms[k], "GBs": b / (ms[k] * 1e-3) / 1e9 if ms[k] > 0 else None, "bytes": b}
    dom = max(stages, key=lambda k: stages[k]["ms"])
    log("stages " + ", ".join(f"{k}={v['ms'] * 1e3:.1f}us" for k, v in stages.items()))
    roof = {"bound": "hbm", "kernel": dom, "achieved": stages[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": stages[dom]["GBs"] / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_launch": stages[dom]["bytes"], "avg_launch_ms": stages[dom]["ms"]}

    extra = {}
    if not args.no_fp64:
        cr64 = ConductorRigid(precision="fp64", **common)
        dt64, it64, st64 = time_steps(cr64, max(1, min(args.steps, 3)), 1)
        F64 = cr64.get_forces()
        errF = float(np.linalg.norm(F32[:, :3] - F64[:, :3]) / np.linalg.norm(F64[:, :3]))
        errq = float(np.linalg.norm(cr32.get_charges() - cr64.get_charges()) / np.linalg.norm(cr64.get_charges()))
        log(f"fp64: {dt64 * 1e3:.2f} ms/eval iters={it64}; fp32 rel force err {errF:.2e} charge err {errq:.2e}")
        extra.update(fp64_evals_per_sec=1.0 / dt64, fp64_ms_per_eval=dt64 * 1e3, fp64_iters=it64[-1],
                     fp64_matvecs=st64["matvecs"], speedup_vs_fp64=dt64 / dt32, rel_force_err_vs_fp64=errF,
                     rel_charge_err_vs_fp64=errq)
        mv = st64["matvecs"]
        nl = cr64.dsys.nlist_host()
        del cr64
    else:
        mv = st32["matvecs"]
        nl = cr32.dsys.nlist_host()
    cpu = None
    if not args.no_cpu_baseline:
        torch.cuda.synchronize()
        log("cpu baseline ...")
        cpu = cpu_baseline(sysdef, prm, mv, nl)
        log(f"cpu baseline {cpu['value']:.4f} evals/s ({cpu['cores']} threads)")

    out = {
        "metric": "force evals/sec at N=1M beads, rel err<=1e-5; GMRES iters plain vs warm-start",
        "value": 1.0 / dt32,
        "unit": "force evals/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 krylov",
        "data": "synthetic (seeded rigid-cluster system)",
        "config": {"workload": f"{args.workload}: N={sysdef.N} beads, {sysdef.N_rigid} rigid bodies x 20, phi=0.3, "
                               f"mesh {mesh[0]}^3, P={cr32.params['P']}, rc={cr32.params['rc']}, xi={prm['xi']}, "
                               f"GMRES restart {w['restart']} tol {w['tol']}, x0=0",
                   "beads": sysdef.N, "mesh": mesh, "gmres_iters": it32[-1], "matvecs": st32["matvecs"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {k: v["ms"] for k, v in stages.items()},
        **extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
