"""Benchmark: force evaluations per second of the ConductorRigid force compute on one GPU.

One step = one force evaluation (computeForces): positions -> GMRES solve for the
bead charges and body potentials -> bead forces.  The configuration is static and
its neighbour list is built once, outside the timed region.  Workload: N = 1,000,000
beads (50,000 rigid bodies x 20 beads, volume fraction 0.3, seeded), mesh 256^3,
P = 8, rc = 4, xi = 0.6, GMRES restart 30, tolerance 1e-6, E0 = (0,0,1), near-field
block preconditioner.

Headline value: FP32-bulk mode (spread / FFT / contract / pair sum in float, Krylov
and constraint sums in double), solved from x0 = 0 every step.  Beside it:
  * the plain FP64 GPU solve on the same system, with the relative force and charge
    error of the FP32-bulk result against it;
  * GMRES iterations from x0 = 0 and from the previous step's solution along a short
    run of rigid random body translations;
  * roofline: the dominant matvec kernel's algorithmic bytes per launch over its
    average launch duration, timed with HIP events on the library's stream;
  * cpu_baseline: the CPU FP64 restatement (oracle/) timed on this host on a
    bounded sample (one full-size matvec, scaled by the solve's matvec count).
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

METRIC = "force evals/sec at N=1M beads, rel err≤1e-5; GMRES iters plain vs warm-start"
HBM_PEAK_GBS = 8000.0

# matvec stage -> kernel it times (FP32-bulk, P = 8)
STAGE_KERNEL = {
    "mesh_clear": "k_fill16",
    "spread": "k_spread_tiled<float, double, 8>",
    "fft_z_r2c": "k_fft_z_r2c<float>",
    "fft_y_fwd": "k_fft_strided<float>",
    "fft_x_scale": "k_fft_x_scale<float>",
    "fft_y_inv": "k_fft_strided<float>",
    "fft_z_c2r": "k_fft_z_c2r<float>",
    "contract": "k_contract_tiled<float, false, 8>",
    "real_space": "k_real_space_csr<float>",
    "constraint_rows": "k_sum_charges + k_set_charge",
}


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 byte-counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "spread_n1m_fp32_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return float(line.split(":")[-1])
    except OSError:
        pass
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stage: str, N: int, G: int, Gh: int, nn: float, B: int, t: int) -> float:
    """Minimal HBM bytes a stage must move: its inputs once and its outputs once.
    t = bytes per bulk scalar (4 FP32-bulk, 8 FP64); a bead record is 4 scalars
    (x, y, z, q).  The mesh clear before the spread's atomic accumulation has no
    algorithmic bytes: its time is overhead of the method."""
    rec = 4 * t
    return {
        "mesh_clear": 0.0,  # clearing before the atomic accumulation is not algorithmic traffic
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
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.matvec(x)
        reps.append(time.perf_counter() - t0)
    dt = float(np.median(reps))
    per_eval = dt * (matvecs_per_eval + 1.0)
    return {"value": 1.0 / per_eval, "unit": "force evals/sec", "cores": int(O.lib().cro_num_threads()),
            "kind": "port",
            "sample": f"median of 3 FP64 matvecs of the CPU restatement on the full N={sysdef.N} system "
                      f"({dt:.3f} s each), scaled by {matvecs_per_eval:.0f}+1 matvecs per force evaluation "
                      f"(the FP64 GPU solve's count)"}


def bench_batch(args) -> int:
    """64 independent boxes (N log-uniform in [2k, 50k], bodies of 8/20/40 beads, volume
    fraction 0.1-0.3, per-box mesh and xi for 1e-6 Ewald error, P = 9), FP32 bulk with
    FP64 Krylov, near-field block preconditioner, restart 30, tolerance 1e-6, x0 = 0.
    One step = one force evaluation of every box.  Timed three ways on the same boxes:
    the lockstep batch with its Arnoldi step replayed from a HIP graph, the same batch
    launched eagerly, and the boxes solved one after another by the single-box path."""
    import torch

    from conductorrigid_amd.compute import BatchConductorRigid, batch_boxes
    from conductorrigid_amd.system import make_batch

    t0 = time.perf_counter()
    items = make_batch(64, seed=4, workers=8)
    N = [s.N for s, _ in items]
    meshes = [p["mesh"][0] for _, p in items]
    log(f"batch: {len(items)} boxes, {sum(N)} beads (N {min(N)}..{max(N)}), meshes {sorted(set(meshes))} "
        f"({time.perf_counter() - t0:.1f} s)")
    boxes = batch_boxes(items, precision="fp32", preconditioner="block", restart=30, max_iter=1000, gmres_tol=1e-6)
    t0 = time.perf_counter()
    tunes = [b.autotune() for b in boxes]
    t_tune = time.perf_counter() - t0
    log(f"autotune of {len(boxes)} boxes: {t_tune:.1f} s")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            st = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, st

    def residual_check(on):
        for b in boxes:
            b.set_residual_check(on)

    res = {}
    for use_graph in (True, False):
        batch = BatchConductorRigid(boxes, use_graph=use_graph)
        # timed: solve + forces; the true-residual check runs on one untimed evaluation
        residual_check(False)
        dt, st = timed(lambda: batch.computeForces(x0="zero"), args.steps, args.warmup)
        residual_check(True)
        st = batch.computeForces(x0="zero")
        key = "batch_graph" if use_graph else "batch_eager"
        res[key] = {"ms_per_eval": dt * 1e3, "iters_max": max(s["iterations"] for s in st),
                    "iters_mean": float(np.mean([s["iterations"] for s in st])),
                    "all_converged": all(s["converged"] == 1 for s in st),
                    "max_true_relres": max(s["true_relres"] for s in st)}
        log(f"{key}: {dt * 1e3:.2f} ms per batch eval; iters mean {res[key]['iters_mean']:.1f} "
            f"max {res[key]['iters_max']}")
        del batch

    def sequential():
        return [b.computeForces(x0="zero") for b in boxes]

    residual_check(False)
    dts, sts = timed(sequential, max(1, min(args.steps, 2)), 1)
    residual_check(True)
    res["sequential"] = {"ms_per_eval": dts * 1e3, "iters_mean": float(np.mean([s["iterations"] for s in sts]))}
    log(f"sequential single-box solves: {dts * 1e3:.2f} ms for all boxes")
    dt = res["batch_graph"]["ms_per_eval"] * 1e-3
    out = {
        "metric": "batched force evals/sec (64 boxes, all boxes per eval)",
        "value": 1.0 / dt,
        "unit": "batch evals/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 krylov",
        "data": "synthetic (seeded rigid-cluster boxes)",
        "config": {"workload": f"batch64: {len(items)} boxes, {sum(N)} beads, N log-uniform [2000, 50000], bodies "
                               f"8/20/40, phi 0.1-0.3, meshes {min(meshes)}^3-{max(meshes)}^3, P=9, tol 1e-6, "
                               f"restart 30, near-field block preconditioner, x0=0",
                   "beads": sum(N), "boxes": len(items)},
        "beads_per_sec": sum(N) / dt,
        "speedup_vs_sequential": res["sequential"]["ms_per_eval"] / res["batch_graph"]["ms_per_eval"],
        "graph_vs_eager": res["batch_eager"]["ms_per_eval"] / res["batch_graph"]["ms_per_eval"],
        "modes": res,
        "autotune_s": t_tune,
        "tiles": [t["tile"] for t in tunes],
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="n1m")
    ap.add_argument("--warm-steps", type=int, default=4, help="evaluated steps of the warm-start comparison (0: skip)")
    ap.add_argument("--train-steps", type=int, default=24, help="trajectory steps the warm-start network trains on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "bytes_read": 4 * n, "bytes_written": 4 * n}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1).", file=sys.stderr)
        return 2

    if args.workload == "batch64":
        return bench_batch(args)

    import ctypes as C

    import torch

    from conductorrigid_amd import _lib
    from conductorrigid_amd.compute import ConductorRigid, update_positions
    from conductorrigid_amd.system import make_workload, translate_bodies

    t0 = time.perf_counter()
    sysdef, w = make_workload(args.workload)
    log(f"system {args.workload}: N={sysdef.N} bodies={sysdef.N_isolated} L={sysdef.L[0]:.3f} "
        f"({time.perf_counter() - t0:.1f} s)")
    prm = dict(xi=w["xi"], errortol=w["errortol"], rc=w["rc"], mesh=w["mesh"], P=w["P"])
    common = dict(field=(0.0, 0.0, 1.0), restart=w["restart"], gmres_tol=w["tol"], max_iter=1000,
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

    cr32 = ConductorRigid(system=sysdef, precision="fp32", **common)
    log(f"params {cr32.params['mesh']} P={cr32.params['P']} rc={cr32.params['rc']} pairs={cr32.dsys.n_pairs}")
    tune32 = cr32.autotune()  # setup, untimed: the tile shape for this system
    log(f"autotune fp32: {tune32}")
    # timed evaluations do what the reference's computeForces does (solve + forces); the
    # true-residual check (one extra operator application) runs on an untimed evaluation
    cr32.set_residual_check(False)
    dt32, it32, st32 = time_steps(cr32, args.steps, args.warmup)
    cr32.set_residual_check(True)
    st32 = cr32.computeForces(0)
    F32 = cr32.get_forces()
    q32 = cr32.get_charges()
    nl32 = cr32.dsys.nlist_host() if args.no_fp64 else None
    log(f"fp32-bulk: {dt32 * 1e3:.2f} ms/eval iters={it32} conv={st32['converged']} "
        f"true_relres={st32['true_relres']:.3e}")

    # per-stage kernel timing (HIP events on the library's stream)
    lib = _lib.load()
    x = torch.from_numpy(np.concatenate([cr32.get_charges(), cr32.get_potentials()])).cuda()
    ms = (C.c_double * _lib.CR_NUM_STAGES)()
    sv = cr32.dsys.view()
    _lib.check(lib.cr_profile_matvec(cr32.ctx, C.byref(sv), x.data_ptr(), args.profile_reps, ms,
                                     _lib.CR_NUM_STAGES, cr32.stream()))
    mesh = cr32.params["mesh"]
    P_used, rc_used = cr32.params["P"], cr32.params["rc"]
    G = int(np.prod(mesh))
    Gh = mesh[0] * mesh[1] * (mesh[2] // 2 + 1)
    nn = cr32.dsys.n_pairs / sysdef.N
    stages = {}
    for k, name in enumerate(_lib.STAGE_NAMES):
        b = algorithmic_bytes(name, sysdef.N, G, Gh, nn, sysdef.N_isolated, 4)
        gbs = b / (ms[k] * 1e-3) / 1e9 if ms[k] > 0 and b > 0 else None
        stages[name] = {"ms": ms[k], "GBs": gbs, "bytes": b, "kernel": STAGE_KERNEL[name]}
    dom = max(stages, key=lambda k: stages[k]["ms"])
    log("stages " + ", ".join(f"{k}={v['ms'] * 1e3:.1f}us" for k, v in stages.items()))
    d = stages[dom]
    achieved = d["GBs"] or 0.0
    traffic = measured_traffic(d["kernel"]) if args.workload == "n1m" else None
    roof = {"bound": "hbm", "kernel": d["kernel"], "stage": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": "profiles/spread_n1m_fp32_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)" if traffic
            else None, "traffic_GBs": traffic / (d["ms"] * 1e-3) / 1e9 if traffic else None,
            "bytes_per_launch": d["bytes"], "avg_launch_ms": d["ms"]}
    sc_bytes = stages["spread"]["bytes"] + stages["contract"]["bytes"]
    sc_ms = stages["mesh_clear"]["ms"] + stages["spread"]["ms"] + stages["contract"]["ms"]
    spread_contract = {"GBs": sc_bytes / (sc_ms * 1e-3) / 1e9, "frac": sc_bytes / (sc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    matvec_ms = sum(v["ms"] for v in stages.values())

    del cr32
    # Warm starts along a Brownian trajectory with |E0| uniform in [0.5, 2] per step: the
    # network is trained on the GPU on solver outputs from the first steps, then each
    # later step is solved from zero, from the previous solution, from the prediction and
    # from the previous solution plus the predicted response to the field change.
    warm = None
    if args.warm_steps > 0:
        import copy

        import conductorrigid_amd.warmstart as WS

        crw = ConductorRigid(system=copy.deepcopy(sysdef), precision="fp32", **common)
        crw.set_tiles(tune32["tile"], tune32["lds_bytes"])
        mlp = WS.WarmStartMLP(device=0, seed=0)
        t0 = time.perf_counter()
        _, pairs = WS.run_trajectory(crw, args.train_steps, 2e-4, collect=range(0, args.train_steps, 2), mlp=mlp,
                                     seed=0)
        hist = WS.fit(mlp, pairs, epochs=2, batch=16384, lr=2e-3)
        t_train = time.perf_counter() - t0
        del pairs
        crw.set_residual_check(False)  # timed like the headline: solve + forces
        iters, wms = WS.compare_starts(crw, mlp, args.warm_steps, 2e-4, seed=1)
        warm = {"protocol": f"Brownian trajectory (dt 2e-4, rigid bodies, contact repulsion), |E0| ~ U[0.5, 2] "
                            f"per step; network trained on {len(range(0, args.train_steps, 2))} frames of the "
                            f"first {args.train_steps} steps ({t_train:.1f} s incl. trajectory), evaluated on the "
                            f"next {args.warm_steps}",
                "iters": {k: float(np.mean(v)) for k, v in iters.items()},
                "ms_per_eval": {k: float(np.mean(v)) for k, v in wms.items()},
                "train_loss": [hist[0], hist[-1]],
                "learned_speedup_vs_zero": float(np.mean(wms["zero"]) / np.mean(wms["learned"]))}
        log("warm starts: " + ", ".join(f"{k} {warm['iters'][k]:.1f} it / {warm['ms_per_eval'][k]:.1f} ms"
                                        for k in warm["iters"]))
        del crw, mlp

    extra = {}
    if not args.no_fp64:
        sys64, _ = make_workload(args.workload)
        cr64 = ConductorRigid(system=sys64, precision="fp64", **common)
        tune64 = cr64.autotune()
        log(f"autotune fp64: {tune64}")
        extra["autotune_fp64"] = tune64
        cr64.set_residual_check(False)
        dt64, it64, st64 = time_steps(cr64, max(1, min(args.steps, 3)), 1)
        F64 = cr64.get_forces()
        errF = float(np.linalg.norm(F32[:, :3] - F64[:, :3]) / np.linalg.norm(F64[:, :3]))
        errq = float(np.linalg.norm(q32 - cr64.get_charges()) / np.linalg.norm(cr64.get_charges()))
        log(f"fp64: {dt64 * 1e3:.2f} ms/eval iters={it64}; fp32 rel force err {errF:.2e} charge err {errq:.2e}")
        extra.update(fp64_evals_per_sec=1.0 / dt64, fp64_ms_per_eval=dt64 * 1e3, fp64_iters=it64[-1],
                     fp64_matvecs=st64["matvecs"], speedup_vs_fp64=dt64 / dt32, rel_force_err_vs_fp64=errF,
                     rel_charge_err_vs_fp64=errq)
        mv = st64["matvecs"]
        nl = cr64.dsys.nlist_host()
        del cr64
    else:
        mv = st32["matvecs"]
        nl = nl32
    cpu = None
    if not args.no_cpu_baseline:
        torch.cuda.synchronize()
        log("cpu baseline ...")
        cpu = cpu_baseline(sysdef, prm, mv, nl)
        log(f"cpu baseline {cpu['value']:.4f} evals/s ({cpu['cores']} threads)")

    out = {
        "metric": METRIC,
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
                               f"mesh {mesh[0]}^3, P={P_used}, rc={rc_used}, xi={prm['xi']}, "
                               f"GMRES restart {w['restart']} tol {w['tol']}, near-field block preconditioner, x0=0",
                   "beads": sysdef.N, "mesh": mesh},
        "gmres_iters": {"zero": it32[-1],
                        "previous": warm["iters"]["previous"] if warm else None,
                        "learned": warm["iters"]["learned"] if warm else None,
                        "learned+previous": warm["iters"]["learned+previous"] if warm else None,
                        "zero_on_trajectory": warm["iters"]["zero"] if warm else None},
        "matvecs_per_eval": st32["matvecs"],
        "true_relres": st32["true_relres"],
        "roofline": roof,
        "spread_contract_hbm": spread_contract,
        "cpu_baseline": cpu,
        "matvec_ms": matvec_ms,
        "stages_ms": {k: v["ms"] for k, v in stages.items()},
        "autotune_fp32": tune32,
        "stages_GBs": {k: v["GBs"] for k, v in stages.items()},
        "warm_start": warm,
        **extra,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
