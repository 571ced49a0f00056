"""Benchmark: force evaluations per second of the ConductorRigid force compute on one GPU.

One step = one force evaluation (computeForces): positions -> GMRES solve for the
bead charges and body potentials -> bead forces.  The configuration is static and
its neighbour list is built once, outside the timed region (its build time is reported
separately).  Workload: N = 1,000,000 beads (50,000 rigid bodies x 20 beads, volume
fraction 0.3, seeded), mesh 256^3, P = 8, rc = 4, xi = 0.6, GMRES restart 30, tolerance
1e-6, E0 = (0,0,1), near-field block preconditioner.

Headline value: FP32-bulk mode (spread / FFT / contract / pair sum in float, Krylov
and constraint sums in double), solved from x0 = 0 every step.  Beside it:
  * the plain FP64 GPU solve on the same system, with the relative force and charge
    error of the FP32-bulk result against it;
  * errors of both precisions against the FP64 direct-Ewald fixture (1000 beads at the
    headline's xi, rc and P);
  * roofline: the dominant matvec kernel's algorithmic bytes per launch over its
    average launch duration, timed with HIP events on the library's stream;
  * cpu_baseline: the CPU FP64 restatement of the reference method (oracle/, diagonal
    preconditioner, CUSP's GMRES rule) on this host: at N = 1M one operator application
    times the reference method's operator count (measured on the GPU); at N = 1000 and
    100k full solves where they fit the time budget;
  * the learned warm start on the 100k-bead Brownian protocol (GMRES iterations and
    times from zero, the previous solution, the network's prediction, and both);
  * the batched-boxes workload (64 boxes) against solving the boxes one by one.
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
    "mesh_clear": "k_charges_fx + k_fx_scale",
    "spread": "k_spread_fx_run<8, 16, 16, 16> + k_fx_gather_zc<8, 16, 16, 16>",
    "fft_z_r2c": "k_fftr_z_r2c<float, 16, 16>",
    "fft_y_fwd": "k_fftr_strided<float, 16, 16, -1>",
    "fft_x_scale": "k_fftr_x_scale<float, 16, 16>",
    "fft_y_inv": "k_fftr_strided<float, 16, 16, 1>",
    "fft_z_c2r": "k_fftr_z_c2r<float, 16, 16>",
    "contract": "k_contract_fx<8, 16, 16, 16, false>",
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
        # FP32 bulk: the fixed-point charge rounding (x in, float charges out); the
        # fixed-point spread writes every node, so there is no clear.  A clear before an
        # atomic accumulation (FP64) is not algorithmic traffic.
        "mesh_clear": 12.0 * N if t == 4 else 0.0,
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


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the CPU baseline leg (the reference method restated in C, FP64)

    O.build()
    return O


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_matvec_seconds(sysdef, prm, nlist, reps: int = 3) -> tuple[float, int]:
    """Median wall time of one FP64 operator application of the CPU restatement."""
    O = _oracle()
    o = O.Oracle(sysdef, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"], nlist=nlist)
    x = np.random.default_rng(0).normal(size=o.n)
    o.matvec(x)  # warm (page in)
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.matvec(x)
        t.append(time.perf_counter() - t0)
    return float(np.median(t)), int(O.lib().cro_num_threads())


def cpu_full_solve(sysdef, prm, nlist, restart: int, tol: float, maxiter: int = 2000, reps: int = 3) -> dict:
    """One force evaluation of the reference method on the CPU restatement: FP64 spectral
    Ewald operator, the reference's diagonal preconditioner and CUSP's restarted GMRES
    stopping rule, then the force pass; median of reps."""
    O = _oracle()
    o = O.Oracle(sysdef, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"], nlist=nlist)
    t, st = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, _, st = o.compute_forces([0.0, 0.0, 1.0], restart=restart, maxiter=maxiter, tol=tol)
        t.append(time.perf_counter() - t0)
    dt = float(np.median(t))
    return {"s_per_eval": dt, "evals_per_sec": 1.0 / dt, "iters": int(st["iters"]),
            "converged": bool(st["converged"]), "cores": int(O.lib().cro_num_threads())}


def batch64_summary(steps: int, warmup: int, sequential: bool = True) -> dict:
    """The batched-boxes workload (64 independent boxes, N log-uniform in [2k, 50k], bodies
    of 8/20/40 beads, volume fraction 0.1-0.3, per-box mesh and xi for 1e-6 Ewald error,
    P = 9; FP32 bulk with FP64 Krylov, near-field block preconditioner, restart 30, tol
    1e-6, x0 = 0).  One step = one force evaluation of every box, timed as the lockstep
    batch with its Arnoldi step replayed from a HIP graph, the same batch launched
    eagerly, and (sequential) the boxes solved one after another by the single-box path."""
    import torch

    from conductorrigid_amd.compute import BatchConductorRigid, batch_boxes
    from conductorrigid_amd.system import make_batch

    t0 = time.perf_counter()
    items = make_batch(64, seed=4, workers=8)
    N = [s.N for s, _ in items]
    meshes = [p["mesh"][0] for _, p in items]
    phis = [p["phi"] for _, p in items]
    log(f"batch: {len(items)} boxes, {sum(N)} beads (N {min(N)}..{max(N)}, phi {min(phis):.3f}..{max(phis):.3f}), "
        f"meshes {sorted(set(meshes))} "
        f"({time.perf_counter() - t0:.1f} s)")
    boxes = batch_boxes(items, precision="fp32", preconditioner="block", restart=30, max_iter=1000, gmres_tol=1e-6)
    t0 = time.perf_counter()
    tunes = [b.autotune() for b in boxes]
    t_tune = time.perf_counter() - t0
    log(f"autotune of {len(boxes)} boxes: {t_tune:.1f} s")

    def timed(fn, n, w):
        for _ in range(w):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = None
        for _ in range(n):
            st = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n, st

    def residual_check(on):
        for b in boxes:
            b.set_residual_check(on)

    res = {}
    for use_graph in (True, False):
        batch = BatchConductorRigid(boxes, use_graph=use_graph)
        # timed: solve + forces; the true-residual check runs on one untimed evaluation
        residual_check(False)
        dt, st = timed(lambda: batch.computeForces(x0="zero"), steps, warmup)
        residual_check(True)
        st = batch.computeForces(x0="zero")
        key = "batch_graph" if use_graph else "batch_eager"
        res[key] = {"ms_per_eval": dt * 1e3, "iters_max": max(x["iterations"] for x in st),
                    "iters_mean": float(np.mean([x["iterations"] for x in st])),
                    "all_converged": all(x["converged"] == 1 for x in st),
                    "max_true_relres": max(x["true_relres"] for x in st)}
        log(f"{key}: {dt * 1e3:.2f} ms per batch eval; iters mean {res[key]['iters_mean']:.1f} "
            f"max {res[key]['iters_max']}")
        del batch
    if sequential:
        residual_check(False)
        dts, sts = timed(lambda: [b.computeForces(x0="zero") for b in boxes], max(1, min(steps, 2)), 1)
        residual_check(True)
        res["sequential"] = {"ms_per_eval": dts * 1e3, "iters_mean": float(np.mean([x["iterations"] for x in sts]))}
        log(f"sequential single-box solves: {dts * 1e3:.2f} ms for all boxes")
    dt = res["batch_graph"]["ms_per_eval"] * 1e-3
    return {"workload": f"batch64: {len(items)} boxes, {sum(N)} beads, N log-uniform [2000, 50000], bodies 8/20/40, "
                        f"phi U[0.1, 0.35] ({min(phis):.3f}-{max(phis):.3f} drawn), meshes {min(meshes)}^3-{max(meshes)}^3, "
                        f"P=9, tol 1e-6, restart 30, "
                        f"near-field block preconditioner, x0=0",
            "boxes": len(items), "beads": sum(N), "ms_per_eval": dt * 1e3, "batch_evals_per_sec": 1.0 / dt,
            "beads_per_sec": sum(N) / dt,
            "speedup_vs_sequential": (res["sequential"]["ms_per_eval"] / res["batch_graph"]["ms_per_eval"]
                                      if "sequential" in res else None),
            "graph_vs_eager": res["batch_eager"]["ms_per_eval"] / res["batch_graph"]["ms_per_eval"],
            "modes": res, "autotune_s": t_tune, "tiles": [t["tile"] for t in tunes]}


def bench_batch(args) -> int:
    b = batch64_summary(args.steps, args.warmup)
    out = {
        "metric": "batched force evals/sec (64 boxes, all boxes per eval)",
        "value": b["batch_evals_per_sec"],
        "unit": "batch evals/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_eval"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 krylov",
        "data": "synthetic (seeded rigid-cluster boxes)",
        "config": {"workload": b["workload"], "beads": b["beads"], "boxes": b["boxes"]},
        **{k: v for k, v in b.items() if k not in ("workload",)},
    }
    print(json.dumps(out), flush=True)
    return 0


# Compile-time tile shapes of the fixed-point spread (k_spread_fx.hip) per window support.
FX_SHAPES = {8: {(16, 16, 16), (16, 8, 16), (16, 16, 32), (32, 16, 16), (16, 32, 32), (8, 8, 8)},
             9: {(16, 16, 16), (16, 8, 16), (8, 8, 8)}}


FXC_SHAPES = {8: {(16, 16, 16), (16, 8, 16), (8, 16, 16), (16, 16, 32), (8, 8, 8)},
              9: {(16, 16, 16), (16, 8, 16), (8, 8, 8)}}


def spread_kernel_name(P: int, tile, mesh) -> str:
    """The FP32-bulk spread stage: the tile kernel and the gather of its blocks (for
    compile-time tile shapes the persistent form, whose blocks carry the z halo)."""
    t = tuple(int(v) for v in tile)
    if t in FX_SHAPES.get(P, ()):
        return f"k_spread_fx_run<{P}, {t[0]}, {t[1]}, {t[2]}> + k_fx_gather_zc<{P}, {t[0]}, {t[1]}, {t[2]}>"
    ys = (t[2] + P) & ~1  # block row pitch
    v = 4 if t[2] % 4 == 0 and ys % 4 == 0 and int(mesh[2]) % 4 == 0 else 1
    return f"k_spread_fx<{P}, 0, 0, 0, false> + k_fx_gather<{v}>"


def contract_kernel_name(P: int, tile) -> str:
    t = tuple(int(v) for v in tile)
    if t in FXC_SHAPES.get(P, ()):
        return f"k_contract_fx<{P}, {t[0]}, {t[1]}, {t[2]}, false>"
    return f"k_contract_fx<{P}, 0, 0, 0, false>"


def direct_ewald_errors() -> dict:
    """FP64 and FP32-bulk solves of the committed 1000-bead direct-Ewald fixture (the
    headline's xi = 0.6, rc = 4, P = 8 on a 27^3 mesh, h = 1.02) against its FP64 direct
    Ewald charges and forces (explicit k-sum, no mesh)."""
    from conductorrigid_amd.compute import ConductorRigid
    from conductorrigid_amd.system import make_system

    f = dict(np.load(os.path.join(ROOT, "tests", "golden", "direct_ewald_n1000.npz"), allow_pickle=False))
    nb, bs, seed = (int(v) for v in f["system"])
    out = {}
    q = {}
    for precision in ("fp64", "fp32"):
        s = make_system(nb, bs, float(f["phi"]), seed=seed, charges="uniform")
        cr = ConductorRigid(system=s, xi=float(f["xi"]), errortol=1e-6, rc=float(f["rc"]), mesh=(27, 27, 27), P=8,
                            precision=precision, restart=30, max_iter=3000, gmres_tol=1e-11, preconditioner="block",
                            field=(0.0, 0.0, 1.0))
        st = cr.computeForces(0, x0="zero")
        qq, F = cr.get_charges(), cr.get_forces()[:, :3]
        q[precision] = (qq, F)
        out[precision] = {"rel_charge_err": float(np.linalg.norm(qq - f["q"]) / np.linalg.norm(f["q"])),
                          "rel_force_err": float(np.linalg.norm(F - f["F"]) / np.linalg.norm(f["F"])),
                          "converged": bool(st["converged"])}
    out["fp32_vs_fp64"] = {
        "rel_charge_err": float(np.linalg.norm(q["fp32"][0] - q["fp64"][0]) / np.linalg.norm(q["fp64"][0])),
        "rel_force_err": float(np.linalg.norm(q["fp32"][1] - q["fp64"][1]) / np.linalg.norm(q["fp64"][1]))}
    out["note"] = ("1000-bead fixture at the headline's xi, rc and P (tests/golden/direct_ewald_n1000.npz). "
                   "Against direct Ewald both precisions carry the method's discretisation error at these "
                   "parameters (the rc = 4 real-space truncation and the mesh); the <= 1e-5 gate holds for FP32 "
                   "bulk against the FP64 method at the same parameters.")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="n1m")
    ap.add_argument("--warm-train", type=int, default=1500,
                    help="trajectory steps the learned warm start trains on (100k beads; 0: skip that leg)")
    ap.add_argument("--warm-eval", type=int, default=500, help="held-out steps it is evaluated on")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-boxes summary")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-Ewald error leg")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--tile", default="", metavar="TX,TY,TZ",
                    help="FP32-bulk tile shape instead of the autotuner's choice (reproducible profiling runs)")
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

    import copy
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
    common = dict(field=(0.0, 0.0, 1.0), restart=w["restart"], gmres_tol=w["tol"], max_iter=1000,
                  x0="zero", preconditioner="block", **prm)

    def time_steps(cr, steps, warmup):
        for _ in range(warmup):
            cr.computeForces(0)
        torch.cuda.synchronize()
        iters = []
        t = time.perf_counter()
        st = None
        for _ in range(steps):
            st = cr.computeForces(0)
            iters.append(st["iterations"])
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, iters, st

    cr32 = ConductorRigid(system=copy.deepcopy(sysdef), precision="fp32", **common)
    # neighbour-list build (HOOMD's side of computeForces; outside the timed evaluation)
    nl_ms = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cr32.dsys.build_nlist()
        torch.cuda.synchronize()
        nl_ms.append(1e3 * (time.perf_counter() - t0))
    nlist_build_ms = float(np.median(nl_ms))
    log(f"params {cr32.params['mesh']} P={cr32.params['P']} rc={cr32.params['rc']} pairs={cr32.dsys.n_pairs} "
        f"nlist build {nlist_build_ms:.2f} ms")
    if args.tile:  # a fixed tiling (profiling runs); otherwise the autotuner's, untimed setup
        tile = [int(v) for v in args.tile.split(",")]
        cr32.set_tiles(tile, 131072)
        tune32 = {"tile": tile, "lds_bytes": 131072, "fixed": True}
    else:
        tune32 = cr32.autotune()
    log(f"autotune fp32: {tune32}")
    # timed evaluations do what the reference's computeForces does (solve + forces); the
    # true-residual check (one extra operator application) runs on an untimed evaluation
    cr32.set_residual_check(False)
    dt32, it32, st32 = time_steps(cr32, args.steps, args.warmup)
    cr32.set_residual_check(True)
    st32 = cr32.computeForces(0)
    F32 = cr32.get_forces()
    q32 = cr32.get_charges()
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
    stage_kernel = dict(STAGE_KERNEL)
    stage_kernel["spread"] = spread_kernel_name(P_used, tune32["tile"], mesh)
    stage_kernel["contract"] = contract_kernel_name(P_used, tune32["tile"])
    stages = {}
    for k, name in enumerate(_lib.STAGE_NAMES):
        b = algorithmic_bytes(name, sysdef.N, G, Gh, nn, sysdef.N_isolated, 4)
        gbs = b / (ms[k] * 1e-3) / 1e9 if ms[k] > 0 and b > 0 else None
        stages[name] = {"ms": ms[k], "GBs": gbs, "bytes": b, "kernel": stage_kernel[name]}
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
    nl_host = cr32.dsys.nlist_host()
    del cr32

    extra = {"nlist_build_ms": nlist_build_ms, "nlist_pairs": int(nl_host[1].sum())}
    # the plain method: FP64 throughout, same preconditioner and start
    if not args.no_fp64:
        cr64 = ConductorRigid(system=copy.deepcopy(sysdef), precision="fp64", **common)
        tune64 = cr64.autotune()
        log(f"autotune fp64: {tune64}")
        cr64.set_residual_check(False)
        dt64, it64, st64 = time_steps(cr64, max(1, min(args.steps, 3)), 1)
        F64 = cr64.get_forces()
        errF = float(np.linalg.norm(F32[:, :3] - F64[:, :3]) / np.linalg.norm(F64[:, :3]))
        errq = float(np.linalg.norm(q32 - cr64.get_charges()) / np.linalg.norm(cr64.get_charges()))
        log(f"fp64: {dt64 * 1e3:.2f} ms/eval iters={it64}; fp32 rel force err {errF:.2e} charge err {errq:.2e}")
        extra.update(autotune_fp64=tune64, fp64_evals_per_sec=1.0 / dt64, fp64_ms_per_eval=dt64 * 1e3,
                     fp64_iters=it64[-1], fp64_matvecs=st64["matvecs"], speedup_vs_fp64=dt64 / dt32,
                     rel_force_err_vs_fp64=errF, rel_charge_err_vs_fp64=errq)
        del cr64

    # the reference method on the GPU (FP64, its diagonal preconditioner and CUSP's stopping
    # rule, restart 30, at most 1000 iterations): its matvec count scales the CPU baseline
    ref = None
    if not args.no_cpu_baseline:
        crd = ConductorRigid(system=copy.deepcopy(sysdef), precision="fp64",
                             **{**common, "preconditioner": "diagonal"})
        crd.set_residual_check(False)
        t0 = time.perf_counter()
        std = crd.computeForces(0)
        torch.cuda.synchronize()
        ref = {"iters": std["iterations"], "matvecs": std["matvecs"], "converged": bool(std["converged"]),
               "gpu_ms": 1e3 * (time.perf_counter() - t0)}
        log(f"reference method on the GPU: {ref}")
        del crd

    cpu = None
    if not args.no_cpu_baseline:
        torch.cuda.synchronize()
        log("cpu baseline ...")
        t_mv, cores = cpu_matvec_seconds(sysdef, prm, nl_host)
        m = ref["matvecs"] + 1
        conv = "reaches" if ref["converged"] else "does not reach"
        cpu = {"value": 1.0 / (t_mv * m), "unit": "force evals/sec", "cores": cores, "kind": "port",
               "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
               "sample": f"median of 3 FP64 operator applications of the CPU restatement of the reference method on "
                         f"the full N={sysdef.N} system ({t_mv:.3f} s each, {cores} threads), times the "
                         f"{ref['matvecs']} applications + 1 force pass of the reference method's solve (FP64, "
                         f"diagonal preconditioner, restart {w['restart']}, tol {w['tol']}) measured on the GPU: "
                         f"it {conv} the tolerance within 1000 iterations"
                         + ("" if ref["converged"] else ", so the value is an upper bound on its rate")}
        log(f"cpu baseline {cpu['value']:.5f} evals/s ({cores} threads)")
        # full CPU solves of the reference method at the smaller configurations
        refcpu = {}
        for wl in ("n1k", "n100k"):
            s2, w2 = make_workload(wl)
            p2 = dict(xi=w2["xi"], errortol=w2["errortol"], rc=w2["rc"], mesh=w2["mesh"], P=w2["P"])
            cr2 = ConductorRigid(system=s2, precision="fp64", field=(0.0, 0.0, 1.0), restart=w2["restart"],
                                 gmres_tol=w2["tol"], max_iter=2000, x0="zero", preconditioner="diagonal", **p2)
            st2 = cr2.computeForces(0)
            nl2 = cr2.dsys.nlist_host()
            del cr2
            t_mv2, _ = cpu_matvec_seconds(s2, p2, nl2, reps=1)
            est = t_mv2 * (st2["matvecs"] + 1)
            if est < 40.0:
                reps = 3 if est < 8.0 else 1
                r = cpu_full_solve(s2, p2, nl2, restart=w2["restart"], tol=w2["tol"], reps=reps)
                r["method"] = f"full FP64 CPU solve + force pass, median of {reps}"
            else:
                r = {"s_per_eval": est, "evals_per_sec": 1.0 / est, "iters": st2["iterations"],
                     "converged": bool(st2["converged"]),
                     "method": "CPU operator application x the reference method's GPU-measured count + 1"}
            r.update(beads=s2.N, gpu_iters_reference_method=st2["iterations"])
            refcpu[wl] = r
            log(f"cpu reference method {wl}: {r}")
        extra["cpu_reference_method"] = refcpu

    # accuracy against the FP64 direct-Ewald oracle (committed 1000-bead fixture)
    if not args.no_direct:
        de = direct_ewald_errors()
        extra["direct_ewald"] = de
        log(f"direct Ewald: {de['fp64']} / {de['fp32']}")

    # learned warm start: the 100k-bead protocol (Brownian trajectory, |E0| ~ U[0.5, 2] per
    # step; the network trained on the GPU on solver outputs of the first steps, evaluated
    # on the held-out later steps from four starts)
    warm = None
    if args.warm_train > 0 and args.warm_eval > 0:
        import conductorrigid_amd.warmstart as WS

        s5, w5 = make_workload("n100k")
        cr5 = ConductorRigid(system=s5, field=(0.0, 0.0, 1.0), xi=w5["xi"], errortol=w5["errortol"], rc=w5["rc"],
                             mesh=w5["mesh"], P=w5["P"], restart=w5["restart"], gmres_tol=w5["tol"], max_iter=2000,
                             precision="fp32", preconditioner="block")
        mlp = WS.WarmStartMLP(device=0, seed=0)
        t0 = time.perf_counter()
        every = 10
        _, pairs = WS.run_trajectory(cr5, args.warm_train, 2e-4, collect=range(0, args.warm_train, every), mlp=mlp,
                                     seed=0)
        t_traj = time.perf_counter() - t0
        t0 = time.perf_counter()
        hist = WS.fit(mlp, pairs, epochs=4, batch=8192, lr=2e-3)
        t_fit = time.perf_counter() - t0
        del pairs
        cr5.set_residual_check(False)  # timed like the headline: solve + forces
        iters, wms = WS.compare_starts(cr5, mlp, args.warm_eval, 2e-4, seed=1)
        warm = {"workload": f"n100k: N={s5.N} beads, {s5.N_rigid} rigid bodies x 20, phi=0.3, mesh 128^3, P=8, "
                            f"rc=4, xi=0.6, FP32 bulk, tol 1e-6",
                "protocol": f"Brownian trajectory (dt 2e-4, rigid bodies, contact repulsion), |E0| ~ U[0.5, 2] per "
                            f"step; network trained on {len(range(0, args.warm_train, every))} frames of the first "
                            f"{args.warm_train} steps, evaluated on the next {args.warm_eval}",
                "iters": {k: float(np.mean(v)) for k, v in iters.items()},
                "ms_per_eval": {k: float(np.mean(v)) for k, v in wms.items()},
                "train_loss": [hist[0], hist[-1]], "trajectory_s": t_traj, "fit_s": t_fit,
                "learned_speedup_vs_zero": float(np.mean(wms["zero"]) / np.mean(wms["learned"]))}
        log("warm starts: " + ", ".join(f"{k} {warm['iters'][k]:.1f} it / {warm['ms_per_eval'][k]:.1f} ms"
                                        for k in warm["iters"]))
        del cr5, mlp

    if not args.no_batch:
        extra["batch64"] = batch64_summary(max(1, min(args.steps, 3)), 1)

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
                        "reference_method_diagonal_precond": ref["iters"] if ref else None,
                        "n100k_trajectory": warm["iters"] if warm else None},
        "matvecs_per_eval": st32["matvecs"],
        "true_relres": st32["true_relres"],
        "rel_force_err_vs_fp64": extra.get("rel_force_err_vs_fp64"),
        "rel_force_err_vs_direct_ewald": extra["direct_ewald"]["fp32"]["rel_force_err"] if "direct_ewald" in extra
        else None,
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
