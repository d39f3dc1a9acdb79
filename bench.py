"""Benchmark of the explicit FP64 time-march on one MI355X.

Workload (default): the coupled 2D model — aluminium block 60 mm x 30 mm under
20 mm of water, 0.025 mm linear triangles (about 7.7M DOF), ASI2D2
fluid-solid interface, 5 mm absorbing layers (quadratic ramp) on three sides,
a seeded void, a 5 MHz toneburst pressure source and 64 monitors,
8000 time steps, FP64. One bench "step" is one full march of all time
steps over the resident operator (bfe_run); inputs are resident in HBM.

value = GDOF-steps/s = n_dof * steps / loop time, with the loop time taken
from HIP events on the march's stream (v6:162-220 interval). The roofline
object prices the dominant kernel (k_step, the fused SpMV + update) with the
algorithmic bytes per DOF-step 12 nnz/n + 4 + 24 + 8k (k = 3 coefficient
streams when damped, 1 otherwise) against the sampled HIP-event duration of
that kernel. cpu_baseline times the oracle's C/OpenMP march on this host on
the first steps of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "GDOF-steps/s + HBM GB/s (% peak); batched per-source speedup; adjoint/fwd cost"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(name, n_t=None):
    import __graft_entry__ as ge
    ge.load_package()
    from bristolfe_v2_amd import configs
    kw = {} if n_t is None else {"n_t": n_t}
    if name == "coupled_2d":
        return configs.coupled_2d(**kw)
    if name == "plate_2d":
        return configs.plate_2d(**kw)
    m = re.fullmatch(r"coupled_2d@([0-9.]+)", name)
    if m:  # reduced element size, for quick runs
        h = float(m.group(1)) * 1e-3
        return configs.coupled_2d(el_size=h, **kw)
    raise SystemExit(f"unknown workload {name}")


def copy_probe(gib=2.0):
    """A plain device copy of known size, for calibrating rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (run under rocprofv3 --pmc)."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "dispatches": 3}))


def algorithmic_bytes_per_dof(nnz, n, damped):
    k = 3 if damped else 1
    return 12.0 * nnz / n + 4.0 + 24.0 + 8.0 * k


def read_traffic(workload):
    """HBM bytes per k_step launch from the committed rocprofv3 PMC summary,
    if one exists for this workload (profiles/k_step_<workload>_pmc.md)."""
    p = ROOT / "profiles" / f"k_step_{workload.split('@')[0]}_pmc.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(c, hi, budget_s=15.0):
    from oracle.march import OracleMarch
    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    K, C, M = c["K"], c["C"], c["M"]
    n_t = len(c["time"])
    t0 = time.perf_counter()
    om = OracleMarch(K, C, M, c["dt"])
    setup = time.perf_counter() - t0
    fi = np.asarray(c["frc_gi"]) - 1
    # probe a few steps, then run a sample sized to the budget
    t0 = time.perf_counter()
    om.run(n_t, fi, c["frcs"], hi, t_end=5)
    per = (time.perf_counter() - t0) / 5
    steps = int(max(10, min(n_t, budget_s / max(per, 1e-9))))
    t0 = time.perf_counter()
    hist, ti = om.run(n_t, fi, c["frcs"], hi, t_end=steps)
    el = time.perf_counter() - t0
    done = steps - (ti - 1)
    val = K.shape[0] * done / el / 1e9
    return {"value": round(val, 4), "unit": "GDOF-steps/s", "cores": threads, "kind": "port",
            "sample": (f"first {steps} of {n_t} time steps of the same workload, oracle/march.c "
                       f"(OpenMP, {threads} threads), loop time {el:.2f} s, setup {setup:.1f} s")}, hist, steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="coupled_2d")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-every", type=int, default=16)
    ap.add_argument("--time-steps", type=int, default=None,
                    help="override the workload's number of time steps (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the march is single-GPU; --gpus {args.gpus} "
              "is not supported)", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    c = build_workload(args.workload, args.time_steps)
    log(f"workload {c['name']}: n={c['K'].shape[0]} nnz={c['K'].nnz} n_t={len(c['time'])} "
        f"setup {time.perf_counter() - t_setup:.1f} s")
    from bristolfe_v2_amd.solver import make_plan

    K, C, M = c["K"], c["C"], c["M"]
    n = K.shape[0]
    n_t = len(c["time"])
    t0 = time.perf_counter()
    plan = make_plan(K, C, M, c["dt"])
    info = plan.info()
    log(f"plan: {info} ({time.perf_counter() - t0:.1f} s)")
    fi = np.asarray(c["frc_gi"]) - 1
    hi = np.asarray(c["hist_gi"]) - 1
    F = c["frcs"]

    for w in range(args.warmup):
        plan.run(n_t, fi, F, hi)
        log(f"warmup {w + 1}/{args.warmup}: loop {plan.stats()['loop_ms']:.1f} ms")
    plan.set_kernel_timing(args.timing_every)
    # bfe_run returns after its stream has drained (hipEventSynchronize), so
    # the wall clock below is bracketed by device synchronisation.
    wall0 = time.perf_counter()
    loops, kern, samples, steps_done = [], 0.0, 0, 0
    hist = None
    for s in range(args.steps):
        hist, _, ti = plan.run(n_t, fi, F, hi)
        st = plan.stats()
        loops.append(st["loop_ms"])
        kern += st["step_kernel_ms"]
        samples += st["step_kernel_samples"]
        steps_done = st["steps"]
        if st["nonfinite"]:
            raise SystemExit(f"non-finite history values ({st['nonfinite']})")
        log(f"step {s + 1}/{args.steps}: loop {st['loop_ms']:.1f} ms, "
            f"{n * st['steps'] / st['loop_ms'] / 1e6:.2f} GDOF-steps/s")
    wall = time.perf_counter() - wall0
    loop_s = float(np.mean(loops)) / 1e3
    value = n * steps_done / loop_s / 1e9
    bpd = algorithmic_bytes_per_dof(info["nnz_k"], n, info["damped"])
    kern_avg_s = kern / max(samples, 1) / 1e3
    achieved = bpd * n / kern_avg_s / 1e9 if samples else None
    traffic = read_traffic(args.workload)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GDOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": c["name"], "n_dof": int(n), "nnz_K": int(info["nnz_k"]),
                   "time_steps": int(n_t), "steps_marched": int(steps_done),
                   "interface_rows": int(info["n_interface"]), "monitors": int(len(hi)),
                   "step": "one full march of all time steps (bfe_run)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     "kernel": "k_step<double,true> (SELL-64 SpMV + central-difference update)",
                     "bytes_per_dof_step": round(bpd, 2),
                     "kernel_avg_us": round(kern_avg_s * 1e6, 2), "kernel_samples": int(samples)},
        "loop_gdof_steps_per_s_wall": round(n * steps_done * args.steps / wall / 1e9, 4),
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        # the sample's check monitors: the forcing DOFs plus the workload's monitors
        hchk = np.concatenate([fi, hi])
        cb, hcpu, sample_steps = cpu_baseline(c, hchk)
        out["cpu_baseline"] = cb
        hg, _, _ = plan.run(sample_steps, fi, F[:, :sample_steps], hchk)
        b = hcpu[:, :sample_steps]
        out["cpu_gpu_rel_l2_on_sample"] = float(np.linalg.norm(hg - b) / max(np.linalg.norm(b), 1e-300))
    else:
        out["cpu_baseline"] = None
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
