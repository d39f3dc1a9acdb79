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
    if name == "cube_3d":
        return configs.cube_3d(**kw)
    m = re.fullmatch(r"(coupled_2d|cube_3d)@([0-9.]+)", name)
    if m:  # other element size (mm), for quick runs
        h = float(m.group(2)) * 1e-3
        return getattr(configs, m.group(1))(el_size=h, **kw)
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


def read_traffic(workload, kernel="k_step"):
    """HBM bytes per step-kernel launch from the committed rocprofv3 PMC
    summary, if one exists for this workload (profiles/<kernel>_<workload>_pmc.md)."""
    p = ROOT / "profiles" / f"{kernel}_{workload.split('@')[0]}_pmc.md"
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


def bench_adjoint(args):
    """Differentiable workload: gradient of the A-scan misfit w.r.t. per-element
    stiffness scale and density; reports t(forward + adjoint incl. recompute and
    gradient) / t(plain forward) and a central-difference check."""
    import __graft_entry__ as ge
    ge.load_package()
    from bristolfe_v2_amd import adjoint as adj
    from bristolfe_v2_amd import configs
    kw = {} if args.time_steps is None else {"n_t": args.time_steps}
    t0 = time.perf_counter()
    c = configs.adjoint_plate(**kw)
    pm, shots = c["model"], c["shots"]
    log(f"workload {c['name']}: n={pm.n} elements={pm.n_el} setup {time.perf_counter() - t0:.1f} s")
    data = adj.forward(pm, shots, c["s_true"], c["rho_true"])
    s0, r0 = np.ones(pm.n_el), pm.rho0.copy()
    # plain forward (no checkpoints), all transmitters
    plan = pm.plan(shots.dt, s0, r0)
    t_plain = []
    for rep in range(args.warmup + args.steps):
        ms = 0.0
        for tx in shots.tx:
            plan.run(shots.n_t, tx, shots.signal, shots.rx)
            ms += plan.stats()["loop_ms"]
        if rep >= args.warmup:
            t_plain.append(ms)
    plan.close()
    t_fwd, t_adj = [], []
    for rep in range(args.warmup + args.steps):
        A, plan, ms_f = adj.forward(pm, shots, s0, r0, keep=True)
        gk, gm, ms_a = adj.backward(plan, A - data)
        plan.close()
        if rep >= args.warmup:
            t_fwd.append(ms_f)
            t_adj.append(ms_a)
        log(f"rep {rep}: forward(ckpt) {ms_f:.1f} ms, adjoint {ms_a:.1f} ms")
    J0 = 0.5 * np.sum((A - data) ** 2)
    tp, tf, ta = float(np.mean(t_plain)), float(np.mean(t_fwd)), float(np.mean(t_adj))
    # central differences on seeded elements (GPU FP64 forward, all transmitters)
    rng = np.random.default_rng(2024)
    cand = np.nonzero(np.abs(gk) > 1e-3 * np.abs(gk).max())[0]
    pick = rng.choice(cand, min(args.fd_elements, cand.size), replace=False)
    errs_s, errs_r = [], []
    for e in pick:
        for which in ("s", "rho"):
            base = s0 if which == "s" else r0
            h = args.fd_step * abs(base[e])
            vals = []
            for sgn in (1, -1):
                th = base.copy()
                th[e] += sgn * h
                Ah = adj.ascans_batched(pm, shots, th if which == "s" else s0,
                                        th if which == "rho" else r0)
                vals.append(0.5 * np.sum((Ah - data) ** 2))
            fd = (vals[0] - vals[1]) / (2 * h)
            g = gk[e] if which == "s" else gm[e]
            (errs_s if which == "s" else errs_r).append(abs(g - fd) / max(abs(fd), 1e-300))
        log(f"fd element {e}: rel err s {errs_s[-1]:.2e} rho {errs_r[-1]:.2e}")
    out = {
        "metric": "adjoint/fwd cost", "value": round((tf + ta) / tp, 4), "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(tf + ta, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": c["name"], "n_dof": int(pm.n), "parameters": 2 * int(pm.n_el),
                   "transmitters": len(shots.tx), "time_steps": shots.n_t,
                   "checkpoint_every": shots.ckpt_every,
                   "step": "one gradient: forward of all transmitters + adjoint"},
        "forward_plain_ms": round(tp, 3), "forward_checkpointed_ms": round(tf, 3),
        "adjoint_ms": round(ta, 3), "misfit": J0,
        "fd_check": {"elements": int(len(pick)), "step_rel": args.fd_step,
                     "max_rel_err_stiffness": float(np.max(errs_s)) if errs_s else None,
                     "max_rel_err_density": float(np.max(errs_r)) if errs_r else None,
                     "median_rel_err": float(np.median(errs_s + errs_r)) if errs_s else None},
    }
    print(json.dumps(out), flush=True)


def bench_fmc(args):
    """Batched multi-source workload: 64 FMC transmitters on the plate plus 8
    subdomain models in one 64-column march, against 64 single-source runs."""
    import __graft_entry__ as ge
    ge.load_package()
    from bristolfe_v2_amd import _lib, autotune, configs
    from bristolfe_v2_amd.solver import make_plan
    from oracle import v6 as ov6
    kw = {} if args.time_steps is None else {"n_t": args.time_steps}
    t0 = time.perf_counter()
    c = configs.fmc_multimodel(**kw)
    K, C, M = c["K"], c["C"], c["M"]
    n, S, n_t = K.shape[0], c["n_src"], len(c["time"])
    fd, fs, hd, F = c["frc_dof"], c["frc_src"], c["hist_dof"], c["frcs"]
    log(f"workload {c['name']}: n={n} nnz={K.nnz} models={c['model_sizes']} "
        f"setup {time.perf_counter() - t0:.1f} s")
    res = {}
    for prec in ("double", "single"):
        plan = make_plan(K, C, M, c["dt"], "vel at curent time step", prec)
        if args.fmc_shape:  # fixed launch shape (short profiled runs): no autotune
            r_, w_, x_, g_, tile_ = (int(v) for v in args.fmc_shape.split(","))
            for key, v in ((_lib.TUNE_MULTI_ROWS_PER_BLOCK, r_), (_lib.TUNE_MULTI_ROWS_PER_WAVE, w_),
                           (_lib.TUNE_XCD_REMAP, x_), (_lib.TUNE_MULTI_GROUP, g_)):
                plan.set_tuning(key, v)
            best, table = dict(rows=r_, rpw=w_, remap=x_, group=g_, tile=tile_), []
        else:
            best, table = autotune.tune_batched(plan, S, fd, fs, F, hd)
        log(f"{prec}: autotune best {best}")
        walls, loops = [], []
        hb = None
        for rep in range(args.warmup + args.steps):
            w0 = time.perf_counter()
            hb, ms = autotune.run_tiled(plan, n_t, S, fd, fs, F, hd, best["tile"])
            w = time.perf_counter() - w0
            if rep >= args.warmup:
                walls.append(w)
                loops.append(ms)
        res[prec] = dict(plan=plan, best=best, table=table, wall=float(np.mean(walls)),
                         loop_ms=float(np.mean(loops)), hist=hb)
        log(f"{prec}: batched {S} columns: loop {np.mean(loops):.1f} ms, wall {np.mean(walls):.2f} s")
    # 64 plain single-source launches (FP64), forcing upload and history download included
    plan = res["double"]["plan"]
    w0 = time.perf_counter()
    plain_loop = 0.0
    worst = 0.0
    for s in range(S):
        sel = fs == s
        h, _, _ = plan.run(n_t, fd[sel], F, hd)
        plain_loop += plan.stats()["loop_ms"]
        b = res["double"]["hist"][:, s, :]
        worst = max(worst, float(np.linalg.norm(h - b) / max(np.linalg.norm(h), 1e-300)))
    plain_wall = time.perf_counter() - w0
    log(f"plain: {S} runs wall {plain_wall:.2f} s loop {plain_loop:.1f} ms; worst column diff {worst:.2e}")
    h64, h32 = res["double"]["hist"], res["single"]["hist"]
    err32 = float(np.linalg.norm(h32 - h64) / np.linalg.norm(h64))
    # oracle spot check on the plate block for two transmitters
    n0 = c["model_sizes"][0]
    K0, C0, M0 = K[:n0, :n0], C[:n0, :n0], M[:n0, :n0]
    q0 = np.nonzero(hd < n0)[0]
    orc = []
    for s in (0, S // 2):
        sel = (fs == s) & (fd < n0)
        r, _, _, _ = ov6.explicit_dynamic_solver_v6(K0, C0, M0, c["time"], fd[sel] + 1, F, hd[q0] + 1)
        orc.append(float(np.linalg.norm(h64[q0, s, :] - r) / np.linalg.norm(r)))
    sub = subdomain_ops(c, h64, K, C, M)
    info = plan.info()
    speedup = plain_wall / res["double"]["wall"]
    bpd = (12.0 * info["nnz_k"] / n + 4.0) / S + 24.0 + (16.0 / S if info["damped"] else 0.0)
    gdofs = n * S * (n_t - 1) / (res["double"]["loop_ms"] / 1e3) / 1e9
    out = {
        "metric": "batched per-source speedup", "value": round(speedup, 3), "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["double"]["wall"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": c["name"], "n_dof": int(n), "models": c["model_sizes"],
                   "columns": int(S), "time_steps": int(n_t),
                   "step": "one 64-column march of all models (bfe_run_batched)"},
        "batched_gdof_steps_per_s": round(gdofs, 3),
        "roofline": {"bound": "hbm", "achieved": round(bpd * n * S * (n_t - 1) / (res["double"]["loop_ms"] / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bpd * n * S * (n_t - 1) / (res["double"]["loop_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("fmc_batched", "k_step_multi"),
                     "kernel": (f"k_step_multi<double> ({S} sources, {res['double']['best']['rpw']} rows "
                                f"per wavefront, {res['double']['best']['group']} rows per gather)"),
                     "bytes_per_dof_step_source": round(bpd, 2)},
        "plain_wall_s": round(plain_wall, 3), "batched_wall_s": round(res["double"]["wall"], 3),
        "plain_loop_ms": round(plain_loop, 1), "batched_loop_ms": round(res["double"]["loop_ms"], 1),
        "speedup_loop_only": round(plain_loop / res["double"]["loop_ms"], 3),
        "batched_vs_single_max_rel_l2": worst,
        "oracle_rel_l2_plate_block": orc,
        "fp32": {"rel_l2_vs_fp64": err32, "wall_s": round(res["single"]["wall"], 3),
                 "speedup_vs_fp64_batched": round(res["double"]["wall"] / res["single"]["wall"], 3),
                 "autotune": res["single"]["best"]},
        "autotune": {"chosen": res["double"]["best"], "candidates": res["double"]["table"]},
        "subdomain_ops": sub,
    }
    for r in res.values():
        r["plan"].close()
    print(json.dumps(out), flush=True)


def subdomain_ops(c, h64, K, C, M):
    """The operations around the batched march, timed wall-clock with their
    host transfers (bfe_convolve, bfe_disps_to_forces) next to the oracle's
    numpy/scipy versions on the same data: the 64 x 64 pristine FMC A-scans
    convolved with the input signal, and the boundary forces of a four-layer
    band of the plate over the whole time axis."""
    import __graft_entry__ as ge
    pkg = ge.load_package()
    from oracle import subdoms as osd
    S, n_t = h64.shape[1], h64.shape[2]
    sig = c["frcs"][0]
    ascans = np.ascontiguousarray(h64[:S].transpose(1, 0, 2).reshape(S * S, n_t))  # (tx, rx) rows
    pkg.subdoms.fn_convolve(ascans[:2], sig, 2)   # plan and code-object warm-up
    w0 = time.perf_counter()
    fmc = pkg.subdoms.fn_convolve(ascans, sig, 2)
    conv_s = time.perf_counter() - w0
    w0 = time.perf_counter()
    ref = osd.convolve(ascans, sig[None, :], 2)
    conv_cpu_s = time.perf_counter() - w0
    conv_err = float(np.abs(fmc - ref).max() / np.abs(ref).max())
    # boundary band: the plate's DOFs on four node columns (the mesher numbers
    # nodes up each column, 101 to a column of the 10 mm plate at 0.1 mm),
    # layers 1..4
    n0 = c["model_sizes"][0]
    per_col = 2 * 101
    band = np.arange(40 * per_col, 44 * per_col)
    lyrs = 1 + (band - band[0]) // per_col
    Kb, Cb, Mb = K[band][:, band], C[band][:, band], M[band][:, band]
    assert band.max() < n0
    rng = np.random.default_rng(3)
    disp = rng.standard_normal((band.size, n_t)) * 1e-12
    pkg.subdoms.fn_convert_disps_to_forces_v2(Kb, Cb, Mb, c["dt"], disp[:, :4], lyrs, "in",
                                              "vel at curent time step")
    w0 = time.perf_counter()
    f, fs, _, _ = pkg.subdoms.fn_convert_disps_to_forces_v2(Kb, Cb, Mb, c["dt"], disp, lyrs, "in",
                                                            "vel at curent time step")
    frc_s = time.perf_counter() - w0
    w0 = time.perf_counter()
    fo, _, _, _ = osd.disps_to_forces(Kb, Cb, Mb, c["dt"], disp, lyrs, "in",
                                      "vel at curent time step")
    frc_cpu_s = time.perf_counter() - w0
    frc_err = float(np.abs(f - fo).max() / np.abs(fo).max())
    log(f"subdomain ops: convolve {S * S} x {n_t} in {conv_s:.3f} s (numpy {conv_cpu_s:.3f} s), "
        f"forces {f.shape} in {frc_s:.3f} s (scipy {frc_cpu_s:.3f} s)")
    return {"fmc_convolve": {"signals": S * S, "samples": n_t, "wall_s": round(conv_s, 4),
                             "numpy_fft_wall_s": round(conv_cpu_s, 4), "max_rel_err": conv_err},
            "disps_to_forces": {"dofs": int(band.size), "force_rows": int(fs.sum()),
                                "samples": n_t, "wall_s": round(frc_s, 4),
                                "scipy_wall_s": round(frc_cpu_s, 4), "max_rel_err": frc_err}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="coupled_2d")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--operator", choices=("auto", "csr", "ebe"), default="auto",
                    help="stiffness operator of the march (auto: element-by-element when the "
                         "mesh has the template form)")
    ap.add_argument("--timing-every", type=int, default=16)
    ap.add_argument("--time-steps", type=int, default=None,
                    help="override the workload's number of time steps (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 2 GiB device copy (counter calibration)")
    ap.add_argument("--fmc-shape", default=None,
                    help="fmc_batched: fixed launch shape rows,rpw,remap,group,tile instead of the autotuner")
    ap.add_argument("--fd-elements", type=int, default=50)
    ap.add_argument("--fd-step", type=float, default=1e-5)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus == 1 and args.workload == "adjoint_2d":
        bench_adjoint(args)
        return
    if args.gpus == 1 and args.workload == "fmc_batched":
        bench_fmc(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the march is single-GPU; --gpus {args.gpus} "
              "is not supported)", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    c = build_workload(args.workload, args.time_steps)
    log(f"workload {c['name']}: n={c['K'].shape[0]} nnz={c['K'].nnz} n_t={len(c['time'])} "
        f"setup {time.perf_counter() - t_setup:.1f} s")
    from bristolfe_v2_amd import ebe
    from bristolfe_v2_amd.solver import make_plan

    K, C, M = c["K"], c["C"], c["M"]
    n = K.shape[0]
    n_t = len(c["time"])
    op = None
    if args.operator != "csr":
        t0 = time.perf_counter()
        op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], n)
        nt = 0 if op is None else sum(g.tmpl.shape[0] for g in op.groups)
        log(f"element operator: {'none' if op is None else f'{nt} templates'} "
            f"({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    plan = make_plan(K, C, M, c["dt"], element_op=op, operator=args.operator)
    info = plan.info()
    log(f"plan: {info} ({time.perf_counter() - t0:.1f} s)")
    use_ebe = info["operator_kind"] == 1
    fi = np.asarray(c["frc_gi"]) - 1
    hi = np.asarray(c["hist_gi"]) - 1
    F = c["frcs"]

    # field frames of the workload (the plate: every 100 steps), reduced to the
    # per-element 'KE' field on the device next to the march
    fe_every = c["fe_options"].get("field_output_every_n_frames", float("inf"))
    fe_every = int(fe_every) if np.isfinite(fe_every) else 0
    if fe_every:
        plan.set_field_elements(c["mod"]["els"], c["gl_lookup"])

    def march():
        if fe_every:
            h, _, ti = plan.run_field_ke(n_t, fi, F, hi, fe_every)
            return h, None, ti
        return plan.run(n_t, fi, F, hi)

    for w in range(args.warmup):
        march()
        log(f"warmup {w + 1}/{args.warmup}: loop {plan.stats()['loop_ms']:.1f} ms")
    plan.set_kernel_timing(args.timing_every)
    # bfe_run returns after its stream has drained (hipEventSynchronize), so
    # the wall clock below is bracketed by device synchronisation.
    wall0 = time.perf_counter()
    loops, kern, samples, steps_done = [], 0.0, 0, 0
    hist = None
    for s in range(args.steps):
        hist, _, ti = march()
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
    pattern = use_ebe and info["ebe_pattern"] != 0
    kbase = "k_step_ebe_pat" if pattern else ("k_step_ebe" if use_ebe else "k_step")
    if use_ebe:
        bpd = ebe.ebe_bytes_per_dof(op, n, info["damped"])
        dmp = "true" if info["damped"] else "false"
        if pattern:
            pname = {1: "PatKuhn24", 2: "PatTriRect", 3: "PatTriIso"}[info["ebe_pattern"]]
            kname = " + ".join(f"k_step_ebe_pat<double,{dmp},{pname},{g.dpn}>" for g in op.groups)
            kname += " (element-by-element, neighbour u in registers, + central-difference update)"
        else:
            kname = " + ".join(f"k_step_ebe<double,{dmp},{g.el_nodes.shape[1]},{g.dpn},8>"
                               for g in op.groups)
            kname += " (element-by-element gather + central-difference update)"
    else:
        bpd = algorithmic_bytes_per_dof(info["nnz_k"], n, info["damped"])
        kname = "k_step<double,true> (SELL-64 SpMV + central-difference update)"
    kern_avg_s = kern / max(samples, 1) / 1e3
    achieved = bpd * n / kern_avg_s / 1e9 if samples else None
    traffic = read_traffic(args.workload, kbase)
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
                   "field_frames_every": fe_every or None,
                   "step": ("one full march of all time steps (bfe_run)" if not fe_every else
                            "one full march of all time steps with the per-element KE field "
                            "every %d steps (bfe_run_field_ke)" % fe_every)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     "kernel": kname,
                     "bytes_per_dof_step": round(bpd, 2),
                     "kernel_avg_us": round(kern_avg_s * 1e6, 2), "kernel_samples": int(samples)},
        "loop_gdof_steps_per_s_wall": round(n * steps_done * args.steps / wall / 1e9, 4),
    }
    # the same march with its time loop captured in one hipGraph (loop time
    # from HIP events around the graph launch; wall includes capture and
    # instantiation)
    plan.set_tuning(7, 1)
    w0 = time.perf_counter()
    hg, _, _ = plan.run(n_t, fi, F, hi)
    gw = time.perf_counter() - w0
    gst = plan.stats()
    plan.set_tuning(7, 0)
    out["graph"] = {"gdof_steps_per_s": round(n * gst["steps"] / gst["loop_ms"] / 1e6, 4),
                    "loop_ms": round(gst["loop_ms"], 3), "wall_s": round(gw, 3),
                    "eager_loop_ms": round(float(np.mean(loops)), 3),
                    "identical_to_eager": bool(np.array_equal(hg, hist))}
    if not use_ebe and n <= 2_000_000:
        # the same march (field reduction included) as one cooperative launch
        # of the persistent multi-step kernel
        plan.set_tuning(9, 1)
        pl, hp = [], None
        for _ in range(max(args.steps, 1)):
            hp, _, _ = march()
            pl.append(plan.stats()["loop_ms"])
        plan.set_tuning(9, 0)
        out["persistent"] = {"gdof_steps_per_s": round(n * steps_done / float(np.mean(pl)) / 1e6, 4),
                             "loop_ms": round(float(np.mean(pl)), 3),
                             "identical_to_eager": bool(np.array_equal(hp, hist))}
    if use_ebe:
        out["config"]["operator"] = (f"element-by-element: {op.n_el} elements in "
                                     f"{info['ebe_groups']} group(s), {info['ebe_templates']} "
                                     f"templates, {info['ebe_slots']} slots")
        # the CSR operator of the same plan, one march, for comparison
        plan.set_tuning(3, 0)
        plan.set_kernel_timing(args.timing_every)
        hc, _, _ = plan.run(n_t, fi, F, hi)
        st = plan.stats()
        bc = algorithmic_bytes_per_dof(info["nnz_k"], n, info["damped"])
        kc = st["step_kernel_ms"] / max(st["step_kernel_samples"], 1) / 1e3
        out["csr_operator"] = {
            "gdof_steps_per_s": round(n * st["steps"] / st["loop_ms"] / 1e6, 4),
            "kernel_avg_us": round(kc * 1e6, 2), "bytes_per_dof_step": round(bc, 2),
            "achieved_GBs": round(bc * n / kc / 1e9, 1),
            "history_rel_diff_vs_ebe": float(np.linalg.norm(hc - hist) / max(np.linalg.norm(hc), 1e-300))}
        plan.set_tuning(3, 1)
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
