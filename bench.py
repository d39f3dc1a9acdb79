"""Benchmark of the explicit FP64 time-march on one MI355X.

Default run (no --workload): the headline line is the largest single-GPU
single-source workload, ``cube_3d`` — a 20 mm steel cube of 48M linear
tetrahedra (24,361,803 DOF), absorbing layers on its four side faces, a
toneburst on a 3 x 3-node patch, 16 monitors, 2000 time steps, FP64. One
bench "step" is one full march of all time steps over the resident operator
(bfe_run); inputs are resident in HBM when the timed region starts.

  value     GDOF-steps/s = n_dof * marched steps / loop time, the loop time
            from HIP events on the march's stream (the v6:162-220 interval)
  roofline  the dominant kernel (the element-by-element pattern kernel when
            the mesh has the template form, else the SELL-64 CSR step):
            algorithmic bytes per DOF-step x n_dof / its HIP-event sampled
            duration; `traffic` and `traffic_frac` from the committed
            rocprofv3 PMC summary of that kernel and workload
            (profiles/<kernel>_<workload>_pmc.md)
  cpu_baseline          oracle/march.c (C/OpenMP, the fused alpha/beta/gamma
            march) on this host's CPU share, on the first steps of the same
            workload
  cpu_baseline_literal  the literal v6 loop u = A f + B u1 + D u2 with A, B, D
            formed as v6:130-139 (oracle/v6.py, SciPy CSR products, one
            thread), on the first steps of the same workload

The same JSON line carries the other components of the metric, each timed in
this run: ``coupled_2d`` (the 7.7M-DOF aluminium/water model with interface
and absorbing layers), ``batched`` (64 full-matrix-capture sources plus 8
subdomain models in one launch against 64 single-source launches of the best
single-source operator; FP32 bulk error; autotuner pick) and ``adjoint``
(gradient of an A-scan misfit: cost against the forward, and a central
finite-difference check). ``--workload NAME`` runs one workload on its own
(cube_3d, coupled_2d, plate_2d, fmc_batched, adjoint_2d).
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


def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


def build_workload(name, n_t=None):
    pkg()
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
    """Kernels of known traffic, for calibrating rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters (run under rocprofv3 --pmc): a plain device copy
    (reads and writes `gib` GiB), a sum (reads `gib` GiB, writes 8 bytes per
    workgroup), a fill (writes `gib` GiB), and two row gathers in a seeded
    random order (index_select, every source row read once): 512-byte rows
    (64 FP64, a 64-column multivector's row) and 64-byte rows (8 FP64), each
    moving gib/2 GiB in and out plus the 8-byte indices; three dispatches
    each."""
    import torch
    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s = 0.0
    for _ in range(3):
        s += float(a.sum())
    for _ in range(3):
        b.fill_(2.0)
    torch.cuda.synchronize()
    g = torch.Generator(device="cpu").manual_seed(7)
    rows = {}
    for width in (64, 8):
        nr = n // 2 // width
        src = a[:nr * width].view(nr, width)
        idx = torch.randperm(nr, generator=g).to("cuda")
        out = b[:nr * width].view(nr, width)
        for _ in range(3):
            torch.index_select(src, 0, idx, out=out)
        torch.cuda.synchronize()
        rows[f"gather_{width * 8}B_rows"] = nr
    print(json.dumps({"probe_bytes": n * 8, "dispatches_each": 3, "sum": s, **rows}))


def algorithmic_bytes_per_dof(nnz, n, damped):
    """CSR step: 12 nnz/n (value + column) + 4 (slice offsets, amortised to
    a row pointer) + 24 (u_n, u_{n-1} read, u_{n+1} written) + 8 k (k = 3
    coefficient streams when damped, else 1)."""
    k = 3 if damped else 1
    return 12.0 * nnz / n + 4.0 + 24.0 + 8.0 * k


def read_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary for this workload, with the calibration uncertainty it states."""
    p = ROOT / "profiles" / f"{kernel}_{workload.split('@')[0]}_pmc.md"
    if not p.exists():
        return None, None
    txt = p.read_text()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    u = re.search(r"traffic_uncertainty:\s*([^\n]+)", txt)
    return (float(m.group(1)) if m else None), (u.group(1).strip() if u else None)


def cpu_threads():
    """Threads of the CPU baseline: the CPU share this process may run on
    (its affinity set, capped by OMP_NUM_THREADS when the host sets it), and
    the host's logical CPU count for the record."""
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    return max(share, 1), os.cpu_count() or 1


def cpu_baseline(c, hi, budget_s=15.0):
    """The oracle's C/OpenMP march on the first steps of the workload."""
    from oracle.march import OracleMarch
    threads, host = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    K, C, M = c["K"], c["C"], c["M"]
    n_t = len(c["time"])
    t0 = time.perf_counter()
    om = OracleMarch(K, C, M, c["dt"])
    setup = time.perf_counter() - t0
    fi = np.asarray(c["frc_gi"]) - 1
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
            "host_logical_cpus": host,
            "sample": (f"first {steps} of {n_t} time steps of the same workload: oracle/march.c, "
                       f"the fused alpha/beta/gamma CSR march (OpenMP, {threads} threads = this "
                       f"process's CPU share of {host} logical CPUs), loop time {el:.2f} s, "
                       f"setup {setup:.1f} s")}, hist, steps


def cpu_baseline_literal(c, hi, budget_s=15.0):
    """The literal v6 loop (oracle/v6.py: A, B, D formed as v6:130-139, then
    u = A f + B u1 + D u2 with SciPy CSR products, which run on one thread)
    on the first steps of the workload. Models whose C is diagonal only (no
    fluid-solid interface): A, B, D are formed row block by row block."""
    from oracle import v6 as ov6
    K, C, M = c["K"], c["C"], c["M"]
    n_t = len(c["time"])
    t0 = time.perf_counter()
    abd = ov6.abd_diagonal_x(K, C, M, c["dt"], "vel at curent time step")
    setup = time.perf_counter() - t0
    fi = np.asarray(c["frc_gi"])
    ts = ov6.ti_start_of(c["frcs"])
    t0 = time.perf_counter()
    ov6.explicit_dynamic_solver_v6(K, C, M, c["time"], fi, c["frcs"], hi + 1, abd=abd, t_end=ts + 1)
    per = (time.perf_counter() - t0) / 2
    steps = int(max(3, min(n_t, budget_s / max(per, 1e-9))))
    t0 = time.perf_counter()
    hist, _, _, ti = ov6.explicit_dynamic_solver_v6(K, C, M, c["time"], fi, c["frcs"], hi + 1,
                                                    abd=abd, t_end=steps)
    el = time.perf_counter() - t0
    done = steps - (ti - 1)
    return {"value": round(K.shape[0] * done / el / 1e9, 5), "unit": "GDOF-steps/s", "cores": 1,
            "kind": "port",
            "sample": (f"first {steps} of {n_t} time steps of the same workload: the literal v6 "
                       f"loop u = A f + B u1 + D u2 (oracle/v6.py, SciPy CSR products on one "
                       f"thread, the analogue of MATLAB's sparse mtimes at v6:185), loop time "
                       f"{el:.2f} s, A/B/D formation {setup:.1f} s")}, hist, steps


def kernel_names(op, info, use_ebe):
    pattern = use_ebe and info["ebe_pattern"] != 0
    kbase = "k_step_ebe_pat" if pattern else ("k_step_ebe" if use_ebe else "k_step")
    dmp = "true" if info["damped"] else "false"
    if pattern:
        pname = {1: "PatKuhn24", 2: "PatTriRect", 3: "PatTriIso"}[info["ebe_pattern"]]
        kname = " + ".join(f"k_step_ebe_pat<double,{dmp},{pname},{g.dpn}>" for g in op.groups)
        kname += " (element-by-element, neighbour u in registers, + central-difference update)"
    elif use_ebe:
        kname = " + ".join(f"k_step_ebe<double,{dmp},{g.el_nodes.shape[1]},{g.dpn},8>"
                           for g in op.groups)
        kname += " (element-by-element gather + central-difference update)"
    else:
        kname = f"k_step<double,1,{dmp}> (SELL-64 SpMV + central-difference update)"
    return kbase, kname


def bench_march(name, reps, warmup, operator="auto", timing_every=16, cpu=True, literal=False,
                csr_compare=True, graph=True, persistent=True, cpu_budget=15.0, n_t=None,
                tag=None):
    """One single-source workload: `reps` timed full marches after `warmup`
    untimed ones. Returns the result dict (headline form)."""
    t_setup = time.perf_counter()
    c = build_workload(name, n_t)
    build_s = time.perf_counter() - t_setup
    log(f"workload {c['name']}: n={c['K'].shape[0]} nnz={c['K'].nnz} n_t={len(c['time'])} "
        f"build {build_s:.1f} s")
    from bristolfe_v2_amd import ebe
    from bristolfe_v2_amd.solver import make_plan

    K, C, M = c["K"], c["C"], c["M"]
    n = K.shape[0]
    n_t = len(c["time"])
    op = None
    t0 = time.perf_counter()
    if operator != "csr":
        op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], n)
    t_op = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = make_plan(K, C, M, c["dt"], element_op=op, operator=operator)
    t_plan = time.perf_counter() - t0
    info = plan.info()
    log(f"plan: {info} (element operator {t_op:.1f} s, plan {t_plan:.1f} s)")
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

    for w in range(warmup):
        march()
        log(f"{name} warmup {w + 1}/{warmup}: loop {plan.stats()['loop_ms']:.1f} ms")
    plan.set_kernel_timing(timing_every)
    # bfe_run returns after its stream has drained (hipEventSynchronize), so
    # the wall clock below is bracketed by device synchronisation.
    wall0 = time.perf_counter()
    loops, kern, samples, steps_done = [], 0.0, 0, 0
    hist = None
    for s in range(reps):
        hist, _, ti = march()
        st = plan.stats()
        loops.append(st["loop_ms"])
        kern += st["step_kernel_ms"]
        samples += st["step_kernel_samples"]
        steps_done = st["steps"]
        if st["nonfinite"]:
            raise SystemExit(f"non-finite history values ({st['nonfinite']})")
        log(f"{name} step {s + 1}/{reps}: loop {st['loop_ms']:.1f} ms, "
            f"{n * st['steps'] / st['loop_ms'] / 1e6:.2f} GDOF-steps/s")
    wall = time.perf_counter() - wall0
    plan.set_kernel_timing(0)
    loop_s = float(np.mean(loops)) / 1e3
    value = n * steps_done / loop_s / 1e9
    kbase, kname = kernel_names(op, info, use_ebe)
    bpd = (ebe.ebe_bytes_per_dof(op, n, info["damped"]) if use_ebe
           else algorithmic_bytes_per_dof(info["nnz_k"], n, info["damped"]))
    kern_avg_s = kern / max(samples, 1) / 1e3
    achieved = bpd * n / kern_avg_s / 1e9 if samples else None
    traffic, unc = read_traffic(tag or name, kbase)
    roof = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "traffic_frac": (round(traffic / kern_avg_s / 1e9 / HBM_PEAK_GBS, 4)
                             if traffic and samples else None),
            "traffic_uncertainty": unc,
            "kernel": kname, "bytes_per_dof_step": round(bpd, 2),
            "kernel_avg_us": round(kern_avg_s * 1e6, 2), "kernel_samples": int(samples)}
    out = {
        "value": round(value, 4), "unit": "GDOF-steps/s",
        "ms_per_step": round(wall / max(reps, 1) * 1e3, 3),
        "config": {"workload": c["name"], "n_dof": int(n), "nnz_K": int(info["nnz_k"]),
                   "time_steps": int(n_t), "steps_marched": int(steps_done),
                   "interface_rows": int(info["n_interface"]), "monitors": int(len(hi)),
                   "field_frames_every": fe_every or None,
                   "step": ("one full march of all time steps (bfe_run)" if not fe_every else
                            "one full march of all time steps with the per-element KE field "
                            "every %d steps (bfe_run_field_ke)" % fe_every)},
        "roofline": roof,
        "loop_gdof_steps_per_s_wall": round(n * steps_done * reps / wall / 1e9, 4),
        "setup_s": {"model_build": round(build_s, 2), "element_operator": round(t_op, 2),
                    "plan": round(t_plan, 2)},
    }
    if use_ebe:
        out["config"]["operator"] = (f"element-by-element: {op.n_el} elements in "
                                     f"{info['ebe_groups']} group(s), {info['ebe_templates']} "
                                     f"templates, {info['ebe_slots']} slots")
    if graph:
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
    if persistent and not use_ebe and n <= 2_000_000:
        # the same march (field reduction included) as one cooperative launch
        # of the persistent multi-step kernel
        plan.set_tuning(9, 1)
        pl, hp = [], None
        for _ in range(max(min(reps, 3), 1)):
            hp, _, _ = march()
            pl.append(plan.stats()["loop_ms"])
        plan.set_tuning(9, 0)
        out["persistent"] = {"gdof_steps_per_s": round(n * steps_done / float(np.mean(pl)) / 1e6, 4),
                             "loop_ms": round(float(np.mean(pl)), 3),
                             "identical_to_eager": bool(np.array_equal(hp, hist))}
    if use_ebe and csr_compare:
        # the CSR operator of the same plan, one march, for comparison
        plan.set_tuning(3, 0)
        plan.set_kernel_timing(timing_every)
        hc, _, _ = plan.run(n_t, fi, F, hi)
        st = plan.stats()
        plan.set_kernel_timing(0)
        bc = algorithmic_bytes_per_dof(info["nnz_k"], n, info["damped"])
        kc = st["step_kernel_ms"] / max(st["step_kernel_samples"], 1) / 1e3
        out["csr_operator"] = {
            "gdof_steps_per_s": round(n * st["steps"] / st["loop_ms"] / 1e6, 4),
            "kernel_avg_us": round(kc * 1e6, 2), "bytes_per_dof_step": round(bc, 2),
            "achieved_GBs": round(bc * n / kc / 1e9, 1),
            "history_rel_diff_vs_ebe": float(np.linalg.norm(hc - hist) / max(np.linalg.norm(hc), 1e-300))}
        plan.set_tuning(3, 1)
    hchk = np.concatenate([fi, hi])  # check monitors: the forcing DOFs and the monitors
    if cpu:
        log(f"{name}: cpu baseline ...")
        cb, hcpu, sample_steps = cpu_baseline(c, hchk, cpu_budget)
        out["cpu_baseline"] = cb
        hg, _, _ = plan.run(sample_steps, fi, F[:, :sample_steps], hchk)
        b = hcpu[:, :sample_steps]
        out["cpu_gpu_rel_l2_on_sample"] = float(np.linalg.norm(hg - b) / max(np.linalg.norm(b), 1e-300))
    if literal:
        log(f"{name}: literal v6 cpu baseline ...")
        cl, hl, sample_steps = cpu_baseline_literal(c, hchk, cpu_budget)
        out["cpu_baseline_literal"] = cl
        hg, _, _ = plan.run(sample_steps, fi, F[:, :sample_steps], hchk)
        b = hl[:, :sample_steps]
        out["cpu_baseline_literal"]["rel_l2_vs_gpu_on_sample"] = float(
            np.linalg.norm(hg - b) / max(np.linalg.norm(b), 1e-300))
    plan.close()
    return out


def bench_fmc(reps=2, warmup=1, autotune_full=False, shape=None, n_t=None):
    """Batched multi-source workload: 64 FMC transmitters on the plate plus 8
    subdomain models in one 64-column march (bfe_run_batched), against 64
    single-source launches (bfe_run) with the best single-source operator."""
    pkg()
    from bristolfe_v2_amd import _lib, autotune, configs, ebe
    from bristolfe_v2_amd.solver import make_plan
    from oracle import v6 as ov6
    t0 = time.perf_counter()
    c = configs.fmc_multimodel(**({} if n_t is None else {"n_t": n_t}))
    K, C, M = c["K"], c["C"], c["M"]
    n, S, n_t = K.shape[0], c["n_src"], len(c["time"])
    fd, fs, hd, F = c["frc_dof"], c["frc_src"], c["hist_dof"], c["frcs"]
    build_s = time.perf_counter() - t0
    log(f"workload {c['name']}: n={n} nnz={K.nnz} models={c['model_sizes']} build {build_s:.1f} s")
    t0 = time.perf_counter()
    op = ebe.element_operator(c["mod"], c["matls"], c["fe_options"], c["gl_lookup"], n)
    t_op = time.perf_counter() - t0
    res = {}
    for prec in ("double", "single"):
        plan = make_plan(K, C, M, c["dt"], "vel at curent time step", prec, element_op=op)
        if shape is not None or prec == "single":
            sh = shape if shape is not None else res["double"]["best"]
            autotune.apply(plan, sh)
            best, table = dict(sh), []
        elif autotune_full:
            best, table = autotune.tune_batched(plan, S, fd, fs, F, hd)
        else:  # a reduced search: the launch-shape and group stages, one repeat
            best, table = autotune.tune_batched(plan, S, fd, fs, F, hd, rows=(1, 2, 4),
                                                rpws=(2, 4, 8), repeats=1, tiles=(64,))
        log(f"{prec}: launch shape {best}")
        walls, loops, hb = [], [], None
        for rep in range(warmup + reps):
            w0 = time.perf_counter()
            hb, ms = autotune.run_tiled(plan, n_t, S, fd, fs, F, hd, best["tile"])
            w = time.perf_counter() - w0
            if rep >= warmup:
                walls.append(w)
                loops.append(ms)
        res[prec] = dict(plan=plan, best=best, table=table, wall=float(np.mean(walls)),
                         loop_ms=float(np.mean(loops)), hist=hb, steps=plan.stats()["steps"])
        log(f"{prec}: batched {S} columns: loop {np.mean(loops):.1f} ms, wall {np.mean(walls):.2f} s")
    # both multi-source operators at the chosen shape, one full march each
    plan = res["double"]["plan"]
    ops = {}
    for name, sel in (("csr", _lib.OPERATOR_CSR), ("element_product", _lib.OPERATOR_EBE),
                      ("node_blocks", _lib.OPERATOR_NODE)):
        plan.set_tuning(_lib.TUNE_MULTI_OPERATOR, sel)
        if sel:
            plan.set_tuning(_lib.TUNE_MULTI_NODES_PER_WAVE, res["double"]["best"].get("npw") or 4)
        hb, ms = autotune.run_tiled(plan, n_t, S, fd, fs, F, hd, res["double"]["best"]["tile"])
        st = plan.stats()
        if plan.info()["multi_operator"] != sel:
            continue
        ops[name] = {"loop_ms": round(ms, 1), "ms_per_step": round(ms / max(st["steps"], 1), 4),
                     "rel_l2_vs_chosen": float(np.linalg.norm(hb - res["double"]["hist"]) /
                                               np.linalg.norm(res["double"]["hist"]))}
        log(f"operator {name}: {ops[name]}")
    autotune.apply(plan, res["double"]["best"])
    # 64 plain single-source launches (FP64) with the single-source march's
    # operator choice ("auto": element-by-element on this mesh), forcing upload
    # and history download included
    plan = res["double"]["plan"]
    sinfo = plan.info()
    single_op = "element-by-element" if sinfo["operator_kind"] == 1 else "CSR"
    w0 = time.perf_counter()
    plain_loop, worst = 0.0, 0.0
    for s in range(S):
        sel = fs == s
        h, _, _ = plan.run(n_t, fd[sel], F, hd)
        plain_loop += plan.stats()["loop_ms"]
        b = res["double"]["hist"][:, s, :]
        worst = max(worst, float(np.linalg.norm(h - b) / max(np.linalg.norm(h), 1e-300)))
    plain_wall = time.perf_counter() - w0
    log(f"plain ({single_op}): {S} runs wall {plain_wall:.2f} s loop {plain_loop:.1f} ms; "
        f"worst column diff {worst:.2e}")
    h64, h32 = res["double"]["hist"], res["single"]["hist"]
    err32 = float(np.linalg.norm(h32 - h64) / np.linalg.norm(h64))
    # oracle spot check on the plate block for one transmitter
    n0 = c["model_sizes"][0]
    K0, C0, M0 = K[:n0, :n0], C[:n0, :n0], M[:n0, :n0]
    q0 = np.nonzero(hd < n0)[0]
    s = S // 2
    sel = (fs == s) & (fd < n0)
    r, _, _, _ = ov6.explicit_dynamic_solver_v6(K0, C0, M0, c["time"], fd[sel] + 1, F, hd[q0] + 1)
    orc = float(np.linalg.norm(h64[q0, s, :] - r) / np.linalg.norm(r))
    info = plan.info()
    speedup = plain_wall / res["double"]["wall"]
    steps = res["double"]["steps"]
    chosen = res["double"]["best"].get("operator")
    if chosen == "ebe":
        # per DOF-step-source: the element data and coefficients once per step
        # for all S sources, u_n / u_{n-1} read and u_{n+1} written per source
        bpd = (ebe.ebe_bytes_per_dof(op, n, info["damped"]) - 24.0) / S + 24.0
    elif chosen == "node":
        # per node its 7 neighbour-class blocks (dpn x dpn), mask, alpha and
        # beta once per step for all S sources; u_n / u_{n-1} / u_{n+1} per source
        per_node = sum(g.n_nodes * (7 * g.dpn * g.dpn * 8 + 4 + (16 if info["damped"] else 8))
                       for g in op.groups)
        bpd = per_node / n / S + 24.0
    else:
        bpd = (12.0 * info["nnz_k"] / n + 4.0) / S + 24.0 + (16.0 / S if info["damped"] else 0.0)
    gb = bpd * n * S * steps / (res["double"]["loop_ms"] / 1e3) / 1e9
    traffic, unc = read_traffic("fmc_batched", "k_step_multi")
    out = {
        "value": round(speedup, 3), "unit": "x (per-source time, 64 single-source launches / "
                                            "one 64-column launch, wall clock)",
        "config": {"workload": c["name"], "n_dof": int(n), "models": c["model_sizes"],
                   "columns": int(S), "time_steps": int(n_t), "steps_marched": int(steps),
                   "single_source_operator": single_op},
        "batched_gdof_steps_per_s": round(n * S * steps / (res["double"]["loop_ms"] / 1e3) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(gb, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gb / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_uncertainty": unc,
                     "kernel": (f"k_step_ebe_multi<double,true,PatTriRect,2,true> ({S} sources, "
                                f"{res['double']['best']['npw']} nodes per wavefront)"
                                if chosen == "ebe" else
                                f"k_step_ebe_multi<double,true,PatTriRect,2,false,false,true> "
                                f"(node-block form, {S} sources, {res['double']['best']['npw']} "
                                f"nodes per wavefront)" if chosen == "node" else
                                f"k_step_multi<double> ({S} sources, {res['double']['best']['rpw']} "
                                f"rows per wavefront, {res['double']['best']['group']} rows per "
                                f"gather)"),
                     "bytes_per_dof_step_source": round(bpd, 2)},
        "plain_wall_s": round(plain_wall, 3), "batched_wall_s": round(res["double"]["wall"], 3),
        "plain_loop_ms": round(plain_loop, 1), "batched_loop_ms": round(res["double"]["loop_ms"], 1),
        "speedup_loop_only": round(plain_loop / res["double"]["loop_ms"], 3),
        "batched_vs_single_max_rel_l2": worst,
        "oracle_rel_l2_plate_block": orc,
        "fp32": {"rel_l2_vs_fp64": err32, "wall_s": round(res["single"]["wall"], 3),
                 "loop_ms": round(res["single"]["loop_ms"], 1),
                 "speedup_vs_fp64_batched": round(res["double"]["wall"] / res["single"]["wall"], 3),
                 "launch_shape": res["single"]["best"]},
        "autotune": {"chosen": res["double"]["best"], "candidates": len(res["double"]["table"]),
                     "search": ("full" if autotune_full else
                                "reduced (shape, group and element-product stages)")},
        "multi_source_operators": ops,
        "setup_s": {"model_build": round(build_s, 2), "element_operator": round(t_op, 2)},
        "subdomain_ops": subdomain_ops(c, h64, K, C, M),
    }
    for r_ in res.values():
        r_["plan"].close()
    return out


def subdomain_ops(c, h64, K, C, M):
    """The operations around the batched march, timed wall-clock with their
    host transfers (bfe_convolve, bfe_disps_to_forces) next to the oracle's
    numpy/scipy versions on the same data: the 64 x 64 pristine FMC A-scans
    convolved with the input signal, and the boundary forces of a four-layer
    band of the plate over the whole time axis."""
    p = pkg()
    from oracle import subdoms as osd
    S, n_t = h64.shape[1], h64.shape[2]
    sig = c["frcs"][0]
    ascans = np.ascontiguousarray(h64[:S].transpose(1, 0, 2).reshape(S * S, n_t))  # (tx, rx) rows
    p.subdoms.fn_convolve(ascans[:2], sig, 2)   # plan and code-object warm-up
    w0 = time.perf_counter()
    fmc = p.subdoms.fn_convolve(ascans, sig, 2)
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
    p.subdoms.fn_convert_disps_to_forces_v2(Kb, Cb, Mb, c["dt"], disp[:, :4], lyrs, "in",
                                            "vel at curent time step")
    w0 = time.perf_counter()
    f, fs, _, _ = p.subdoms.fn_convert_disps_to_forces_v2(Kb, Cb, Mb, c["dt"], disp, lyrs, "in",
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


def bench_adjoint(reps=2, warmup=1, fd_elements=4, fd_step=1e-5, n_t=None):
    """Differentiable workload: gradient of the A-scan misfit w.r.t.
    per-element stiffness scale and density. Reports t(checkpointed forward +
    adjoint incl. recompute and gradient) against t(plain forward), the
    gradient batched over the transmitters (the value, against the batched
    forward) and per transmitter (against 16 single-source forwards), and a
    central-difference check on seeded elements."""
    pkg()
    from bristolfe_v2_amd import adjoint as adj
    from bristolfe_v2_amd import configs
    kw = {} if n_t is None else {"n_t": n_t}
    t0 = time.perf_counter()
    c = configs.adjoint_plate(**kw)
    pm, shots = c["model"], c["shots"]
    log(f"workload {c['name']}: n={pm.n} elements={pm.n_el} build {time.perf_counter() - t0:.1f} s")
    data = adj.forward(pm, shots, c["s_true"], c["rho_true"])
    s0, r0 = np.ones(pm.n_el), pm.rho0.copy()
    # plain forward (no checkpoints): every transmitter on its own, and all
    # transmitters in one batched march
    plan = pm.plan(shots.dt, s0, r0)
    t_plain, t_batch = [], []
    fi = np.concatenate(shots.tx)
    fsrc = np.concatenate([np.full(t.size, k) for k, t in enumerate(shots.tx)])
    for rep in range(warmup + reps):
        ms = 0.0
        for tx in shots.tx:
            plan.run(shots.n_t, tx, shots.signal, shots.rx)
            ms += plan.stats()["loop_ms"]
        plan.run_batched(shots.n_t, len(shots.tx), fi, fsrc, shots.signal, shots.rx)
        if rep >= warmup:
            t_plain.append(ms)
            t_batch.append(plan.stats()["loop_ms"])
    plan.close()
    # the gradient: all transmitters in one checkpointed multi-source march
    # and one reverse sweep (batched), and one march and sweep per
    # transmitter
    t_fwd, t_adj, t_fwd1, t_adj1 = [], [], [], []
    for rep in range(warmup + reps):
        A, plan, ms_f = adj.forward_batched(pm, shots, s0, r0, keep=True)
        gk, gm, ms_a = adj.backward_batched(plan, A - data)
        plan.close()
        A1, plan, ms_f1 = adj.forward(pm, shots, s0, r0, keep=True)
        gk1, gm1, ms_a1 = adj.backward(plan, A1 - data)
        plan.close()
        if rep >= warmup:
            t_fwd.append(ms_f)
            t_adj.append(ms_a)
            t_fwd1.append(ms_f1)
            t_adj1.append(ms_a1)
        log(f"adjoint rep {rep}: batched forward(ckpt) {ms_f:.1f} ms, adjoint {ms_a:.1f} ms; "
            f"per transmitter {ms_f1:.1f} + {ms_a1:.1f} ms")
    J0 = 0.5 * np.sum((A - data) ** 2)
    tp, tb = float(np.mean(t_plain)), float(np.mean(t_batch))
    tf, ta = float(np.mean(t_fwd)), float(np.mean(t_adj))
    tf1, ta1 = float(np.mean(t_fwd1)), float(np.mean(t_adj1))
    g_agree = max(float(np.linalg.norm(gk - gk1) / np.linalg.norm(gk1)),
                  float(np.linalg.norm(gm - gm1) / np.linalg.norm(gm1)))
    # central differences on seeded elements (GPU FP64 batched forward)
    rng = np.random.default_rng(2024)
    cand = np.nonzero(np.abs(gk) > 1e-3 * np.abs(gk).max())[0]
    pick = rng.choice(cand, min(fd_elements, cand.size), replace=False)
    errs_s, errs_r = [], []
    for e in pick:
        for which in ("s", "rho"):
            base = s0 if which == "s" else r0
            h = fd_step * abs(base[e])
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
    return {
        "value": round((tf + ta) / tb, 4),
        "unit": ("x (batched checkpointed forward + batched adjoint) / (batched forward of all "
                 "transmitters, the fastest forward)"),
        "vs_sequential_forward": round((tf + ta) / tp, 4),
        "per_transmitter_gradient_vs_sequential_forward": round((tf1 + ta1) / tp, 4),
        "batched_vs_per_transmitter_gradient_rel_l2": g_agree,
        "config": {"workload": c["name"], "n_dof": int(pm.n), "parameters": 2 * int(pm.n_el),
                   "transmitters": len(shots.tx), "time_steps": shots.n_t,
                   "checkpoint_every": shots.ckpt_every,
                   "step": "one gradient: forward of all transmitters + adjoint"},
        "forward_plain_ms": round(tp, 3), "forward_batched_ms": round(tb, 3),
        "forward_checkpointed_ms": round(tf, 3), "adjoint_ms": round(ta, 3),
        "per_transmitter": {"forward_checkpointed_ms": round(tf1, 3), "adjoint_ms": round(ta1, 3)},
        "misfit": J0,
        "fd_check": {"elements": int(len(pick)), "step_rel": fd_step,
                     "max_rel_err_stiffness": float(np.max(errs_s)) if errs_s else None,
                     "max_rel_err_density": float(np.max(errs_r)) if errs_r else None,
                     "median_rel_err": float(np.median(errs_s + errs_r)) if errs_s else None},
    }


def line(body, args, metric=METRIC, higher=True):
    out = {"metric": metric, "value": body.pop("value"), "unit": body.pop("unit"), "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": body.pop("ms_per_step", None), "higher_is_better": higher,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    out.update(body)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=None,
                    help="one workload on its own (default: the cube_3d headline with the "
                         "coupled_2d, batched and adjoint components in the same line)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-components", action="store_true",
                    help="default run: the headline only")
    ap.add_argument("--operator", choices=("auto", "csr", "ebe"), default="auto",
                    help="stiffness operator of the march (auto: element-by-element when the "
                         "mesh has the template form)")
    ap.add_argument("--timing-every", type=int, default=16)
    ap.add_argument("--time-steps", type=int, default=None,
                    help="override the workload's number of time steps (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the counter-calibration kernels (2 GiB copy, sum, fill, row gathers)")
    ap.add_argument("--fmc-shape", default=None,
                    help="fmc_batched: fixed launch shape rows,rpw,remap,group,tile[,npw] instead "
                         "of the autotuner (npw > 0: the multi-source element product with npw "
                         "nodes per wavefront)")
    ap.add_argument("--fd-elements", type=int, default=None)
    ap.add_argument("--fd-step", type=float, default=1e-5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the march is single-GPU; --gpus {args.gpus} "
              "is not supported)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    wl = args.workload
    if wl == "adjoint_2d":
        body = bench_adjoint(args.steps, args.warmup, args.fd_elements or 50, args.fd_step,
                             args.time_steps)
        body["ms_per_step"] = round(body["forward_checkpointed_ms"] + body["adjoint_ms"], 3)
        print(json.dumps(line(body, args, "adjoint/fwd cost", higher=False)), flush=True)
        return
    if wl == "fmc_batched":
        shape = None
        if args.fmc_shape:
            v_ = [int(v) for v in args.fmc_shape.split(",")]
            r_, w_, x_, g_, tile_ = v_[:5]
            npw_ = v_[5] if len(v_) > 5 else 0
            shape = dict(rows=r_, rpw=w_, remap=x_, group=g_, tile=tile_,
                         operator="ebe" if npw_ > 0 else "csr", npw=npw_)
        body = bench_fmc(args.steps, args.warmup, autotune_full=True, shape=shape,
                         n_t=args.time_steps)
        body["ms_per_step"] = round(body["batched_wall_s"] * 1e3, 3)
        print(json.dumps(line(body, args, "batched per-source speedup")), flush=True)
        return
    if wl is not None:
        body = bench_march(wl, args.steps, args.warmup, args.operator, args.timing_every,
                           cpu=not args.no_cpu, literal=not args.no_cpu and wl == "cube_3d",
                           n_t=args.time_steps)
        print(json.dumps(line(body, args)), flush=True)
        return
    # default: the cube headline, then the other metric components
    t0 = time.perf_counter()
    body = bench_march("cube_3d", args.steps, args.warmup, args.operator, args.timing_every,
                       cpu=not args.no_cpu, literal=not args.no_cpu, n_t=args.time_steps)
    if not args.no_components:
        log(f"headline done at {time.perf_counter() - t0:.0f} s; coupled_2d ...")
        cb = bench_march("coupled_2d", 3, 1, args.operator, args.timing_every, cpu=False,
                         literal=False, graph=False, persistent=False)
        body["coupled_2d"] = {k: cb[k] for k in ("value", "unit", "config", "roofline",
                                                 "csr_operator", "setup_s") if k in cb}
        log(f"coupled_2d done at {time.perf_counter() - t0:.0f} s; batched ...")
        body["batched"] = bench_fmc(reps=2, warmup=1)
        log(f"batched done at {time.perf_counter() - t0:.0f} s; adjoint ...")
        body["adjoint"] = bench_adjoint(reps=2, warmup=1, fd_elements=args.fd_elements or 4,
                                        fd_step=args.fd_step)
        log(f"all components done at {time.perf_counter() - t0:.0f} s")
    print(json.dumps(line(body, args)), flush=True)


if __name__ == "__main__":
    main()
