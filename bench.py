#!/usr/bin/env python3
"""Throughput of one baroclinic step of the MOM5 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is the reference's order (ocean_model.F90:1441-2108, the part this
build implements): the tracer update (advection into th_tendency, explicit
update, vert_diffuse_implicit/invtri) for every tracer, then the barotropic
subcycle with its time-averaging tail.  Inputs are seeded synthetic fields
resident in HBM before timing starts.

--workload mixed (default): 3600 x 2700 x 75, T and S in FP64 plus 10 passive
tracers in FP32, 120 barotropic subcycles, advanced by the batched step
(mom_gpu_step_tracers + update_ocean_barotropic).  The same case is then run
through the plain per-tracer FP64 path for the same number of steps: that run
gives the FP64 rate the batched step is compared with, the bitwise check of
T and S, and the FP64 reference for each passive tracer's error.  Masked
tracer content (compensated FP64 sums) is reported before and after.

--workload batched_fp64: 1440 x 1080 x 50, 32 FP64 tracers through the batched
step.  --workload ensemble: 24 regional domains as one ensemble.  --workload
plain_fp64: 1440 x 1080 x 50 plain FP64 path (T, S MDFL-Sweby, one Quicker
tracer, 80 subcycles).  --workload small: 90 x 60 x 25, MDFL only.

Prints one JSON line: tracer-cell-updates/s of the step (value), the roofline
of the dominant kernel (HIP-event timed on the context's stream) and a CPU
baseline (the CPU restatement on a bounded sample of the workload, on one core
and on every usable core, with the host's CPU model and core count).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# vector ALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 FP32 (not
# packed) or FP64 instruction per SIMD every 4 cycles
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4

WORKLOADS = {
    "mixed": dict(name="mixed_3600x2700x75", ni=3600, nj=2700, nk=75, land_frac=0.30,
            nts=120, n_passive=10, mixed=True, passive_dtype=np.float32, steps=18, warmup=2,
            schemes=["mdfl_sweby"] * 12, sample=dict(ni=240, nj=180)),
    "batched_fp64": dict(name="batched_fp64_1440x1080x50", ni=1440, nj=1080, nk=50, land_frac=0.30,
            nts=80, n_passive=30, mixed=True, passive_dtype=np.float64, steps=20, warmup=2,
            schemes=["mdfl_sweby"] * 32, sample=dict(ni=240, nj=180)),
    "ensemble": dict(name="ensemble_24_regional", ndom=24, n_passive=6, ensemble=True, steps=20,
            warmup=2, sample=dict(ni=200, nj=150, nk=40, nts=80, regional=True)),
    "plain_fp64": dict(name="plain_fp64_1440x1080x50", ni=1440, nj=1080, nk=50, land_frac=0.30,
            nts=80, n_passive=1, schemes=["mdfl_sweby", "mdfl_sweby", "quicker"],
            sample=dict(ni=480, nj=360)),
    "small": dict(name="plain_fp64_90x60x25", ni=90, nj=60, nk=25, land_frac=0.25, nts=0,
            n_passive=0, schemes=["mdfl_sweby", "mdfl_sweby"], sample=dict(ni=90, nj=60)),
}


def make(w, ni=None, nj=None, seed=0x4D4F4D35):
    from mom5_amd.synth import make_case
    s = w.get("sample", {})
    if s.get("regional"):              # a closed regional basin like the ensemble's members
        return make_case(ni or s["ni"], nj or s["nj"], s["nk"], cyclic_x=False, lat_s=20.0,
                         lat_n=20.0 + 0.1 * (nj or s["nj"]), lon_span=0.1 * (ni or s["ni"]),
                         land_frac=0.25, nts=s["nts"], n_passive=w["n_passive"], seed=seed)
    return make_case(ni or w["ni"], nj or w["nj"], w["nk"], land_frac=w["land_frac"],
                     nts=w["nts"] or 80, n_passive=w["n_passive"], barotropic=w["nts"] > 0,
                     seed=seed)


def mdfl_bytes_per_launch(ni, nj, nk):
    """Algorithmic bytes of one MDFL launch: per cell T read, th_tendency
    read+write, uhrho_et/vhrho_nt/wrho_bt/rho_dzt read (8 B each); per column
    six FP64 metrics and int32 kmt."""
    return ni * nj * nk * (8 + 16 + 32) + ni * nj * 52


def host_cpu():
    """Model name (lscpu / /proc/cpuinfo), nproc and the cores this process
    may use (its affinity mask, capped by OMP_NUM_THREADS when set)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    nproc = os.cpu_count()
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        usable = min(usable, int(os.environ["OMP_NUM_THREADS"]))
    return dict(model=model, nproc=nproc, usable=max(1, usable))


def cpu_step_time(w, threads, steps=4, budget=10.0):
    """Median time of one whole step of the CPU restatement (serial build, or
    its all-core OpenMP build) on the bounded sample of the workload: at
    least `steps` steps after one warm-up and about `budget` s of timed work."""
    from oracle.oracle import Oracle
    s = w["sample"]
    case = make(w, ni=s["ni"], nj=s["nj"], seed=0x4D4F4D35 + 7)
    o = Oracle(case, threads=threads)
    f = case.fields
    Ts = [np.array(t, dtype=np.float64) for _, t in case.tracers]
    names = [n for n, _ in case.tracers]
    schemes = w.get("schemes") or ["mdfl_sweby"] * len(Ts)
    mins = [(-2.0 if n == "temp" else 1.0 if n == "salt" else -0.1) for n in names]
    maxs = [(32.0 if n == "temp" else 42.0 if n == "salt" else 1.0) for n in names]

    def step(Ts):
        # compute_tmask_limit only feeds Quicker (ocean_tracer_advect.F90:
        # 2609-2625, 3006-3111): formed when a Quicker tracer is present
        lim = o.tmask_limit(Ts, mins, maxs, 0, 1, True) if "quicker" in schemes else None
        out = []
        for n, T in enumerate(Ts):
            th = np.zeros_like(T)
            if schemes[n] == "quicker":
                th[:, 1:-1, 1:-1] -= o.horz_quicker(T, T, lim[n])
                th[:, 1:-1, 1:-1] -= o.vert_quicker(T, T, lim[n])
            else:
                o.horz_advect_mdfl(T, th, case.dt, lim=1.0)
            out.append(o.update_explicit(T, th, case.dt))
        o.vert_diffuse_implicit(out, f["stf"], f["btf"], case.dt, index_salt=1)
        if case.dtbt > 0:
            o.update_ocean_barotropic(case.nts, case.dtbt)
        return out

    Ts = step(Ts)                       # warm-up
    times = []
    while len(times) < steps or (sum(times) < budget and len(times) < 64):
        t0 = time.perf_counter()
        Ts = step(Ts)
        times.append(time.perf_counter() - t0)
    cells = case.ni * case.nj * case.nk * len(Ts)
    sample = (f"{case.ni}x{case.nj}x{case.nk}, {len(Ts)} tracers, full step incl. barotropic "
              f"nts={case.nts}; median of {len(times)} steps after 1 warm-up "
              f"({sum(times):.1f} s timed)")
    return cells / float(np.median(times)), sample


def cpu_baseline(w):
    """The CPU restatement timed on this host: one core (the serial build)
    and every usable core (the same source built with OpenMP, parallel over
    levels, rows and points; OMP_NUM_THREADS threads).  The all-core figure
    is the reported value."""
    host = host_cpu()
    n = host["usable"]
    os.environ["OMP_NUM_THREADS"] = str(n)   # read when the OpenMP build loads
    v1, s1 = cpu_step_time(w, threads=False)
    vn, sn = cpu_step_time(w, threads=True)
    return dict(value=vn, unit="tracer-cell-updates/s", cores=n, kind="port", sample=sn,
                single_core=dict(value=v1, cores=1, sample=s1),
                host=dict(model=host["model"], nproc=host["nproc"], threads_used=n))


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def step_kernel_bytes(ni, nj, nk, elem, group):
    """Algorithmic bytes of one fused tracer-step launch (advection +
    explicit update of `group` tracers): per cell uhrho_et, vhrho_nt,
    wrho_bt, rho_dzt, rho_dztr read (8 B each) and each tracer read and
    written once (elem B each way); per column the FP64 metrics and kmt."""
    return ni * nj * nk * (40 + 2 * elem * group) + ni * nj * 52


def vdiff_kernel_bytes(ni, nj, nk, elem, ntr):
    """Algorithmic bytes of one implicit-diffusion launch over `ntr`
    tracers (the invtri routine-boundary minimum): per cell each tracer read
    and written once and dcb, rho_dzt, dzwt read once (24 B); per column
    stf and btf of each tracer and kmt."""
    return ni * nj * nk * (2 * elem * ntr + 24) + ni * nj * (16 * ntr + 4)


def mixed_step_bytes(ni, nj, nk, n64, n32, barotropic=True):
    """Algorithmic bytes of one whole step, each distinct input read once and
    each output written once:
      C*(sum over tracers of 2*s_n + 8*S3) + C2*(52 + 16*N_tr) [+ C2*156],
    C = ni*nj*nk cells, C2 = ni*nj columns, s_n the tracer's storage size,
    S3 = 8 shared FP64 3-D arrays passed (uhrho_et, vhrho_nt, wrho_bt,
    rho_dzt, rho_dztr, dzwt, two diff_cbt planes), 52 B/column of metrics and
    kmt, 16 B/column/tracer of stf and btf, 156 B/column of barotropic inputs
    and outputs."""
    C, C2 = ni * nj * nk, ni * nj
    ntr = n64 + n32
    return (C * (16 * n64 + 8 * n32 + 8 * 8) + C2 * (52 + 16 * ntr)
            + (C2 * 156 if barotropic else 0))


def time_launches(g, fn, reps=10):
    fn()
    g.sync()
    g.mark(2)
    for _ in range(reps):
        fn()
    g.mark(3)
    return g.elapsed_ms(2, 3) / reps


def run_ensemble(args, w):
    """Config 4: 24 seeded regional domains (T, S FP64 + 6 FP32 passive
    each, own dt and subcycle count) advanced as one ensemble, timed from
    graph replays, against the same domains one after another on the plain
    per-tracer FP64 path."""
    import dataclasses
    import gc
    from mom5_amd.ensemble import EnsembleStep, make_regional_ensemble, regional_shapes
    from mom5_amd.step import BatchedStep, PlainStep
    steps = w["steps"] if args.steps is None else args.steps
    warm = w["warmup"] if args.warmup is None else args.warmup
    shapes = regional_shapes(w["ndom"])
    t0 = time.time()
    cases = make_regional_ensemble(shapes, n_passive=w["n_passive"])
    members = [BatchedStep(c, gpu=c.info["gpu"]) for c in cases]
    ens = EnsembleStep(members)
    g = ens.gpu
    ncells = ens.cells()
    ntc = ens.tracer_cells()
    log(f"generated {len(cases)} regional domains ({ncells / 1e6:.1f} M cells, "
        f"{ntc / 1e6:.0f} M tracer-cells) in {time.time() - t0:.1f}s")
    ens.run(warm)
    g.sync()
    g.mark(0)
    ens.run(2)
    g.mark(1)
    ms_eager = g.elapsed_ms(0, 1) / 2
    graph = ens.capture(2)
    g.mark(0)
    for _ in range(steps // 2):
        graph.launch()
    if steps % 2:
        ens.run(1)
    g.mark(1)
    ms = g.elapsed_ms(0, 1) / steps
    nodes = graph.nodes
    graph.close()
    warm += 2
    value = ntc / (ms / 1e3)
    log(f"ensemble step {ms:.2f} ms from graph replays ({nodes} nodes per 2 steps; eager "
        f"{ms_eager:.2f} ms) = {value:.3e} tracer-cell-updates/s")

    # the same domains one after another on the plain FP64 path
    total_plain_ms = 0.0
    ts_differ = 0
    worst = 0.0
    for k, c in enumerate(cases):
        m = members[k]
        mg = c.info["gpu"]
        pc = dataclasses.replace(c, tracers=[(nm, c.info["tracer"](i, np.float64))
                                             for i, (nm, _) in enumerate(c.tracers)])
        plain = PlainStep(pc, schemes=["mdfl_sweby"] * m.ntr, gpu=mg)
        plain.run(warm)
        mg.sync()
        mg.mark(0)
        plain.run(steps)
        mg.mark(1)
        total_plain_ms += mg.elapsed_ms(0, 1) / steps
        for n in range(m.ntr):
            e = mg.tracer_error(m.tracer(n), plain.tracer(n))
            if n < 2:
                ts_differ += e["cells_differ"]
            else:
                worst = max(worst, e["normwise"])
        del plain, pc
        gc.collect()
    value_plain = ntc / (total_plain_ms / 1e3)
    log(f"plain FP64, domains one after another: {total_plain_ms:.2f} ms per step "
        f"({value_plain:.3e}); T/S cells differing {ts_differ}; worst passive error {worst:.2e}")
    step_bytes = sum(mixed_step_bytes(c.ni, c.nj, c.nk, 2, len(c.tracers) - 2) for c in cases)
    achieved = step_bytes / (ms / 1e3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="whole ensemble step (all phases)", bytes_per_launch=step_bytes)
    extra = dict(timed_as="hipGraph replays of 2 captured ensemble steps",
                 graph_nodes_per_2_steps=nodes, eager_ms_per_step=ms_eager,
                 plain_fp64_sequential_value=value_plain,
                 plain_fp64_sequential_ms_per_step=total_plain_ms,
                 speedup_vs_plain_fp64_sequential=value / value_plain,
                 ts_cells_differing_from_plain_fp64=ts_differ,
                 passive_max_rel_err_normwise=worst, errors_after_steps=warm + steps)
    cfg = {"workload": w["name"], "domains": len(cases),
           "shapes": [[s["ni"], s["nj"], s["nk"], s["nts"]] for s in shapes],
           "cells": ncells, "tracers_per_domain": {"f64": 2, "f32": w["n_passive"]},
           "path": "ensemble (mom_gpu_ensemble_*: work-queue tracer passes)",
           "tuning": {k: g.get_tuning(k) for k in TUNING_KEYS}}
    return dict(value=value, ms=ms, steps=steps, warmup=warm, dtype="f64+f32", cfg=cfg,
                roof=roof, extra=extra, ntr=None)


def measured_traffic(workload, key):
    """HBM bytes per dispatch of kernel `key` on `workload`, from the
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs committed under profiles/
    (scaled by the copy-kernel calibration recorded there); None if that
    kernel and launch shape were not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_mixed_3600x2700x75.json")
    if workload is None or not os.path.exists(path):
        return None, None
    with open(path) as fh:
        rec = json.load(fh)
    k = rec.get("kernels", {}).get(key)
    if rec.get("workload") != workload or k is None:
        return None, None
    return k["traffic_bytes"], os.path.relpath(path, ROOT)


def measured_valu(workload, key, kernel_ms):
    """The vector-ALU issue fraction of kernel `key`: its SQ_INSTS_VALU per
    dispatch (wave instructions, whole device, from the rocprofv3 --pmc run
    committed under profiles/) over what the device issues at peak in
    kernel_ms; None if not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_mixed_3600x2700x75.json")
    if workload is None or not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    k = rec.get("kernels", {}).get(key)
    if rec.get("workload") != workload or k is None or "sq" not in k:
        return None
    insts = k["sq"]["insts_valu"]
    return dict(insts_per_launch=insts, peak_wave_insts_per_s=VALU_PEAK_WAVE_INSTS,
                frac=round(insts / (kernel_ms / 1e3) / VALU_PEAK_WAVE_INSTS, 4),
                source=os.path.relpath(path, ROOT))


def run_mixed(args, w):
    import dataclasses
    import gc
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.step import BatchedStep, PlainStep
    from mom5_amd.synth import make_case_device
    steps = w["steps"] if args.steps is None else args.steps
    warm = w["warmup"] if args.warmup is None else args.warmup
    ni, nj, nk = w["ni"], w["nj"], w["nk"]
    if args.shape:
        ni, nj, nk = (int(x) for x in args.shape.split("x"))
    cells = ni * nj * nk
    t0 = time.time()
    case = make_case_device(ni, nj, nk, land_frac=w["land_frac"], nts=w["nts"],
                            n_passive=w["n_passive"], passive_dtype=w["passive_dtype"])
    g = case.info["gpu"]
    log(f"generated {ni}x{nj}x{nk} on the device in {time.time() - t0:.1f}s "
        f"(dt={case.dt:.3f}s, dtbt={case.dtbt:.4f}s, nts={case.nts})")
    bs = BatchedStep(case, gpu=g)
    ntr, names = bs.ntr, bs.names
    rho = bs.dev["rho_dzt"]
    tuned = None
    if not args.no_autotune:
        from mom5_amd.tune import autotune
        t0 = time.time()
        tuned = autotune(bs)
        log(f"autotuned in {time.time() - t0:.1f}s: " + "; ".join(
            f"{k} {v['best']} {v['best_ms']:.3f} ms (default {v['default_ms']:.3f})"
            for k, v in tuned.items()))
        for k, v in tuned.items():
            log(f"  {k} candidates: " + ", ".join(
                "/".join(str(x) for kk, x in t.items() if kk != "ms") + f" {t['ms']:.2f}"
                for t in v["tried"]))
    c0 = [g.tracer_content(bs.tracer(n), rho) for n in range(ntr)]
    bs.run(warm)
    g.sync()
    g.mark(0)
    bs.run(2)
    g.mark(1)
    ms_eager = g.elapsed_ms(0, 1) / 2
    graph = bs.capture(2)                # two steps: time levels return to place
    g.mark(0)
    for _ in range(steps // 2):
        graph.launch()
    if steps % 2:                        # an odd count ends with one eager step
        bs.run(1)
    g.mark(1)
    ms = g.elapsed_ms(0, 1) / steps
    value = ntr * cells / (ms / 1e3)
    c1 = [g.tracer_content(bs.tracer(n), rho) for n in range(ntr)]
    nodes = graph.nodes
    graph.close()
    warm += 2                            # the eager pair also advanced the state
    log(f"batched step {ms:.2f} ms from graph replays ({nodes} nodes per 2 steps; eager "
        f"{ms_eager:.2f} ms) = {value:.3e} tracer-cell-updates/s")

    # the fused tracer-step kernels, each timed alone on the context's stream
    d = bs.dev
    kw = dict(uhrho_et=d["uhrho_et"], vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
              rho_dzt_tau=d["rho_dzt"], rho_dztr=d["rho_dztr"], dtime=case.dt)
    is64 = [bs.tracer(n).element_size() == 8 for n in range(ntr)]
    n64 = sum(is64)
    n32 = ntr - n64
    cand = []
    for want64, elem in ((True, 8), (False, 4)):
        grp = g.get_tuning("step_group_f64" if want64 else "step_group_f32")
        rows = g.get_tuning("step_rows_per_wave" if want64 else "step_rows_per_wave_f32")
        waves = g.get_tuning("step_waves_f64" if want64 else "step_waves_f32")
        # the tile shapes the library builds (others fall back to 8 wavefronts)
        shapes = {(2, 12), (1, 16)} if want64 else {(1, 16), (2, 16), (2, 12)}
        if (rows, waves) not in shapes:
            waves = 8
        key = f"k_step<{'double' if want64 else 'float'}, {grp}, {rows}, {waves}>"
        idx = [n for n in range(ntr) if is64[n] == want64][:grp]
        if len(idx) < grp:
            continue
        src = [bs.tracer(n) for n in idx]
        dst = [bs.T[n][1 - bs.cur] for n in idx]
        k_ms = time_launches(g, lambda: g.step_tracers(src, dst, **kw))
        nl = -(-(n64 if want64 else n32) // grp)
        cand.append(dict(kernel=f"{key} (fused MDFL-Sweby advection + explicit update; "
                                f"template args: type, tracers, rows per wave, waves)",
                         key=key,
                         kernel_ms=k_ms, bytes_per_launch=step_kernel_bytes(ni, nj, nk, elem, grp),
                         launches_per_step=nl))
        log(f"{cand[-1]['kernel']}: {k_ms:.3f} ms/launch x {nl} per step")
        del src, dst
    if n32 and bs.diffuse:
        # the implicit diffusion: the tracer step with minus without it.  With
        # vdiff_merge (T, S and up to 12 FP32 tracers) it is one launch of
        # k_vdiff_mix for every tracer; otherwise the FP32 tracers' k_vdiff_inc.
        twin = g.get_tuning("vdiff_solver") == 1 and 2 <= nk <= 80
        merged = not twin and g.get_tuning("vdiff_merge") == 1 and n32 <= 12 and n64 <= 2
        pf = "true" if g.get_tuning("vdiff_prefetch") else "false"
        if twin:
            idx = [n for n in range(ntr) if not is64[n]]
            h = -(-nk // 2)
            h = next(v for v in (8, 16, 24, 32, 40) if v >= h)
            key = f"k_vdiff_twin<{h}>"
            what = ("implicit vertical diffusion of the FP32 tracers, two wavefronts per "
                    "column, levels in registers")
            vbytes = vdiff_kernel_bytes(ni, nj, nk, 4, min(n32, 12))
            nvd = -(-n32 // 12)             # solver groups of up to 12 tracers
        elif merged:
            idx = list(range(ntr))
            nb = 4 if n32 <= 4 else 8 if n32 <= 8 else 10 if n32 <= 10 else 12
            key = f"k_vdiff_mix<{pf}, {nb}>"
            what = "implicit vertical diffusion of every tracer in one launch"
            vbytes = ni * nj * nk * (8 * n32 + 16 * n64 + 32) + ni * nj * (16 * ntr + 4)
            nvd = 1
        else:
            idx = [n for n in range(ntr) if not is64[n]]
            key = f"k_vdiff_inc<{pf}>"
            what = "implicit vertical diffusion of the FP32 tracers, increment form"
            vbytes = vdiff_kernel_bytes(ni, nj, nk, 4, min(n32, 12))
            nvd = -(-n32 // 12)             # solver groups of up to 12 tracers
        src = [bs.tracer(n) for n in idx]
        dst = [bs.T[n][1 - bs.cur] for n in idx]
        kwd = dict(kw, stfs=[bs.stf[n] for n in idx], btfs=[bs.btf[n] for n in idx],
                   dzwt=d["dzwt"], diff_cbt=d["diff_cbt"],
                   index_salt=bs.index_salt if merged else -1)
        t_with = time_launches(g, lambda: g.step_tracers(src, dst, **kwd))
        t_adv = time_launches(g, lambda: g.step_tracers(src, dst, **kw))
        cand.append(dict(kernel=f"{key} ({what}; timed as the tracer step with minus without "
                                f"it)", key=key, kernel_ms=(t_with - t_adv) / nvd,
                         bytes_per_launch=vbytes, launches_per_step=nvd))
        log(f"{cand[-1]['kernel']}: {cand[-1]['kernel_ms']:.3f} ms/launch x {nvd} per step")
        del src, dst
    dom = max(cand, key=lambda c: c["kernel_ms"] * c["launches_per_step"])
    achieved = dom["bytes_per_launch"] / (dom["kernel_ms"] / 1e3) / 1e9
    dkey = dom.pop("key")
    wname = w["name"] if not args.shape else None
    traffic, tsrc = measured_traffic(wname, dkey)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, **{
                    k: (round(v, 4) if isinstance(v, float) else v) for k, v in dom.items()})
    if tsrc:
        roof["traffic_source"] = tsrc
    # what limits the kernel: the larger of its HBM fraction (real traffic over
    # the kernel time) and its vector-ALU issue fraction
    valu = measured_valu(wname, dkey, dom["kernel_ms"])
    if valu:
        roof["valu_issue"] = valu
        hbm_real = (traffic / (dom["kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS) if traffic else 0.0
        roof["hbm_traffic_frac"] = round(hbm_real, 4)
        roof["limiter"] = "valu issue" if valu["frac"] > hbm_real else "hbm"
    step_bytes = mixed_step_bytes(ni, nj, nk, n64, n32)
    step_frac = step_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS

    bt_check = barotropic_blocked_vs_subcycle(g, bs) if bs.barotropic else None
    if bt_check:
        log(f"barotropic at full size: blocked (H={bt_check['blocked_subcycles_per_launch']}) "
            f"vs per-subcycle: {bt_check['points_differ']} differing values; "
            f"max |eta_t_bar| {bt_check['eta_t_bar_max_abs_m']:.3f} m")

    # keep the batched results (FP64 on the device, FP32 on the host), free the rest
    fin64 = {n: bs.tracer(n) for n in range(ntr) if is64[n]}
    fin32 = {n: bs.tracer(n).to_host() for n in range(ntr) if not is64[n]}
    tuning_used = {k: g.get_tuning(k) for k in TUNING_KEYS}
    del bs, cand
    case.tracers = []
    gc.collect()
    # a fresh context for the plain path: the mixed one holds its solver
    # scratch (one FP32 plane per passive tracer), which the plain run's
    # twelve FP64 tracers need the room of
    g.close()
    g = OceanGPU(case.dom, case.grid)

    # the plain per-tracer FP64 path from the same initial state
    pcase = dataclasses.replace(case, tracers=[
        (names[n], case.info["tracer"](n, np.float64)) for n in range(ntr)])
    plain = PlainStep(pcase, schemes=["mdfl_sweby"] * ntr, gpu=g)
    del pcase
    if not args.no_autotune:
        from mom5_amd.tune import autotune
        autotune(plain, barotropic=False)
    pc0 = [g.tracer_content(plain.tracer(n), rho) for n in range(ntr)]
    plain.run(warm)
    g.sync()
    g.mark(0)
    plain.run(steps)
    g.mark(1)
    ms_plain = g.elapsed_ms(0, 1) / steps
    value_plain = ntr * cells / (ms_plain / 1e3)
    pc1 = [g.tracer_content(plain.tracer(n), rho) for n in range(ntr)]
    log(f"plain FP64 step {ms_plain:.2f} ms ({value_plain:.3e} tracer-cell-updates/s)")
    ref = {n: plain.tracer(n) for n in range(ntr)}
    del plain
    gc.collect()

    errors = {}
    for n in range(ntr):
        if is64[n]:
            e = g.tracer_error(fin64[n], ref[n])
            errors[names[n]] = dict(cells_differ_from_plain_fp64=e["cells_differ"],
                                    cells=e["cells"])
            continue
        a = DeviceArray.from_host(fin32[n])
        e = g.tracer_error(a, ref[n])
        e2 = g.tracer_error(a, ref[n], rel_floor=1e-6 * e["max_ref"])
        errors[names[n]] = dict(max_rel_normwise=e["normwise"], rms_rel=e["rms_rel"],
                                max_rel_cell_where_ref_ge_1em6_max=e2["max_rel_cell"])
        del a
    drift = {names[n]: dict(mixed=abs(c1[n] - c0[n]) / abs(c0[n]),
                            plain_fp64=abs(pc1[n] - pc0[n]) / abs(pc0[n])) for n in range(ntr)}
    worst = max((e["max_rel_normwise"] for e in errors.values() if "max_rel_normwise" in e),
                default=0.0)
    ts_equal = all(e.get("cells_differ_from_plain_fp64", 0) == 0 for e in errors.values())
    log(f"T/S bitwise equal to plain FP64: {ts_equal}; worst passive normwise error {worst}")
    extra = dict(timed_as="hipGraph replays of 2 captured steps", graph_nodes_per_2_steps=nodes,
                 eager_ms_per_step=ms_eager, autotune=_tune_summary(tuned),
                 plain_fp64_value=value_plain, plain_fp64_ms_per_step=ms_plain,
                 speedup_vs_plain_fp64=value / value_plain, step_algorithmic_bytes=step_bytes,
                 step_hbm_frac=round(step_frac, 4), ts_bitwise_equal_plain_fp64=ts_equal,
                 passive_max_rel_err_normwise=worst, errors_after_steps=warm + steps,
                 tracer_errors=errors, content_drift_rel=drift,
                 barotropic_full_size_check=bt_check)
    name = w["name"] if not args.shape else f"{w['name']}@{args.shape}"
    cfg = {"workload": name, "grid": [ni, nj, nk], "tracers": ntr,
           "tracer_dtypes": {"f64": n64, "f32": n32}, "schemes": "mdfl_sweby (all)",
           "barotropic_nts": case.nts,
           "path": ("batched mixed" if n32 else "batched FP64") + " (mom_gpu_step_tracers)",
           "tuning": tuning_used}
    return dict(value=value, ms=ms, steps=steps, warmup=warm,
                dtype="f64+f32" if n32 else "f64", cfg=cfg, roof=roof, extra=extra, ntr=ntr)


def barotropic_blocked_vs_subcycle(g, bs):
    """update_ocean_barotropic at the workload's full size and subcycle count
    from the step's current state, once with the tuned subcycles per launch
    (temporally blocked kernel) and once per subcycle (0): every value of
    eta_t_bar(taup1), udrho(taup1) and conv_rho_ud_t compared bit for bit,
    and the largest |eta_t_bar| reported (initial eta ~ N(0, 0.1 m))."""
    from mom5_amd import DeviceArray
    knob = "barotropic_subcycles_per_launch"
    H = g.get_tuning(knob)
    hs = (H, 0) if H != 0 else (0, 4)
    res = {}
    for h in hs:
        g.set_tuning(knob, h)
        eb, cv = DeviceArray(g.shape2), DeviceArray(g.shape2)
        ud = DeviceArray((2,) + g.shape2)
        g.update_ocean_barotropic(**dict(bs.bt_args(), eta_t_bar_taup1=eb, udrho_taup1=ud,
                                         conv_rho_ud_t=cv))
        g.sync()
        res[h] = [x.to_host() for x in (eb, ud, cv)]
        del eb, ud, cv
    g.set_tuning(knob, H)
    a, b = res[hs[0]], res[hs[1]]
    differ = sum(int(np.count_nonzero(x.view(np.int64) != y.view(np.int64)))
                 for x, y in zip(a, b))
    eta = a[0]
    finite = bool(np.isfinite(eta).all() and np.isfinite(a[1]).all())
    emax = float(np.abs(eta).max()) if finite else float("inf")
    return dict(blocked_subcycles_per_launch=hs[0], compared_with_subcycles_per_launch=hs[1],
                nts=bs.case.nts, points_compared=int(sum(x.size for x in a)),
                points_differ=differ, eta_t_bar_max_abs_m=emax,
                eta_bounded=bool(finite and emax < 10.0))


def _tune_summary(rep):
    if not rep:
        return None
    return {k: dict(best=v["best"], best_ms=round(v["best_ms"], 4),
                    default=v["default"], default_ms=round(v["default_ms"], 4))
            for k, v in rep.items()}


TUNING_KEYS = ("barotropic_subcycles_per_launch", "mdfl_rows_per_wave", "step_group_f64",
               "step_group_f32", "step_rows_per_wave", "step_rows_per_wave_f32",
               "step_waves_f32", "step_waves_f64", "vdiff_prefetch", "vdiff_block",
               "vdiff_merge", "vdiff_solver", "step_streams", "barotropic_land_tiles")


def run_plain(args, w):
    from mom5_amd.step import PlainStep
    steps = 20 if args.steps is None else args.steps
    warm = 3 if args.warmup is None else args.warmup
    t0 = time.time()
    case = make(w)
    log(f"generated {w['name']} in {time.time() - t0:.1f}s (dt={case.dt:.3f}s, "
        f"dtbt={case.dtbt:.4f}s, nts={case.nts})")
    st = PlainStep(case, schemes=w["schemes"])
    g = st.gpu
    tuned = None
    if not args.no_autotune:
        from mom5_amd.tune import autotune
        tuned = autotune(st)
    st.run(warm)
    g.sync()
    g.mark(0)
    st.run(steps)
    g.mark(1)
    ms = g.elapsed_ms(0, 1) / steps
    ntr = st.ntr
    cells = case.ni * case.nj * case.nk
    value = ntr * cells / (ms / 1e3)

    # dominant kernel: one MDFL-Sweby launch, timed alone on the same stream
    d = st.dev
    th = g.field3()
    T0 = st.tracer(0)
    k_ms = time_launches(g, lambda: g.horz_advect_tracer(
        "mdfl_sweby", T0, th, uhrho_et=d["uhrho_et"], vhrho_nt=d["vhrho_nt"],
        wrho_bt=d["wrho_bt"], rho_dzt_tau=d["rho_dzt"], dtime=case.dt))
    kb = mdfl_bytes_per_launch(case.ni, case.nj, case.nk)
    achieved = kb / (k_ms / 1e3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="k_mdfl (MDFL-Sweby advection, one tracer)",
                kernel_ms=round(k_ms, 4), bytes_per_launch=kb)
    log(f"step {ms:.3f} ms; mdfl launch {k_ms:.4f} ms")
    cfg = {"workload": w["name"], "grid": [case.ni, case.nj, case.nk], "tracers": ntr,
           "schemes": w["schemes"], "barotropic_nts": case.nts, "path": "plain per-tracer FP64",
           "tuning": {k: g.get_tuning(k) for k in TUNING_KEYS}}
    return dict(value=value, ms=ms, steps=steps, warmup=warm, dtype="f64", cfg=cfg, roof=roof,
                extra=dict(autotune=_tune_summary(tuned)), ntr=ntr)


def profile_kernels(args, w):
    """Launch only the kernels of one step, three dispatches each, for
    rocprofv3 (--kernel-trace --stats, --pmc): k_copy of one 3-D FP64 field
    (the calibration of FETCH_SIZE and WRITE_SIZE: it reads and writes exactly
    its size), one pass of each fused tracer-step group, the implicit
    diffusion of every tracer and the barotropic subcycles of one step."""
    from mom5_amd import DeviceArray
    from mom5_amd.synth import make_case_device
    ni, nj, nk = w["ni"], w["nj"], w["nk"]
    if args.shape:
        ni, nj, nk = (int(x) for x in args.shape.split("x"))
    case = make_case_device(ni, nj, nk, land_frac=w["land_frac"], nts=w["nts"],
                            n_passive=w["n_passive"], passive_dtype=np.float32)
    g = case.info["gpu"]
    f = case.fields
    kw = dict(uhrho_et=f["uhrho_et"], vhrho_nt=f["vhrho_nt"], wrho_bt=f["wrho_bt"],
              rho_dzt_tau=f["rho_dzt"], rho_dztr=f["rho_dztr"], dtime=case.dt)
    src = f["rho_dzt"]
    dst = DeviceArray(src.shape)
    for _ in range(3):
        g.copy(dst, src)
    g.sync()
    log(f"k_copy x3: {src.nbytes} bytes read and written per dispatch")
    trs = [t for _, t in case.tracers]
    for size, knob in ((8, "step_group_f64"), (4, "step_group_f32")):
        grp = g.get_tuning(knob)
        sel = [t for t in trs if t.element_size() == size][:grp]
        out = [DeviceArray(t.shape, t.dtype) for t in sel]
        for _ in range(3):
            g.step_tracers(sel, out, **kw)
        g.sync()
        log(f"k_step<{'double' if size == 8 else 'float'},G={grp}> x3: algorithmic "
            f"{step_kernel_bytes(ni, nj, nk, size, grp)} bytes per dispatch")
    # the implicit diffusion of every tracer of the step (with the advection passes)
    out = [DeviceArray(t.shape, t.dtype) for t in trs]
    ns = len(trs)
    up = lambda a: a if isinstance(a, DeviceArray) else DeviceArray.from_host(a)
    kwd = dict(kw, stfs=[up(a) for a in f["stf"][:ns]], btfs=[up(a) for a in f["btf"][:ns]],
               dzwt=up(f["dzwt"]), diff_cbt=up(f["diff_cbt"]), index_salt=1)
    for _ in range(3):
        g.step_tracers(trs, out, **kwd)
    g.sync()
    log(f"implicit diffusion of {ns} tracers x3 (k_vdiff_mix when merged)")
    # the barotropic subcycles of one step (k_bt_block launches and their tails)
    up2 = {k: up(f[k]) for k in ("eta_t_bar", "udrho", "forcing_bt", "thicku", "patm", "pme",
                                  "river", "source", "eta_smooth", "rho_g")}
    eb, cv = DeviceArray(g.shape2), DeviceArray(g.shape2)
    ud = DeviceArray((2,) + g.shape2)
    for _ in range(3):
        g.update_ocean_barotropic(nts=case.nts, dtbt=case.dtbt, eta_t_bar_tau=up2["eta_t_bar"],
                                  udrho_tau=up2["udrho"], forcing_bt=up2["forcing_bt"],
                                  thicku=up2["thicku"], patm=up2["patm"], pme=up2["pme"],
                                  river=up2["river"], source=up2["source"],
                                  eta_smooth=up2["eta_smooth"], rho_g=up2["rho_g"],
                                  eta_t_bar_taup1=eb, udrho_taup1=ud, conv_rho_ud_t=cv)
    g.sync()
    log(f"update_ocean_barotropic x3 (nts={case.nts}, "
        f"{g.get_tuning('barotropic_subcycles_per_launch')} subcycles per launch)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 18 for the mixed workload, 20 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 2 for the mixed workload, 3 otherwise)")
    ap.add_argument("--workload", default="mixed", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the kernels measured with rocprofv3 --pmc (mixed workload)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing the candidates")
    ap.add_argument("--shape", default=None,
                    help="NIxNJxNK: another grid for the mixed workload (quick runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    if args.profile_kernels:
        profile_kernels(args, w)
        return
    if not w.get("ensemble"):
        r = run_mixed(args, w) if w.get("mixed") else run_plain(args, w)

    if w.get("ensemble") and not args.profile_kernels:
        r = run_ensemble(args, w)
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        cpu = cpu_baseline(w)

    line = {"metric": "tracer-cell-updates/s", "value": r["value"],
            "unit": "tracer-cell-updates/s", "n_gpus": 1, "steps": r["steps"],
            "warmup": r["warmup"], "ms_per_step": r["ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": r["dtype"], "data": "synthetic",
            "config": r["cfg"], "roofline": r["roof"], "cpu_baseline": cpu}
    line.update(r["extra"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
