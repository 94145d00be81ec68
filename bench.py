#!/usr/bin/env python3
"""Throughput of one baroclinic step of the MOM5 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]

A step is the reference's order (ocean_model.F90:1441-2108, the part this
build implements): compute_tmask_limit; per tracer horz/vert advection into
th_tendency and the explicit update; vert_diffuse_implicit (invtri) for all
tracers; the barotropic subcycle with its time-averaging tail.  Inputs are
seeded synthetic fields resident in HBM before timing starts.

Prints one JSON line: tracer-cell-updates/s for the whole step (value), the
roofline of the dominant kernel (MDFL-Sweby advection, HIP-event timed on the
context's stream), and a CPU baseline: the CPU restatement timed on a bounded
sample on the host (single thread).
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

WORKLOADS = {
    1: dict(name="cfg1_plain_fp64_1440x1080x50", ni=1440, nj=1080, nk=50, land_frac=0.30,
            nts=80, n_passive=1, schemes=["mdfl_sweby", "mdfl_sweby", "quicker"],
            sample=dict(ni=480, nj=360)),
    0: dict(name="cfg0_plain_fp64_90x60x25", ni=90, nj=60, nk=25, land_frac=0.25, nts=0,
            n_passive=0, schemes=["mdfl_sweby", "mdfl_sweby"], sample=dict(ni=90, nj=60)),
}


def make(w, ni=None, nj=None, seed=0x4D4F4D35):
    from mom5_amd.synth import make_case
    return make_case(ni or w["ni"], nj or w["nj"], w["nk"], land_frac=w["land_frac"],
                     nts=w["nts"] or 80, n_passive=w["n_passive"], barotropic=w["nts"] > 0,
                     seed=seed)


def mdfl_bytes_per_launch(ni, nj, nk):
    """Algorithmic bytes of one MDFL launch: per cell T read, th_tendency
    read+write, uhrho_et/vhrho_nt/wrho_bt/rho_dzt read (8 B each); per column
    six FP64 metrics and int32 kmt."""
    return ni * nj * nk * (8 + 16 + 32) + ni * nj * 52


def cpu_baseline(w, steps=8):
    """The CPU restatement on a bounded sample of the same workload."""
    from oracle.oracle import Oracle
    s = w["sample"]
    case = make(w, ni=s["ni"], nj=s["nj"], seed=0x4D4F4D35 + 7)
    o = Oracle(case)
    f = case.fields
    Ts = [np.array(t, dtype=np.float64) for _, t in case.tracers]
    names = [n for n, _ in case.tracers]
    mins = [(-2.0 if n == "temp" else 1.0 if n == "salt" else -0.1) for n in names]
    maxs = [(32.0 if n == "temp" else 42.0 if n == "salt" else 1.0) for n in names]

    def step(Ts):
        lim = o.tmask_limit(Ts, mins, maxs, 0, 1, True)
        out = []
        for n, T in enumerate(Ts):
            th = np.zeros_like(T)
            if w["schemes"][n] == "quicker":
                th[:, 1:-1, 1:-1] -= o.horz_quicker(T, T, lim[n])
                th[:, 1:-1, 1:-1] -= o.vert_quicker(T, T, lim[n])
            else:
                o.horz_advect_mdfl(T, th, case.dt, lim=1.0)
            out.append(o.update_explicit(T, th, case.dt))
        o.vert_diffuse_implicit(out, f["stf"], f["btf"], case.dt, index_salt=1)
        if w["nts"]:
            o.update_ocean_barotropic(case.nts, case.dtbt)
        return out

    Ts = step(Ts)                       # warm-up
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        Ts = step(Ts)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    cells = case.ni * case.nj * case.nk * len(Ts)
    return dict(value=cells / dt, unit="tracer-cell-updates/s", cores=1, kind="port",
                sample=f"{case.ni}x{case.nj}x{case.nk}, {len(Ts)} tracers, full step incl. "
                       f"barotropic nts={case.nts}; median of {steps} steps after 1 warm-up "
                       f"({sum(times):.1f} s timed)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not an implementation.", file=sys.stderr)
        sys.exit(2)

    from mom5_amd.step import PlainStep
    w = WORKLOADS[args.config]
    t0 = time.time()
    case = make(w)
    print(f"# generated {w['name']} in {time.time() - t0:.1f}s (dt={case.dt:.3f}s, "
          f"dtbt={case.dtbt:.4f}s, nts={case.nts})", file=sys.stderr, flush=True)
    st = PlainStep(case, schemes=w["schemes"])
    g = st.gpu
    st.run(args.warmup)
    g.sync()
    g.mark(0)
    st.run(args.steps)
    g.mark(1)
    ms_total = g.elapsed_ms(0, 1)
    ms = ms_total / args.steps
    ntr = st.ntr
    cells = case.ni * case.nj * case.nk
    value = ntr * cells / (ms / 1e3)

    # dominant kernel: one MDFL-Sweby launch, timed alone on the same stream
    d = st.dev
    th = g.field3()
    T0 = st.tracer(0)
    reps = 10
    g.horz_advect_tracer("mdfl_sweby", T0, th, uhrho_et=d["uhrho_et"], vhrho_nt=d["vhrho_nt"],
                         wrho_bt=d["wrho_bt"], rho_dzt_tau=d["rho_dzt"], dtime=case.dt)
    g.sync()
    g.mark(2)
    for _ in range(reps):
        g.horz_advect_tracer("mdfl_sweby", T0, th, uhrho_et=d["uhrho_et"],
                             vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
                             rho_dzt_tau=d["rho_dzt"], dtime=case.dt)
    g.mark(3)
    k_ms = g.elapsed_ms(2, 3) / reps
    kb = mdfl_bytes_per_launch(case.ni, case.nj, case.nk)
    achieved = kb / (k_ms / 1e3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="k_mdfl (MDFL-Sweby advection, one tracer)",
                kernel_ms=round(k_ms, 4), bytes_per_launch=kb)
    print(f"# step {ms:.3f} ms; mdfl launch {k_ms:.4f} ms", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        print("# timing the CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w)

    line = {"metric": "tracer-cell-updates/s", "value": value,
            "unit": "tracer-cell-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": w["name"], "grid": [case.ni, case.nj, case.nk],
                       "tracers": ntr, "schemes": w["schemes"], "barotropic_nts": case.nts,
                       "path": "plain per-tracer FP64"},
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
