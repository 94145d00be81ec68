"""Autotuner for the launch-shape knobs of a context.

The knobs (``mom_gpu_set_tuning``) change tile heights, how many tracers
share one pass and how many barotropic subcycles one launch advances; they
never change results (the GPU tests check bit-for-bit equality across
them).  The best setting depends on the grid and the tracer mix, so the
tuner times each candidate on the caller's own arrays, with HIP events on
the context's stream, and keeps the fastest.

    from mom5_amd.tune import autotune
    report = autotune(step)            # a BatchedStep or PlainStep

Each kernel family is tuned on its own (FP64 tracer passes, FP32 tracer
passes, the barotropic subcycle, the MDFL kernel of the plain path), so the
search is a sum, not a product, of the candidate counts.
"""
from __future__ import annotations

import itertools


def _time(gpu, fn, reps):
    fn()                        # first launch of a new shape: code object load
    gpu.sync()
    gpu.mark(6)
    for _ in range(reps):
        fn()
    gpu.mark(7)
    return gpu.elapsed_ms(6, 7) / reps


def _search(gpu, knobs, candidates, fn, reps):
    saved = {k: gpu.get_tuning(k) for k in knobs}
    results = []
    for vals in candidates:
        for k, v in zip(knobs, vals):
            gpu.set_tuning(k, v)
        results.append((_time(gpu, fn, reps), dict(zip(knobs, vals))))
    best_ms, best = min(results, key=lambda r: r[0])
    for k, v in best.items():
        gpu.set_tuning(k, v)
    default = next(ms for ms, c in results if c == saved)
    return dict(best=best, best_ms=best_ms, default=saved, default_ms=default,
                tried=[dict(ms=round(ms, 4), **c) for ms, c in results])


def autotune(step, reps=3, barotropic=True):
    """Tune the knobs that the step's kernels use; returns a report per
    kernel family and leaves the context set to the fastest choices."""
    from .step import BatchedStep
    g, d, c = step.gpu, step.dev, step.case
    report = {}
    if isinstance(step, BatchedStep):
        kw = dict(uhrho_et=d["uhrho_et"], vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
                  rho_dzt_tau=d["rho_dzt"], rho_dztr=d["rho_dztr"], dtime=c.dt)
        cur, nxt = step.cur, 1 - step.cur
        fp32 = [c_ + (8,) for c_ in itertools.product((1, 2, 4), (2, 4))] + \
            [(1, 2, 16), (2, 2, 16), (4, 2, 16), (4, 1, 16), (6, 1, 16), (8, 1, 16),
             (10, 1, 16), (5, 2, 12)]
        for tag, size, knobs, cands in (
                ("fp64", 8, ("step_group_f64", "step_rows_per_wave", "step_waves_f64"),
                 [c_ + (8,) for c_ in itertools.product((1, 2, 4), (2, 4))] +
                 [(1, 2, 12), (2, 2, 12), (1, 1, 16), (2, 1, 16), (4, 1, 16)]),
                ("fp32", 4, ("step_group_f32", "step_rows_per_wave_f32", "step_waves_f32"),
                 fp32)):
            idx = [n for n in range(step.ntr) if step.tracer(n).element_size() == size]
            if not idx:
                continue
            src = [step.T[n][cur] for n in idx]
            dst = [step.T[n][nxt] for n in idx]
            report[tag] = _search(g, knobs, cands, lambda: g.step_tracers(src, dst, **kw), reps)
        if step.diffuse and any(step.tracer(n).element_size() == 4 for n in range(step.ntr)):
            # the implicit diffusion of a step with FP32 tracers: the whole
            # tracer step (tuned advection passes) per candidate
            a = step.step_args()
            src_all, dst_all = a.pop("T_taum1s"), a.pop("T_taup1s")
            # the register-resident solver (vdiff_solver 1) ignores the
            # other three knobs
            report["vdiff"] = _search(
                g, ("vdiff_prefetch", "vdiff_block", "vdiff_merge", "vdiff_solver"),
                [(pf, bx, mg, 0) for mg in (0, 1) for pf in (0, 1) for bx in (64, 256)] +
                [(g.get_tuning("vdiff_prefetch"), g.get_tuning("vdiff_block"),
                  g.get_tuning("vdiff_merge"), 1)],
                lambda: g.step_tracers(src_all, dst_all, **a), reps)
    else:
        th = step.th
        T0 = step.tracer(0)
        report["mdfl"] = _search(
            g, ("mdfl_rows_per_wave",), [(1,), (2,), (4,)],
            lambda: g.horz_advect_tracer("mdfl_sweby", T0, th, uhrho_et=d["uhrho_et"],
                                         vhrho_nt=d["vhrho_nt"], wrho_bt=d["wrho_bt"],
                                         rho_dzt_tau=d["rho_dzt"], dtime=c.dt), reps)
    if isinstance(step, BatchedStep) and step.barotropic and any(
            step.tracer(n).element_size() == 4 for n in range(step.ntr)):
        # FP32 tracers on a second stream: timed as whole steps (tracer call
        # and barotropic call, each writing the next time level only)
        a = step.step_args()
        src_all, dst_all = a.pop("T_taum1s"), a.pop("T_taup1s")
        bt = step.bt_args()
        report["streams"] = _search(
            g, ("step_streams",), [(1,), (2,)],
            lambda: (g.step_tracers(src_all, dst_all, **a), g.update_ocean_barotropic(**bt)),
            reps)
    if barotropic and getattr(step, "barotropic", c.dtbt > 0.0) and c.dtbt > 0.0:
        cur, nxt = step.cur, 1 - step.cur
        report["barotropic"] = _search(
            g, ("barotropic_subcycles_per_launch",), [(n,) for n in range(0, 13)],
            lambda: g.update_ocean_barotropic(
                nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=step.eta_bar[cur],
                udrho_tau=step.udrho[cur], forcing_bt=d["forcing_bt"], thicku=d["thicku"],
                patm=d["patm"], pme=d["pme"], river=d["river"], source=d["source"],
                eta_smooth=d["eta_smooth"], rho_g=d["rho_g"],
                eta_t_bar_taup1=step.eta_bar[nxt], udrho_taup1=step.udrho[nxt],
                conv_rho_ud_t=step.conv), reps)
    return report
