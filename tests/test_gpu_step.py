"""The batched tracer step (mom_gpu_step_tracers).

* FP64 tracers: T(taup1) bit-for-bit equal to the per-tracer path
  (horz_advect_tracer into a zeroed th_tendency, explicit update,
  vert_diffuse_implicit) for every grouping/tile knob, and within 1e-12
  relative per cell of the CPU restatement of the same sequence (the
  kernels' limited fluxes are division-free).
* FP32 passive tracers: error against the FP64 restatement, normwise
  (max |err| / max |T|) within 1e-5, per-cell relative error reported.
"""
import numpy as np
import pytest
from _cases import assert_close_cells

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["cyc_130x70x12", "regional_41x29x9"])
def setup(request, oracle_lib):
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    if request.param == "cyc_130x70x12":
        case = make_case(130, 70, 12, land_frac=0.3, nts=10, seed=13, n_passive=3,
                         stf_scale=1e-3)
    else:
        case = make_case(41, 29, 9, land_frac=0.2, cyclic_x=False, lat_s=20.0, lat_n=50.0,
                         lon_span=30.0, nts=16, seed=12, n_passive=2, stf_scale=1e-3)
    print(f"[case {request.param}]", flush=True)
    orc = oracle_lib.Oracle(case)
    gpu = OceanGPU(case.dom, case.grid)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    yield case, orc, gpu, dev
    gpu.close()


def _oracle_step(case, orc, Ts):
    f = case.fields
    outs = []
    for T in Ts:
        th = np.zeros_like(T)
        orc.horz_advect_mdfl(T, th, case.dt, lim=1.0)
        outs.append(orc.update_explicit(T, th, case.dt))
    orc.vert_diffuse_implicit(outs, f["stf"][:len(Ts)], f["btf"][:len(Ts)], case.dt, index_salt=1)
    return outs


def _run_step(case, gpu, dev, host_tracers, dtypes):
    from mom5_amd import DeviceArray
    f = case.fields
    tin = [DeviceArray.from_host(t.astype(dt)) for t, dt in zip(host_tracers, dtypes)]
    tout = [DeviceArray(t.shape, dt) for t, dt in zip(host_tracers, dtypes)]
    for a in tout:
        a.zero()
    stf = [DeviceArray.from_host(s) for s in f["stf"][:len(host_tracers)]]
    btf = [DeviceArray.from_host(s) for s in f["btf"][:len(host_tracers)]]
    gpu.step_tracers(tin, tout, uhrho_et=dev["uhrho_et"], vhrho_nt=dev["vhrho_nt"],
                     wrho_bt=dev["wrho_bt"], rho_dzt_tau=dev["rho_dzt"], dtime=case.dt,
                     stfs=stf, btfs=btf, dzwt=dev["dzwt"], diff_cbt=dev["diff_cbt"],
                     index_salt=1, rho_dztr=dev["rho_dztr"])
    gpu.sync()
    return [a.to_host() for a in tout]


@pytest.mark.parametrize("group,rows,waves", [(1, 4, 8), (2, 2, 8), (1, 2, 8), (2, 2, 12),
                                              (1, 2, 12), (2, 1, 16), (1, 1, 16), (4, 1, 16)])
def test_fp64_step_bitwise_equal_to_plain_path_close_to_oracle(setup, group, rows, waves):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    gpu.set_tuning("step_group_f64", group)
    gpu.set_tuning("step_rows_per_wave", rows)
    gpu.set_tuning("step_waves_f64", waves)
    Ts = [case.tracers[0][1], case.tracers[1][1]]
    got = _run_step(case, gpu, dev, Ts, [np.float64, np.float64])
    ref = _oracle_step(case, orc, Ts)
    # per-tracer path on the GPU
    f = case.fields
    plain = []
    for T in Ts:
        dT, th, out = DeviceArray.from_host(T), gpu.field3(), gpu.field3()
        gpu.horz_advect_tracer("mdfl_sweby", dT, th, uhrho_et=dev["uhrho_et"],
                               vhrho_nt=dev["vhrho_nt"], wrho_bt=dev["wrho_bt"],
                               rho_dzt_tau=dev["rho_dzt"], dtime=case.dt)
        gpu.update_tracer_explicit(dT, th, dev["rho_dzt"], dev["rho_dztr"], case.dt, out)
        plain.append(out)
    gpu.vert_diffuse_implicit(plain, [DeviceArray.from_host(s) for s in f["stf"][:2]],
                              [DeviceArray.from_host(s) for s in f["btf"][:2]], dev["diff_cbt"],
                              dev["rho_dzt"], dev["dzwt"], case.dt, index_salt=1)
    gpu.sync()
    gpu.set_tuning("step_waves_f64", 8)
    for n in range(2):
        a = got[n][:, 1:-1, 1:-1]
        assert np.array_equal(a, plain[n].to_host()[:, 1:-1, 1:-1]), f"tracer {n} vs plain"
        assert_close_cells(a, ref[n][:, 1:-1, 1:-1], what=f"tracer {n} vs oracle")


@pytest.mark.parametrize("group,rows", [(2, 4), (4, 2), (1, 4)])
def test_fp32_passive_error_vs_fp64(setup, group, rows):
    case, orc, gpu, dev = setup
    gpu.set_tuning("step_group_f32", group)
    gpu.set_tuning("step_rows_per_wave_f32", rows)
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    got = _run_step(case, gpu, dev, Ts, dtypes)
    ref = _oracle_step(case, orc, Ts)
    tm = case.grid["tmask"][:, 1:-1, 1:-1] > 0
    for n in range(2):
        assert_close_cells(got[n][:, 1:-1, 1:-1], ref[n][:, 1:-1, 1:-1], what=f"tracer {n}")
    for n in range(2, len(Ts)):
        g = got[n][:, 1:-1, 1:-1].astype(np.float64)[tm]
        r = ref[n][:, 1:-1, 1:-1][tm]
        err = np.abs(g - r)
        normwise = err.max() / np.abs(r).max()
        rms = np.sqrt((err ** 2).mean()) / np.sqrt((r ** 2).mean())
        cell = np.quantile(err / np.maximum(np.abs(r), 1e-30), [0.5, 0.99, 1.0])
        print(f"tracer {n}: normwise max {normwise:.2e}, rms {rms:.2e}, per-cell rel "
              f"median/p99/max {cell[0]:.2e}/{cell[1]:.2e}/{cell[2]:.2e}")
        assert normwise <= 1e-5
        assert rms <= 1e-6


def test_step_with_other_tendencies(setup):
    """th_tendency inputs (the other tendency terms the step adds to the
    advection): FP64 tracers within 1e-12 per cell of the CPU restatement of
    th + advection + explicit update; FP32 tracers bit for bit the FP32 update's
    definition, T(taup1) = [update without th] + th * (dtime / rho(taup1)),
    each operand rounded to FP32 (the library applies that last addition in
    a second pass), at every cell of the compute domain."""
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    f = case.fields
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    rng = np.random.default_rng(17)
    ths = [1e-3 * rng.standard_normal(t.shape) for t in Ts]
    kw = dict(uhrho_et=dev["uhrho_et"], vhrho_nt=dev["vhrho_nt"], wrho_bt=dev["wrho_bt"],
              rho_dzt_tau=dev["rho_dzt"], dtime=case.dt, rho_dztr=dev["rho_dztr"])
    outs = {}
    for with_th in (False, True):
        tin = [DeviceArray.from_host(t.astype(dt)) for t, dt in zip(Ts, dtypes)]
        tout = [DeviceArray(t.shape, dt) for t, dt in zip(Ts, dtypes)]
        for a in tout:
            a.zero()
        extra = dict(th_tendencies=[DeviceArray.from_host(t) for t in ths]) if with_th else {}
        gpu.step_tracers(tin, tout, **kw, **extra)
        gpu.sync()
        outs[with_th] = [a.to_host()[:, 1:-1, 1:-1] for a in tout]
    for n in range(2):
        th = ths[n].copy()
        orc.horz_advect_mdfl(Ts[n], th, case.dt, lim=1.0)
        want = orc.update_explicit(Ts[n], th, case.dt)[:, 1:-1, 1:-1]
        assert_close_cells(outs[True][n], want, what=f"tracer {n} with th")
    c32 = (case.dt * f["rho_dztr"][:, 1:-1, 1:-1]).astype(np.float32)
    for n in range(2, len(Ts)):
        th32 = ths[n][:, 1:-1, 1:-1].astype(np.float32)
        want = (outs[False][n] + th32 * c32).astype(np.float32)
        assert np.array_equal(outs[True][n].view(np.uint32), want.view(np.uint32)), n
        assert not np.array_equal(outs[True][n], outs[False][n])


def test_two_streams_equal_one_stream():
    """step_streams = 2 (FP32 tracers on the context's second stream,
    overlapping the FP64 tracers and the barotropic calls) gives the bits of
    step_streams = 1, with diffusion and other th_tendency terms, over steps
    that alternate the two modes on one context (the joins at the
    entry points order them) and with reads in between."""
    from mom5_amd import DeviceArray
    from mom5_amd.step import BatchedStep
    from mom5_amd.synth import make_case
    case = make_case(90, 50, 11, land_frac=0.3, nts=8, seed=31, n_passive=4, stf_scale=1e-3)
    print("[two streams vs one: 90x50x11, 2 FP64 + 4 FP32, barotropic nts=8]", flush=True)
    dt = [np.float64, np.float64] + [np.float32] * 4
    import dataclasses
    from mom5_amd import OceanGPU
    g = OceanGPU(case.dom, case.grid)

    def member():
        c = dataclasses.replace(case, tracers=[(nm, t.astype(dt[i]))
                                               for i, (nm, t) in enumerate(case.tracers)])
        m = BatchedStep(c, gpu=g)
        rng = np.random.default_rng(8)
        m.th = [DeviceArray.from_host(1e-3 * rng.standard_normal(case.tracers[0][1].shape))
                for _ in range(m.ntr)]
        return m

    one, two = member(), member()
    for k in range(4):
        g.set_tuning("step_streams", 2)
        two.run(1)
        if k == 1:
            # a read of the FP32 state between steps (joins the second stream)
            g.tracer_content(two.tracer(3), two.dev["rho_dzt"])
        g.set_tuning("step_streams", 1)
        one.run(1)
    g.set_tuning("step_streams", 2)
    two.run(2)
    g.set_tuning("step_streams", 1)
    one.run(2)
    g.sync()
    for n in range(one.ntr):
        assert np.array_equal(one.tracer(n).to_host(), two.tracer(n).to_host()), n
    assert np.array_equal(one.eta_bar[one.cur].to_host(), two.eta_bar[two.cur].to_host())
    assert np.array_equal(one.udrho[one.cur].to_host(), two.udrho[two.cur].to_host())
    g.close()


def test_fp32_results_independent_of_launch_shape(setup):
    """Grouping, tile height and the column solver's load schedule change
    which threads do the work and when, not the arithmetic of any tracer:
    FP32 results are bit-for-bit the same."""
    case, orc, gpu, dev = setup
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    gpu.set_tuning("vdiff_solver", 0)
    base = None
    for group, rows, waves, pf in [(1, 2, 8, 0), (2, 2, 8, 1), (4, 2, 8, 0), (1, 4, 8, 1),
                                   (2, 4, 8, 0), (4, 4, 8, 1), (1, 2, 16, 0), (2, 2, 16, 1),
                                   (4, 2, 16, 0), (4, 2, 16, 1), (10, 1, 16, 1), (8, 1, 16, 0),
                                   (4, 1, 16, 1), (5, 2, 12, 0)]:
        gpu.set_tuning("step_group_f32", group)
        gpu.set_tuning("step_rows_per_wave_f32", rows)
        gpu.set_tuning("step_waves_f32", waves)
        gpu.set_tuning("vdiff_prefetch", pf)
        gpu.set_tuning("vdiff_block", 256 if group % 2 else 64)
        gpu.set_tuning("vdiff_merge", 1 if rows == 2 else 0)
        got = [g[:, 1:-1, 1:-1] for g in _run_step(case, gpu, dev, Ts, dtypes)]
        if base is None:
            base = got
        for n in range(len(Ts)):
            assert np.array_equal(got[n], base[n]), (group, rows, waves, pf, n)
    # the register-resident solver: its own bits, the same for every shape
    gpu.set_tuning("vdiff_solver", 1)
    base1 = None
    for group, rows, waves in [(1, 2, 8), (4, 2, 16), (2, 1, 16), (4, 1, 16), (5, 2, 12)]:
        gpu.set_tuning("step_group_f32", group)
        gpu.set_tuning("step_rows_per_wave_f32", rows)
        gpu.set_tuning("step_waves_f32", waves)
        got = [g[:, 1:-1, 1:-1] for g in _run_step(case, gpu, dev, Ts, dtypes)]
        if base1 is None:
            base1 = got
        for n in range(len(Ts)):
            assert np.array_equal(got[n], base1[n]), (group, rows, waves, n)
    for n in range(2):                 # FP64 tracers do not depend on it
        assert np.array_equal(base1[n], base[n])
    gpu.set_tuning("vdiff_solver", 0)
    gpu.set_tuning("step_waves_f32", 8)


@pytest.mark.parametrize("nk,mix", [(2, 1.0), (3, 1.0), (9, 1.0), (16, 1.0), (17, 1.0), (33, 1.0),
                                    (64, 1.0), (75, 1.0), (75, 100.0), (80, 1.0), (81, 1.0)])
def test_fp32_twin_solver(oracle_lib, nk, mix):
    """vdiff_solver 1 (two wavefronts per column, levels in registers) at
    every register-slot size, odd and even nk (an inert slot on one half),
    shallow and full columns, weak and strong mixing (diff_cbt x 100: up to
    1 m2/s); nk = 81 is beyond the register solver and takes solver 0.  FP32
    tracers within the FP32 bound of the FP64 restatement, FP64 tracers within
    1e-12 per cell, and within the bound of solver 0's results."""
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    case = make_case(70, 18, nk, land_frac=0.3, nts=8, seed=100 + nk, n_passive=5,
                     stf_scale=1e-3)
    case.fields["diff_cbt"] = case.fields["diff_cbt"] * mix
    print(f"[twin solver 70x18x{nk} mix x{mix}]", flush=True)
    gpu = OceanGPU(case.dom, case.grid)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    got = {}
    for solver in (0, 1):
        gpu.set_tuning("vdiff_solver", solver)
        got[solver] = [g[:, 1:-1, 1:-1] for g in _run_step(case, gpu, dev, Ts, dtypes)]
    gpu.close()
    ref = [r[:, 1:-1, 1:-1] for r in _oracle_step(case, oracle_lib.Oracle(case), Ts)]
    for n in range(2):
        assert_close_cells(got[1][n], ref[n], what=f"tracer {n}")
    for n in range(2, len(Ts)):
        scale = np.abs(ref[n]).max()
        e1 = np.abs(got[1][n].astype(np.float64) - ref[n]).max() / scale
        e0 = np.abs(got[0][n].astype(np.float64) - ref[n]).max() / scale
        d10 = np.abs(got[1][n].astype(np.float64) - got[0][n]).max() / scale
        print(f"tracer {n}: solver 1 {e1:.2e}, solver 0 {e0:.2e}, |1 - 0| {d10:.2e}")
        assert e1 <= 1e-5 and d10 <= 2e-5, (n, e1, d10)
        if nk > 80:
            assert np.array_equal(got[1][n], got[0][n])


@pytest.mark.parametrize("cyclic", [True, False])
def test_fp32_large_groups_bitwise_and_accurate(oracle_lib, cyclic):
    """Ten FP32 tracers in one pass (1 row x 16 wavefronts), two passes of
    five (2 x 12) and the 4 + 4 + 2 passes of 2 x 16 tiles give the same
    bits; the result is within the FP32 bound of the FP64 restatement."""
    from mom5_amd import OceanGPU
    from mom5_amd.synth import make_case
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    case = make_case(100, 52, 10, land_frac=0.3, nts=10, seed=21, n_passive=10, stf_scale=1e-3,
                     **kw)
    print(f"[ten FP32 tracers 100x52x10 cyclic={cyclic}]", flush=True)
    from mom5_amd import DeviceArray
    gpu = OceanGPU(case.dom, case.grid)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    base = None
    for group, rows, waves in [(4, 2, 16), (10, 1, 16), (8, 1, 16), (6, 1, 16), (5, 2, 12)]:
        gpu.set_tuning("step_group_f32", group)
        gpu.set_tuning("step_rows_per_wave_f32", rows)
        gpu.set_tuning("step_waves_f32", waves)
        got = [g[:, 1:-1, 1:-1] for g in _run_step(case, gpu, dev, Ts, dtypes)]
        if base is None:
            base = got
        for n in range(len(Ts)):
            assert np.array_equal(got[n], base[n]), (group, rows, waves, n)
    ref = _oracle_step(case, oracle_lib.Oracle(case), Ts)
    for n in range(2, len(Ts)):
        r = ref[n][:, 1:-1, 1:-1]
        err = np.abs(base[n].astype(np.float64) - r).max() / np.abs(r).max()
        assert err <= 1e-5, (n, err)
    gpu.close()


@pytest.mark.parametrize("rows,waves", [(1, 16), (2, 8), (4, 8)])
def test_fp64_full_groups_of_four(oracle_lib, rows, waves):
    """Eight FP64 tracers at step_group_f64 = 4: two full passes of the
    four-tracer kernel, against the oracle's per-tracer sequence."""
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    from _cases import oracle_tracer_step
    case = make_case(97, 45, 9, land_frac=0.3, nts=8, seed=44, n_passive=6, stf_scale=1e-3)
    print(f"[8 FP64 tracers 97x45x9, group 4, {rows} rows x {waves} waves]", flush=True)
    gpu = OceanGPU(case.dom, case.grid)
    gpu.set_tuning("step_group_f64", 4)
    gpu.set_tuning("step_rows_per_wave", rows)
    gpu.set_tuning("step_waves_f64", waves)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    Ts = [t for _, t in case.tracers]
    assert len(Ts) == 8
    got = _run_step(case, gpu, dev, Ts, [np.float64] * 8)
    gpu.close()
    f = case.fields
    ref = oracle_tracer_step(oracle_lib.Oracle(case), Ts, f["stf"][:8], f["btf"][:8], case.dt)
    for n in range(8):
        assert_close_cells(got[n][:, 1:-1, 1:-1], ref[n][:, 1:-1, 1:-1], what=f"tracer {n}")


@pytest.mark.parametrize("nk", [2, 3])
def test_shallow_domains_default_shapes(oracle_lib, nk):
    """Two and three levels (the fewest the library and invtri take), where
    every level index the passes clamp (the tracers three levels down, w and
    rho_dzt one level ahead, the deferred update of the level before) meets the
    bottom: FP64 tracers bit for bit the per-tracer path and within 1e-12 per
    cell of the CPU restatement, FP32 tracers within the FP32 bound of it."""
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    case = make_case(70, 18, nk, land_frac=0.3, nts=8, seed=300 + nk, n_passive=3,
                     stf_scale=1e-3)
    print(f"[shallow domain 70x18x{nk}]", flush=True)
    gpu = OceanGPU(case.dom, case.grid)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    Ts = [t for _, t in case.tracers]
    dtypes = [np.float64, np.float64] + [np.float32] * (len(Ts) - 2)
    got = [g[:, 1:-1, 1:-1] for g in _run_step(case, gpu, dev, Ts, dtypes)]
    f = case.fields
    plain = []
    for T in Ts[:2]:
        dT, th, out = DeviceArray.from_host(T), gpu.field3(), gpu.field3()
        gpu.horz_advect_tracer("mdfl_sweby", dT, th, uhrho_et=dev["uhrho_et"],
                               vhrho_nt=dev["vhrho_nt"], wrho_bt=dev["wrho_bt"],
                               rho_dzt_tau=dev["rho_dzt"], dtime=case.dt)
        gpu.update_tracer_explicit(dT, th, dev["rho_dzt"], dev["rho_dztr"], case.dt, out)
        plain.append(out)
    gpu.vert_diffuse_implicit(plain, [DeviceArray.from_host(s) for s in f["stf"][:2]],
                              [DeviceArray.from_host(s) for s in f["btf"][:2]], dev["diff_cbt"],
                              dev["rho_dzt"], dev["dzwt"], case.dt, index_salt=1)
    gpu.sync()
    ref = [r[:, 1:-1, 1:-1] for r in _oracle_step(case, oracle_lib.Oracle(case), Ts)]
    for n in range(2):
        assert np.array_equal(got[n], plain[n].to_host()[:, 1:-1, 1:-1]), f"tracer {n} vs plain"
        assert_close_cells(got[n], ref[n], what=f"tracer {n} vs oracle")
    for n in range(2, len(Ts)):
        err = np.abs(got[n].astype(np.float64) - ref[n]).max() / np.abs(ref[n]).max()
        assert err <= 1e-5, (n, err)
    gpu.close()
