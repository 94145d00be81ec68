"""Whole steps over several time steps: the batched step against the plain
per-tracer path (both through the C-ABI), on host-built and device-built
cases.

* all-FP64: every tracer and the barotropic time-averaged state bit-for-bit
  equal after 3 steps (MDFL-Sweby for every tracer);
* mixed: T and S bit-for-bit equal to the plain FP64 path, FP32 passive
  tracers within 1e-5 normwise (max |err| / max |T|) and 1e-6 rms of the
  plain FP64 result after 3 steps;
* masked tracer content (compensated sums) conserved to round-off when the
  surface and bottom fluxes vanish.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 3


def _host(a):
    return a.cpu().numpy() if hasattr(a, "cpu") else a.to_host()


def test_batched_fp64_equals_plain_over_steps():
    from mom5_amd.step import BatchedStep, PlainStep
    from mom5_amd.synth import make_case
    case = make_case(96, 50, 10, land_frac=0.3, nts=12, seed=21, n_passive=2, stf_scale=1e-3)
    print("[case 96x50x10 host-built, 4 FP64 tracers]", flush=True)
    plain = PlainStep(case, schemes=["mdfl_sweby"] * 4)
    batched = BatchedStep(case, gpu=plain.gpu)
    plain.run(NSTEPS)
    batched.run(NSTEPS)
    plain.gpu.sync()
    for n in range(4):
        a = _host(plain.tracer(n))[:, 1:-1, 1:-1]
        b = _host(batched.tracer(n))[:, 1:-1, 1:-1]
        assert np.array_equal(a, b), f"tracer {n}"
    assert np.array_equal(_host(plain.eta_bar[plain.cur]), _host(batched.eta_bar[batched.cur]))
    assert np.array_equal(_host(plain.udrho[plain.cur]), _host(batched.udrho[batched.cur]))


@pytest.mark.parametrize("cyclic", [True, False])
def test_device_case_mixed_vs_plain_fp64(cyclic):
    import dataclasses
    from mom5_amd.step import BatchedStep, PlainStep
    from mom5_amd.synth import make_case_device
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    ni, nj, nk = 150, 90, 14
    case = make_case_device(ni, nj, nk, seed=5, nts=10, n_passive=3, **kw)
    print(f"[case {ni}x{nj}x{nk} device-built cyclic={cyclic}, 2 FP64 + 3 FP32]", flush=True)
    g = case.info["gpu"]
    names = [n for n, _ in case.tracers]
    batched = BatchedStep(case, gpu=g)
    rho = batched.dev["rho_dzt"]
    c0 = [g.tracer_content(batched.tracer(n), rho) for n in range(5)]
    batched.run(NSTEPS)
    c1 = [g.tracer_content(batched.tracer(n), rho) for n in range(5)]
    pcase = dataclasses.replace(case, tracers=[
        (names[n], case.info["tracer"](n, np.float64)) for n in range(5)])
    plain = PlainStep(pcase, schemes=["mdfl_sweby"] * 5, gpu=g)
    plain.run(NSTEPS)
    g.sync()
    kmt = case.grid["kmt"]
    wet = (np.arange(1, nk + 1)[:, None, None] <= kmt[None])[:, 1:-1, 1:-1]
    for n in range(5):
        ref = plain.tracer(n).to_host()[:, 1:-1, 1:-1]
        got = batched.tracer(n).to_host()[:, 1:-1, 1:-1]
        e = g.tracer_error(batched.tracer(n), plain.tracer(n))
        if n < 2:
            assert np.array_equal(got, ref), names[n]
            assert e["cells_differ"] == 0 and e["cells"] == int(wet.sum())
        else:
            r = ref[wet]
            err = np.abs(got.astype(np.float64)[wet] - r)
            normwise = err.max() / np.abs(r).max()
            rms = np.sqrt((err ** 2).mean() / (r ** 2).mean())
            print(f"{names[n]}: normwise {normwise:.2e}, rms {rms:.2e}")
            assert normwise <= 1e-5 and rms <= 1e-6
            # the device-side error report agrees with the host computation
            assert e["max_abs"] == pytest.approx(err.max(), rel=1e-12)
            assert e["normwise"] == pytest.approx(normwise, rel=1e-12)
            assert e["rms_rel"] == pytest.approx(rms, rel=1e-9)
            assert e["cells"] == r.size
        drift = abs(c1[n] - c0[n]) / abs(c0[n])
        print(f"{names[n]}: content drift {drift:.2e}")
        assert drift <= (1e-13 if n < 2 else 1e-8)


def test_device_case_matches_host_case_transports():
    """The device-built transports repeat make_case's construction: same
    time steps, uhrho_et/vhrho_nt/wrho_bt equal to rounding of sin."""
    from mom5_amd.synth import make_case, make_case_device
    for kw in ({}, dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0,
                        barotropic=False)):
        print(f"[transports 70x44x11 {kw}]", flush=True)
        c = make_case(70, 44, 11, seed=3, **kw)
        d = make_case_device(70, 44, 11, seed=3, **kw)
        assert c.dt == pytest.approx(d.dt, rel=1e-12) and c.dtbt == d.dtbt
        for k in ("uhrho_et", "vhrho_nt", "wrho_bt", "rho_dzt", "rho_dztr", "dzwt"):
            a, b = c.fields[k], d.fields[k].to_host()
            assert a.shape == b.shape, k
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), k
        T = d.tracers[0][1].to_host()
        tm = np.arange(1, 12)[:, None, None] <= c.grid["kmt"][None]
        assert np.all(T[~tm] == 0.0) and np.all(T[tm] != 0.0)
        d.info["gpu"].close()
