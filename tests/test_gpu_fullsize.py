"""Full-size checks on the GPU, at shapes the CPU restatement cannot follow in
a test's time, through properties that hold at any size:

* 1440x1080x50 (T and S MDFL-Sweby, a Quicker tracer, 80 barotropic
  subcycles): two steps of the batched step (mom_gpu_step_tracers) equal two
  steps of the per-tracer path bit for bit, for every tracer and for the
  barotropic state; the per-tracer path itself is checked against the CPU
  restatement at the smaller sizes of test_gpu_parity.py.
* 3600x2700 with 120 subcycles: the temporally blocked barotropic (5
  subcycles a launch) equals the per-subcycle kernels bit for bit at every
  value of eta_t_bar(taup1), udrho(taup1) and conv_rho_ud_t, and eta_t_bar
  stays bounded (initial eta ~ N(0, 0.1 m)).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(a.view(np.int64), b.view(np.int64))


def test_batched_step_equals_plain_path_1440x1080x50():
    from mom5_amd.synth import make_case_device
    from mom5_amd.step import BatchedStep, PlainStep
    print("[full size: 1440x1080x50, MDFL, MDFL, Quicker, nts=80, two steps]", flush=True)
    case = make_case_device(1440, 1080, 50, land_frac=0.25, nts=80, n_passive=1,
                            passive_dtype=np.float64, seed=0x4D4F4D35)
    g = case.info["gpu"]
    schemes = ["mdfl_sweby", "mdfl_sweby", "quicker"]
    plain = PlainStep(case, schemes=schemes, gpu=g)
    fresh = [(name, case.info["tracer"](n, np.float64)) for n, (name, _) in
             enumerate(case.tracers)]
    batched = BatchedStep(case, gpu=g, tracers=fresh, schemes=schemes)
    plain.run(2)
    batched.run(2)
    g.sync()
    for n, name in enumerate(plain.names):
        a = plain.tracer(n).to_host()[:, 1:-1, 1:-1]
        b = batched.tracer(n).to_host()[:, 1:-1, 1:-1]
        assert np.isfinite(a).all(), name
        assert _bits_equal(a, b), f"{name}: batched step differs from the per-tracer path"
    for what in ("eta_bar", "udrho"):
        a = getattr(plain, what)[plain.cur].to_host()
        b = getattr(batched, what)[batched.cur].to_host()
        assert _bits_equal(a, b), what
    g.close()


def test_blocked_barotropic_equals_subcycles_3600x2700():
    from mom5_amd import DeviceArray
    from mom5_amd.synth import make_case_device
    from mom5_amd.step import BatchedStep
    print("[full size: barotropic 3600x2700, nts=120, blocked H=5 vs per subcycle]", flush=True)
    case = make_case_device(3600, 2700, 4, land_frac=0.30, nts=120, n_passive=0)
    g = case.info["gpu"]
    bs = BatchedStep(case, gpu=g)
    res = {}
    for h in (5, 0):
        g.set_tuning("barotropic_subcycles_per_launch", h)
        eb, cv = DeviceArray(g.shape2), DeviceArray(g.shape2)
        ud = DeviceArray((2,) + g.shape2)
        g.update_ocean_barotropic(**dict(bs.bt_args(), eta_t_bar_taup1=eb, udrho_taup1=ud,
                                         conv_rho_ud_t=cv))
        g.sync()
        res[h] = [x.to_host() for x in (eb, ud, cv)]
    for name, a, b in zip(("eta_t_bar", "udrho", "conv_rho_ud_t"), res[5], res[0]):
        assert _bits_equal(a, b), f"{name}: blocked differs from the per-subcycle path"
    eta = res[5][0]
    assert np.isfinite(eta).all() and np.abs(eta).max() < 10.0
    g.close()
