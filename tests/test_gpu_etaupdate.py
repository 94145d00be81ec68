"""mom_gpu_eta_and_pbot_update and mom_gpu_barotropic_closure against the
CPU restatement, bit for bit (integer views of the FP64 results), on a
global cyclic and a closed regional case, depth- and pressure-based, B- and
C-grid gradients."""
import numpy as np
import pytest

from _cases import bt_tmasken

pytestmark = pytest.mark.gpu


def _case(cyclic):
    from mom5_amd.synth import make_case
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    return make_case(66, 40, 6, seed=43, land_frac=0.3, barotropic=False, **kw)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _ucell(rng, shape):
    NJ, NI = shape
    dte, dtw, dus, dun = (rng.uniform(2.0e4, 6.0e4, shape) for _ in range(4))
    area = np.ones(shape)
    area[:-1, :-1] = (dte[:-1, :-1] * dus[:-1, :-1] + dtw[:-1, 1:] * dus[:-1, :-1]
                      + dte[1:, :-1] * dun[:-1, :-1] + dtw[1:, 1:] * dun[:-1, :-1])
    return dict(dte=dte, dtw=dtw, dus=dus, dun=dun, daur=1.0 / area)


@pytest.mark.parametrize("cyclic", [True, False])
@pytest.mark.parametrize("pressure", [False, True])
def test_eta_and_pbot_update_bitwise(oracle_lib, cyclic, pressure):
    from mom5_amd import DeviceArray, OceanGPU
    c = _case(cyclic)
    print(f"[eta_and_pbot_update 66x40 cyclic={cyclic} pressure={pressure}]", flush=True)
    sh = (c.nj + 2, c.ni + 2)
    rng = np.random.default_rng(11)
    x0 = rng.standard_normal(sh) * (1.0e4 if pressure else 0.5)
    conv = rng.standard_normal(sh) * 1e-3
    pme = rng.standard_normal(sh) * 1e-4
    river = rng.random(sh) * 1e-4
    src = rng.standard_normal(sh) * 1e-5
    dpatm = rng.standard_normal(sh) * 1e-2 if pressure else None
    ht = rng.uniform(50.0, 5500.0, sh) if pressure else None
    prev = rng.standard_normal(sh)
    want = oracle_lib.Oracle(c).eta_and_pbot_update(
        x0, conv, pme, river, src, 1800.0, pressure_based=pressure, dpatm_dt=dpatm, ht=ht,
        x_taup1=prev)
    g = OceanGPU(c.dom, c.grid)
    H = DeviceArray.from_host
    xd, td = H(prev), DeviceArray(sh)
    pb = DeviceArray(sh) if pressure else None
    g.eta_and_pbot_update(H(x0), xd, td, conv_rho_ud_t_tau=H(conv), pme=H(pme), river=H(river),
                          source=H(src), dtime=1800.0, pressure_based=pressure,
                          dpatm_dt=H(dpatm) if pressure else None, ht=H(ht) if pressure else None,
                          pbot_t_taup1=pb)
    g.sync()
    assert np.array_equal(_bits(xd.to_host()), _bits(want[0]))
    assert np.array_equal(_bits(td.to_host()), _bits(want[1]))
    if pressure:
        assert np.array_equal(_bits(pb.to_host()), _bits(want[2]))
    g.close()


@pytest.mark.parametrize("cyclic", [True, False])
@pytest.mark.parametrize("pressure,cgrid", [(False, False), (True, False), (False, True)])
def test_barotropic_closure_bitwise(oracle_lib, cyclic, pressure, cgrid):
    from mom5_amd import DeviceArray, OceanGPU
    c = _case(cyclic)
    print(f"[barotropic closure 66x40 cyclic={cyclic} pressure={pressure} cgrid={cgrid}]",
          flush=True)
    sh = (c.nj + 2, c.ni + 2)
    rng = np.random.default_rng(12)
    uc = _ucell(rng, sh)
    eta = rng.standard_normal(sh) * 0.5
    ab = rng.standard_normal(sh) * 1.0e4
    patm = 1.0e5 + rng.standard_normal(sh) * 100.0
    rho_g = 9.80 * (1025.0 + rng.standard_normal(sh))
    ht = rng.uniform(50.0, 5500.0, sh)
    tm = bt_tmasken(c) if cgrid else None
    want = oracle_lib.Oracle(c).barotropic_closure(
        eta, patm, rho_g, uc, anompb_bar=ab, pressure_based=pressure, ht=ht, tmasken=tm)
    g = OceanGPU(c.dom, c.grid)
    g.set_remap_metrics(**uc)
    H = DeviceArray.from_host
    eu, ps = DeviceArray(sh), DeviceArray(sh)
    gps, gab = DeviceArray((2,) + sh), DeviceArray((2,) + sh)
    g.barotropic_closure(H(eta), H(patm), H(rho_g), eu, ps, gps, anompb_bar=H(ab),
                         grad_anompb=gab, pressure_based=pressure, ht=H(ht),
                         tmasken=H(tm) if cgrid else None)
    g.sync()
    for name, got, w in zip(("eta_u", "ps", "grad_ps", "grad_anompb"), (eu, ps, gps, gab), want):
        assert np.array_equal(_bits(got.to_host()), _bits(w)), name
    g.close()


def test_closure_needs_metrics():
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd._lib import MomGpuError
    c = _case(True)
    g = OceanGPU(c.dom, c.grid)
    sh = (c.nj + 2, c.ni + 2)
    z = DeviceArray(sh)
    with pytest.raises(MomGpuError, match="REMAP_BT_TO_BU metrics not set"):
        g.barotropic_closure(z, z, z, DeviceArray(sh), DeviceArray(sh), DeviceArray((2,) + sh))
    g.close()
