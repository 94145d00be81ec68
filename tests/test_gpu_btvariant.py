"""mom_gpu_update_ocean_barotropic_variant (B/C-grid x depth/pressure-based
split-explicit schemes and the time-mean tail) against the CPU
restatement, bit for bit (integer views), on a global cyclic and a closed
regional case; and the B-grid depth-based variant against the blocked
mom_gpu_update_ocean_barotropic."""
import numpy as np
import pytest

from _cases import bt_variant_fields

pytestmark = pytest.mark.gpu


def _case(cyclic):
    from mom5_amd.synth import make_case
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    return make_case(70, 44, 3, seed=61, land_frac=0.3, barotropic=True, **kw)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _inputs(c, cgrid, press, seed):
    rng = np.random.default_rng(seed)
    fl = bt_variant_fields(c, rng, cgrid, scale=0.3)
    sh = fl["patm"].shape
    tm = c.grid["tmask"][0]
    fl["x_tau"] = fl["x_tau"] + 0.01 * rng.standard_normal(sh) * tm
    fl["river"] = 1e-4 * rng.random(sh) * tm
    fl["source"] = 1e-5 * rng.standard_normal(sh) * tm
    fl["smooth"] = 1e-6 * rng.standard_normal(sh) * tm
    fl["patm"] = 1.0e5 + 50.0 * rng.standard_normal(sh)
    fl["rho_g"] = 9.80 * (1025.0 + rng.standard_normal(sh))
    fl["dpatm_dt"] = 1e-2 * rng.standard_normal(sh)
    if press:
        fl["x_tau"] = 1035.0 * 9.80 * fl["x_tau"]
        fl["xbar_tau"] = 1035.0 * 9.80 * fl["xbar_tau"]
        fl["thick"] = 1035.0 * fl["thick"]
    return fl


def _run_gpu(c, fl, cgrid, press, nts, stepping_a, initsum):
    from mom5_amd import DeviceArray, OceanGPU
    H = DeviceArray.from_host
    g = OceanGPU(c.dom, c.grid)
    sh = fl["patm"].shape
    xb, ud, cv = DeviceArray(sh), DeviceArray((2,) + sh), DeviceArray(sh)
    kw = {k: H(v) for k, v in fl.items() if v is not None and k != "udrho" and k != "xbar_tau"}
    g.update_ocean_barotropic_variant(
        cgrid=cgrid, pressure_based=press, nts=nts, dtbt=c.dtbt, x_bar_tau=H(fl["xbar_tau"]),
        udrho_tau=H(fl["udrho"]), x_bar_taup1=xb, udrho_taup1=ud, conv_rho_ud_t=cv,
        stepping_a=stepping_a, initsum_with_bar=initsum, **kw)
    g.sync()
    out = (xb.to_host(), ud.to_host(), cv.to_host())
    g.close()
    return out


@pytest.mark.parametrize("cyclic", [True, False])
@pytest.mark.parametrize("cgrid,press", [(False, False), (True, False), (False, True),
                                         (True, True)])
def test_variant_bitwise(oracle_lib, cyclic, cgrid, press):
    c = _case(cyclic)
    for stepping_a, initsum in ((False, True), (True, False)):
        print(f"[barotropic variant 70x44 cyclic={cyclic} cgrid={cgrid} press={press} "
              f"stepping_a={stepping_a} initsum={initsum}]", flush=True)
        fl = _inputs(c, cgrid, press, seed=7 + 2 * cgrid + press)
        want = oracle_lib.Oracle(c).update_ocean_barotropic_variant(
            9, c.dtbt, cgrid=cgrid, press=press, fields=fl, stepping_a=stepping_a,
            initsum_with_bar=initsum)
        got = _run_gpu(c, fl, cgrid, press, 9, stepping_a, initsum)
        for name, a, b in zip(("x_bar", "udrho", "conv_rho_ud_t"), got, want):
            assert np.array_equal(_bits(a), _bits(b)), name


def test_bgrid_depth_variant_equals_blocked_path():
    from mom5_amd import DeviceArray, OceanGPU
    c = _case(True)
    print("[barotropic variant vs blocked path 70x44]", flush=True)
    f = c.fields
    fl = dict(x_tau=f["eta_t"], xbar_tau=f["eta_t_bar"], udrho=f["udrho"],
              forcing_bt=f["forcing_bt"], thick=f["thicku"], patm=f["patm"], rho_g=f["rho_g"],
              pme=f["pme"], river=f["river"], source=f["source"], smooth=f["eta_smooth"])
    got = _run_gpu(c, fl, False, False, c.nts, False, True)
    H = DeviceArray.from_host
    g = OceanGPU(c.dom, c.grid)
    sh = f["patm"].shape
    xb, ud, cv = DeviceArray(sh), DeviceArray((2,) + sh), DeviceArray(sh)
    g.update_ocean_barotropic(
        nts=c.nts, dtbt=c.dtbt, eta_t_bar_tau=H(f["eta_t_bar"]), udrho_tau=H(f["udrho"]),
        forcing_bt=H(f["forcing_bt"]), thicku=H(f["thicku"]), patm=H(f["patm"]),
        pme=H(f["pme"]), river=H(f["river"]), source=H(f["source"]),
        eta_smooth=H(f["eta_smooth"]), rho_g=H(f["rho_g"]), eta_t_bar_taup1=xb,
        udrho_taup1=ud, conv_rho_ud_t=cv)
    g.sync()
    for name, a, b in zip(("eta_t_bar", "udrho", "conv"), got,
                          (xb.to_host(), ud.to_host(), cv.to_host())):
        assert np.array_equal(_bits(a), _bits(b)), name
    g.close()
