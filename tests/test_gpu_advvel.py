"""mom_gpu_advection_velocity against the CPU restatement of
ocean_advection_velocity (B-grid branch), bit for bit, on a global cyclic
and a closed regional case, with and without the thickness-source terms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["cyclic_70x44x9", "regional_41x29x7"])
@pytest.mark.parametrize("sources", [True, False])
def test_advection_velocity_bitwise(oracle_lib, kind, sources):
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    if kind == "cyclic_70x44x9":
        c = make_case(70, 44, 9, seed=31, land_frac=0.3, barotropic=False)
    else:
        c = make_case(41, 29, 7, seed=32, land_frac=0.2, cyclic_x=False, lat_s=20.0, lat_n=50.0,
                      lon_span=30.0, barotropic=False)
    print(f"[advection velocity {kind} sources={sources}]", flush=True)
    nk, NJ, NI = c.nk, c.nj + 2, c.ni + 2
    rng = np.random.default_rng(7)
    u = rng.standard_normal((2, nk, NJ, NI)) * 0.2
    rdu = 1035.0 * c.grid["dzt"][:, None, None] * (1.0 + 0.01 * rng.random((nk, NJ, NI)))
    pme = 1e-4 * rng.standard_normal((NJ, NI))
    river = 1e-4 * rng.random((NJ, NI))
    rt = rng.standard_normal((nk, NJ, NI)) * 1e-3 if sources else None
    ms = rng.standard_normal((nk, NJ, NI)) * 1e-3 if sources else None
    ref = oracle_lib.Oracle(c).advection_velocity(u, rdu, pme, river, rt, ms)
    g = OceanGPU(c.dom, c.grid)
    d = {k: DeviceArray.from_host(v) for k, v in
         dict(u=u, rdu=rdu, pme=pme, river=river).items()}
    srcs = [DeviceArray.from_host(a) if a is not None else None for a in (rt, ms)]
    ue, vn, div = g.field3(), g.field3(), g.field3()
    wb = DeviceArray((nk + 1, NJ, NI))
    g.advection_velocity(d["u"], d["rdu"], d["pme"], d["river"], ue, vn, wb,
                         rho_dzt_tendency=srcs[0], mass_source=srcs[1], diverge_t=div)
    g.sync()
    for name, got, want in zip(("uhrho_et", "vhrho_nt", "wrho_bt", "diverge_t"),
                               (ue, vn, wb, div), ref):
        assert np.array_equal(got.to_host(), want), name
    g.close()


@pytest.mark.parametrize("cyclic", [True, False])
def test_barotropic_forcing_bitwise(oracle_lib, cyclic):
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    c = make_case(66, 40, 8, seed=41, land_frac=0.3, barotropic=False, **kw)
    print(f"[barotropic forcing 66x40x8 cyclic={cyclic}]", flush=True)
    nk, NJ, NI = c.nk, c.nj + 2, c.ni + 2
    rng = np.random.default_rng(5)
    accel = rng.standard_normal((2, nk, NJ, NI))
    smf = rng.standard_normal((2, NJ, NI))
    bmf = rng.standard_normal((2, NJ, NI))
    ref = oracle_lib.Oracle(c).barotropic_forcing(accel, smf, bmf)
    g = OceanGPU(c.dom, c.grid)
    out = DeviceArray((2, NJ, NI))
    g.barotropic_forcing(DeviceArray.from_host(accel), DeviceArray.from_host(smf),
                         DeviceArray.from_host(bmf), out)
    g.sync()
    assert np.array_equal(out.to_host(), ref)
    g.close()
