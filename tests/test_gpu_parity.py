"""GPU parity: every C-ABI entry point against the CPU restatement (oracle/)
on the same seeded inputs.  FP64 results are bit-identical (the kernels and
the oracle share the reference's evaluation order and neither fuses
multiply-adds) except MDFL-Sweby / DST-linear advection, whose kernels form
the limited flux without its divisions and the update factors with
reciprocals: their tendencies and fluxes must agree to 1e-12 of their scale
(rho_dzt*max|T|/dtime per cell; the largest flux of the level); and
vert_diffuse_implicit, whose FP64 solver eliminates each column from both
ends (1e-12 relative per cell; invtri itself keeps the reference's order and
is compared bit for bit).  The compensated content sum must agree to 1e-15
relative.
"""
import numpy as np
import pytest
from _cases import assert_close_cells, assert_close_scaled

pytestmark = pytest.mark.gpu

CASES = {
    # name: (ni, nj, nk, kwargs)
    "small_90x60x25": (90, 60, 25, dict(land_frac=0.25, barotropic=False)),
    "odd_cyclic_37x23x7": (37, 23, 7, dict(land_frac=0.3, nts=12, seed=11)),
    "regional_41x29x9": (41, 29, 9, dict(land_frac=0.2, cyclic_x=False, lat_s=20.0,
                                         lat_n=50.0, lon_span=30.0, nts=16, seed=12)),
    "tiles_130x70x12": (130, 70, 12, dict(land_frac=0.3, nts=10, seed=13, n_passive=1)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def setup(request, oracle_lib):
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    ni, nj, nk, kw = CASES[request.param]
    print(f"[case {request.param}]", flush=True)
    case = make_case(ni, nj, nk, stf_scale=1e-3, pme_scale=1e-4, **kw)
    orc = oracle_lib.Oracle(case)
    gpu = OceanGPU(case.dom, case.grid)
    dev = {k: DeviceArray.from_host(v) for k, v in case.fields.items()
           if isinstance(v, np.ndarray)}
    yield case, orc, gpu, dev
    gpu.close()


def same(a, b):
    """Bit-level equality up to the sign of zero (NaN fails)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape
    eq = (a == b)
    if not eq.all():
        bad = np.argwhere(~eq)
        i = tuple(bad[0])
        raise AssertionError(f"{len(bad)} mismatches, first at {i}: {a[i]!r} vs {b[i]!r} "
                             f"(max abs diff {np.nanmax(np.abs(a - b)):.3e})")


@pytest.mark.parametrize("rows_per_wave", [4, 2, 1])
@pytest.mark.parametrize("scheme,lim", [("mdfl_sweby", 1.0), ("dst_linear", 0.0)])
def test_horz_advect_mdfl(setup, scheme, lim, rows_per_wave):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    gpu.set_tuning("mdfl_rows_per_wave", rows_per_wave)
    rng = np.random.default_rng(5)
    for name, T in case.tracers[:2]:
        th0 = rng.standard_normal(orc.shape3)
        th_ref = th0.copy()
        wrk_ref = orc.horz_advect_mdfl(T, th_ref, case.dt, lim=lim)
        _, (fx_ref, fy_ref, fz_ref) = orc.advect_mdfl_sweby(T, case.dt, lim=lim, want_flux=True)
        dT = DeviceArray.from_host(T)
        dth = DeviceArray.from_host(th0)
        wrk = gpu.field3()
        fx, fy, fz = gpu.field3(), gpu.field3(), gpu.field3()
        gpu.horz_advect_tracer(scheme, dT, dth, uhrho_et=dev["uhrho_et"], vhrho_nt=dev["vhrho_nt"],
                               wrho_bt=dev["wrho_bt"], rho_dzt_tau=dev["rho_dzt"], dtime=case.dt,
                               wrk1=wrk, flux_x=fx, flux_y=fy, flux_z=fz)
        gpu.sync()
        th_scale = case.fields["rho_dzt"] * np.abs(T).max() / case.dt
        got_th = dth.to_host()
        assert_close_scaled(got_th, th_ref, th_scale, what=f"{name} th_tendency")
        assert_close_scaled(wrk.to_host(), wrk_ref, th_scale, what=f"{name} wrk1")
        for lab, a, b in (("flux_x", fx.to_host(), fx_ref), ("flux_y", fy.to_host(), fy_ref),
                          # flux_z is only defined on the compute domain
                          ("flux_z", fz.to_host()[:, 1:-1, 1:-1], fz_ref[:, 1:-1, 1:-1])):
            lev = np.abs(b).max(axis=(1, 2), keepdims=True)
            assert_close_scaled(a, b, np.where(lev > 0, lev, 1.0), what=f"{name} {lab}")
        # vertical part of a 3-D scheme is a no-op
        gpu.vert_advect_tracer(scheme, dT, dth)
        gpu.sync()
        same(dth.to_host(), got_th)


def test_advect_sweby_all(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    Ts = [t for _, t in case.tracers]
    rng = np.random.default_rng(6)
    th0 = [rng.standard_normal(orc.shape3) for _ in Ts]
    th_ref = [a.copy() for a in th0]
    wrk_ref = orc.sweby_all(Ts, th_ref, case.dt)
    dTs = [DeviceArray.from_host(t) for t in Ts]
    dth = [DeviceArray.from_host(a) for a in th0]
    dw = [gpu.field3() for _ in Ts]
    gpu.advect_tracer_sweby_all(dTs, dth, uhrho_et=dev["uhrho_et"], vhrho_nt=dev["vhrho_nt"],
                                wrho_bt=dev["wrho_bt"], rho_dzt_tau=dev["rho_dzt"], dtime=case.dt,
                                wrk1s=dw)
    gpu.sync()
    for n in range(len(Ts)):
        th_scale = case.fields["rho_dzt"] * np.abs(Ts[n]).max() / case.dt
        assert_close_scaled(dth[n].to_host(), th_ref[n], th_scale, what=f"th {n}")
        assert_close_scaled(dw[n].to_host(), wrk_ref[n], th_scale, what=f"wrk1 {n}")


def _limits(case):
    mins = [2.0, 34.8] + [-0.1] * (len(case.tracers) - 2)
    maxs = [20.0, 35.2] + [0.01] * (len(case.tracers) - 2)
    return mins, maxs


def test_compute_tmask_limit(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    Ts = [t for _, t in case.tracers]
    mins, maxs = _limits(case)
    ref = orc.tmask_limit(Ts, mins, maxs, 0, 1, True)
    dTs = [DeviceArray.from_host(t) for t in Ts]
    outs = [gpu.field3() for _ in Ts]
    gpu.compute_tmask_limit(dTs, mins, maxs, outs, 0, 1, True)
    gpu.sync()
    for n in range(len(Ts)):
        same(outs[n].to_host(), ref[n])
    assert ref[0].sum() > 0, "limits chosen so that some cells are flagged"


@pytest.mark.parametrize("ntr", [4, 7])
def test_compute_tmask_limit_many_tracers(setup, ntr):
    """Four and seven tracers (the kernel takes several tracers' loads at
    once: whole groups of four, and a remainder), each with its own range, T
    and S masks joined: the flags equal the CPU restatement's bit for bit."""
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    base = [t for _, t in case.tracers]
    rng = np.random.default_rng(ntr)
    Ts = [base[n % len(base)] * (1.0 + 0.1 * n) + rng.standard_normal(base[0].shape) * 0.05
          for n in range(ntr)]
    mins = [float(np.quantile(t, 0.1)) for t in Ts]
    maxs = [float(np.quantile(t, 0.9)) for t in Ts]
    ref = orc.tmask_limit(Ts, mins, maxs, 0, 1, True)
    dTs = [DeviceArray.from_host(t) for t in Ts]
    outs = [gpu.field3() for _ in Ts]
    gpu.compute_tmask_limit(dTs, mins, maxs, outs, 0, 1, True)
    gpu.sync()
    for n in range(ntr):
        same(outs[n].to_host(), ref[n])
    assert all(r.sum() > 0 for r in ref)


@pytest.mark.parametrize("limit_with_upwind", [False, True])
def test_quicker(setup, limit_with_upwind):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    Ts = [t for _, t in case.tracers]
    mins, maxs = _limits(case)
    lim = orc.tmask_limit(Ts, mins, maxs, 0, 1, True)
    T = Ts[0]
    rng = np.random.default_rng(7)
    Ttau = T + 0.01 * rng.standard_normal(T.shape) * case.grid["tmask"]
    res_h, (fx_ref, fy_ref) = orc.horz_quicker(T, Ttau, lim[0], limit_with_upwind, want_flux=True)
    res_v, fz_ref = orc.vert_quicker(T, Ttau, lim[0], want_flux=True)
    th0 = rng.standard_normal(orc.shape3)
    th_ref = th0.copy()
    th_ref[:, 1:-1, 1:-1] += -res_h
    wrk_h = np.zeros(orc.shape3)
    wrk_h[:, 1:-1, 1:-1] = -res_h
    dT, dTt = DeviceArray.from_host(T), DeviceArray.from_host(Ttau)
    dlim = DeviceArray.from_host(lim[0])
    dth = DeviceArray.from_host(th0)
    wrk, fx, fy, fz = gpu.field3(), gpu.field3(), gpu.field3(), gpu.field3()
    gpu.horz_advect_tracer("quicker", dT, dth, T_tau=dTt, uhrho_et=dev["uhrho_et"],
                           vhrho_nt=dev["vhrho_nt"], tmask_limit=dlim,
                           limit_with_upwind=limit_with_upwind, wrk1=wrk, flux_x=fx, flux_y=fy)
    gpu.sync()
    same(dth.to_host(), th_ref)
    same(wrk.to_host(), wrk_h)
    same(fx.to_host(), fx_ref)
    same(fy.to_host(), fy_ref)
    th_ref[:, 1:-1, 1:-1] += -res_v
    wrk_v = np.zeros(orc.shape3)
    wrk_v[:, 1:-1, 1:-1] = -res_v
    gpu.vert_advect_tracer("quicker", dT, dth, T_tau=dTt, wrho_bt=dev["wrho_bt"],
                           tmask_limit=dlim, wrk1=wrk, flux_z=fz)
    gpu.sync()
    same(dth.to_host(), th_ref)
    same(wrk.to_host(), wrk_v)
    same(fz.to_host()[:, 1:-1, 1:-1], fz_ref[:, 1:-1, 1:-1])


def test_update_tracer_explicit(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    T = case.tracers[0][1]
    rng = np.random.default_rng(8)
    th = rng.standard_normal(orc.shape3)
    ref = orc.update_explicit(T, th, case.dt)
    out = gpu.field3()
    gpu.update_tracer_explicit(DeviceArray.from_host(T), DeviceArray.from_host(th),
                               dev["rho_dzt"], dev["rho_dztr"], case.dt, out)
    gpu.sync()
    same(out.to_host()[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1])


def test_invtri(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    f = case.fields
    z0 = case.tracers[0][1].copy()
    dcb = 1035.0 * f["diff_cbt"][0]
    ref = orc.invtri(z0.copy(), f["stf"][0], f["btf"][0], dcb, case.dt, f["rho_dzt"], f["dzwt"])
    dz = DeviceArray.from_host(z0)
    kmt = DeviceArray.from_host(case.grid["kmt"].astype(np.int32))
    gpu.invtri(dz, DeviceArray.from_host(f["stf"][0]), DeviceArray.from_host(f["btf"][0]),
               DeviceArray.from_host(dcb), case.dt, kmt, DeviceArray.from_host(case.grid["tmask"]),
               dev["rho_dzt"], dev["dzwt"], 1.0)
    gpu.sync()
    got = dz.to_host()
    same(got[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1])
    same(got[:, 0, :], z0[:, 0, :])   # halo untouched


def test_vert_diffuse_implicit(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    f = case.fields
    Ts = [t.copy() for _, t in case.tracers]
    ref = orc.vert_diffuse_implicit([t.copy() for t in Ts], f["stf"], f["btf"], case.dt,
                                    index_salt=1)
    dTs = [DeviceArray.from_host(t) for t in Ts]
    dstf = [DeviceArray.from_host(s) for s in f["stf"]]
    dbtf = [DeviceArray.from_host(s) for s in f["btf"]]
    gpu.vert_diffuse_implicit(dTs, dstf, dbtf, dev["diff_cbt"], dev["rho_dzt"], dev["dzwt"],
                              case.dt, index_salt=1)
    gpu.sync()
    # FP64 tracers take the register-resident two-ended elimination for
    # 2 <= nk <= 80: the reference's system and expressions, another order
    for n in range(len(Ts)):
        assert_close_cells(dTs[n].to_host()[:, 1:-1, 1:-1], ref[n][:, 1:-1, 1:-1],
                           what=f"tracer {n}")


@pytest.mark.parametrize("per_launch", [0, 1, 4, 7, 12])
@pytest.mark.parametrize("stepping_a,initsum", [(False, True), (True, False)])
def test_update_ocean_barotropic(setup, stepping_a, initsum, per_launch):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    if case.dtbt == 0.0:
        pytest.skip("advection-only case")
    gpu.set_tuning("barotropic_subcycles_per_launch", per_launch)
    nts = case.nts
    eta_ref, ud_ref, conv_ref = orc.update_ocean_barotropic(nts, case.dtbt, stepping_a=stepping_a,
                                                            initsum_with_bar=initsum)
    shp = orc.shape2
    eb = DeviceArray(shp)
    ud = DeviceArray((2,) + shp)
    cv = DeviceArray(shp)
    f = dev
    gpu.update_ocean_barotropic(nts=nts, dtbt=case.dtbt, eta_t_tau=f["eta_t"],
                                eta_t_bar_tau=f["eta_t_bar"], udrho_tau=f["udrho"],
                                forcing_bt=f["forcing_bt"], thicku=f["thicku"], patm=f["patm"],
                                pme=f["pme"], river=f["river"], source=f["source"],
                                eta_smooth=f["eta_smooth"], rho_g=f["rho_g"], eta_t_bar_taup1=eb,
                                udrho_taup1=ud, conv_rho_ud_t=cv, stepping_a=stepping_a,
                                initsum_with_bar=initsum)
    gpu.sync()
    same(eb.to_host(), eta_ref)
    same(ud.to_host(), ud_ref)
    same(cv.to_host(), conv_ref)


@pytest.mark.parametrize("exact", [True, False])
def test_blocked_barotropic_grid_reciprocals(oracle_lib, exact):
    """The blocked barotropic forms dxur, dyur as 1/(dxu + epsln), 1/(dyu +
    epsln) (ocean_grids.F90:949-951) only when the context's grid holds those
    values bit for bit, and loads them otherwise: with the grid's own
    reciprocals, and with a third of them moved by 2**-40 relative (which
    forming them would not reproduce; the halo columns then take the cyclic
    wrap, the barotropic precondition), the result equals the CPU
    restatement, which reads the grid's arrays, bit for bit."""
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    case = make_case(60, 40, 4, land_frac=0.25, nts=12, seed=77, pme_scale=1e-4)
    if not exact:
        rng = np.random.default_rng(3)
        for key in ("dxur", "dyur"):
            a = case.grid[key].copy()
            pick = rng.random(a.shape) < 0.3
            a[pick] = a[pick] * (1.0 + 2.0 ** -40)
            a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
            case.grid[key] = a
    print(f"[blocked barotropic, grid reciprocals exact={exact}]", flush=True)
    orc = oracle_lib.Oracle(case)
    gpu = OceanGPU(case.dom, case.grid)
    f = {k: DeviceArray.from_host(v) for k, v in case.fields.items() if isinstance(v, np.ndarray)}
    gpu.set_tuning("barotropic_subcycles_per_launch", 5)
    eta_ref, ud_ref, conv_ref = orc.update_ocean_barotropic(case.nts, case.dtbt)
    shp = orc.shape2
    eb, ud, cv = DeviceArray(shp), DeviceArray((2,) + shp), DeviceArray(shp)
    gpu.update_ocean_barotropic(nts=case.nts, dtbt=case.dtbt, eta_t_tau=f["eta_t"],
                                eta_t_bar_tau=f["eta_t_bar"], udrho_tau=f["udrho"],
                                forcing_bt=f["forcing_bt"], thicku=f["thicku"], patm=f["patm"],
                                pme=f["pme"], river=f["river"], source=f["source"],
                                eta_smooth=f["eta_smooth"], rho_g=f["rho_g"], eta_t_bar_taup1=eb,
                                udrho_taup1=ud, conv_rho_ud_t=cv)
    gpu.sync()
    same(eb.to_host(), eta_ref)
    same(ud.to_host(), ud_ref)
    same(cv.to_host(), conv_ref)
    gpu.close()


def _land_tiles(case, nsub, cyclic=True):
    """Tiles of the blocked barotropic (64x32 regions, nsub subcycles) whose
    every point is inert: both masks 0, forcing and velocity zeros."""
    g, f = case.grid, case.fields
    inert = ((g["kmt"] == 0) & (g["kmu"] == 0) & (f["forcing_bt"][0] == 0.0)
             & (f["forcing_bt"][1] == 0.0) & (f["udrho"][0] == 0.0) & (f["udrho"][1] == 0.0))
    NJ, NI = inert.shape
    ni = NI - 2
    TX, TY, LO = 64 - 2 * nsub - 1, 32 - 2 * nsub - 1, nsub + 1
    xs = range(1, 1 + ni, TX) if cyclic else range(0, NI, TX)
    n = 0
    for y0 in range(0, NJ, TY):
        for x0 in xs:
            cols = np.arange(x0 - LO, x0 - LO + 64)
            cols = 1 + (cols - 1) % ni if cyclic else cols[(cols >= 0) & (cols < NI)]
            rows = np.arange(y0 - LO, y0 - LO + 32)
            rows = rows[(rows >= 0) & (rows < NJ)]
            n += bool(inert[np.ix_(rows, cols)].all())
    return n


@pytest.mark.parametrize("stepping_a,cyclic", [(False, True), (True, True), (False, False)])
def test_blocked_barotropic_land_tiles(oracle_lib, stepping_a, cyclic):
    """On a grid with 60% land, regions of the blocked barotropic made of
    inert points only (both masks 0, zero forcing and velocity) take the
    short path that only accumulates the height: with it and without it
    (launches of 5, 5 and 2 subcycles, so the tile flags are formed for two
    tilings) the result equals the CPU restatement bit for bit, signed zeros
    included (eta, sf and the forcing hold zeros of both signs on land), on
    a cyclic domain and on a closed basin."""
    from mom5_amd import DeviceArray, OceanGPU
    from mom5_amd.synth import make_case
    case = make_case(600, 300, 4, land_frac=0.6, nts=12, seed=5, pme_scale=1e-4,
                     cyclic_x=cyclic)
    assert _land_tiles(case, 5, cyclic) > 0 and _land_tiles(case, 2, cyclic) > 0
    print(f"[blocked barotropic, land tiles, stepping_a={stepping_a}, cyclic={cyclic}]",
          flush=True)
    orc = oracle_lib.Oracle(case)
    gpu = OceanGPU(case.dom, case.grid)
    f = {k: DeviceArray.from_host(v) for k, v in case.fields.items() if isinstance(v, np.ndarray)}
    gpu.set_tuning("barotropic_subcycles_per_launch", 5)
    eta_ref, ud_ref, conv_ref = orc.update_ocean_barotropic(case.nts, case.dtbt,
                                                            stepping_a=stepping_a)
    shp = orc.shape2
    got = {}
    for land in (1, 0):
        gpu.set_tuning("barotropic_land_tiles", land)
        eb, ud, cv = DeviceArray(shp), DeviceArray((2,) + shp), DeviceArray(shp)
        gpu.update_ocean_barotropic(nts=case.nts, dtbt=case.dtbt, eta_t_tau=f["eta_t"],
                                    eta_t_bar_tau=f["eta_t_bar"], udrho_tau=f["udrho"],
                                    forcing_bt=f["forcing_bt"], thicku=f["thicku"],
                                    patm=f["patm"], pme=f["pme"], river=f["river"],
                                    source=f["source"], eta_smooth=f["eta_smooth"],
                                    rho_g=f["rho_g"], eta_t_bar_taup1=eb, udrho_taup1=ud,
                                    conv_rho_ud_t=cv, stepping_a=stepping_a)
        gpu.sync()
        got[land] = [x.to_host() for x in (eb, ud, cv)]
        same(got[land][0], eta_ref)
        same(got[land][1], ud_ref)
        same(got[land][2], conv_ref)
    for a, b in zip(got[1], got[0]):     # the short path's bits, zero signs too
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    gpu.close()


def test_tracer_content(setup):
    from mom5_amd import DeviceArray
    case, orc, gpu, dev = setup
    for _, T in case.tracers:
        ref = orc.tracer_content(T)
        got = gpu.tracer_content(DeviceArray.from_host(T), dev["rho_dzt"])
        assert abs(got - ref) <= 1e-15 * abs(ref), (got, ref)
