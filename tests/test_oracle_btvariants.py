"""The CPU restatement of the four split-explicit barotropic schemes
(B/C-grid x depth/pressure-based; ocean_barotropic.F90:3225-4693) pinned by:

* the B-grid depth-based variant equals the separately pinned
  orc_update_ocean_barotropic bit for bit (same scheme, other code);
* C-grid volume budget on a closed basin: sum(dat*eta_t_bar) over the
  compute domain = S0 + dtbt/rho0 * sum(dat*steady_forcing) * nts/2 (the
  time mean of the subcycles; a wrong DIV_UD index breaks the telescoping);
* C-grid inertial oscillation on an f-plane: uniform (U, V) far from the
  walls follows the textbook third-order Adams-Bashforth recurrence
  u' = f v, v' = -f u with the start-up levels equal to the first one
  (a wrong level rotation or weight fails);
* C-grid pressure gradient: one subcycle with f = 0, gamma = 0 and a
  quadratic eta gives u = -dtbt * H * rho_g * (eta(i+1) - eta(i))/dxte at
  wet east faces, exactly (pins the forward difference, 1/dxte, the face
  mask and the sign);
* pressure-based = depth-based scaled: anompb = rho0*grav*eta, mass =
  rho0*thickness, rho_g = rho0*grav, patm = dpatm_dt = 0 give
  anompb_bar = rho0*grav*eta_t_bar and the same udrho (to rounding), on
  both grids (a missing grav or 1/rho0 fails by orders of magnitude).
"""
import numpy as np

from _cases import bt_tmasken as _tmasken, bt_variant_fields as _fields, uniform_case


def _regional(cyclic=False):
    from mom5_amd.synth import make_case
    kw = {} if cyclic else dict(cyclic_x=False, lat_s=20.0, lat_n=50.0, lon_span=30.0)
    return make_case(40, 30, 3, seed=51, land_frac=0.25, barotropic=True, **kw)


def test_bgrid_depth_equals_pinned_scheme(oracle_lib):
    for cyc in (True, False):
        c = _regional(cyc)
        o = oracle_lib.Oracle(c)
        f = c.fields
        for stepping_a in (True, False):
            ref = o.update_ocean_barotropic(12, c.dtbt, stepping_a=stepping_a)
            fl = dict(x_tau=f["eta_t"], xbar_tau=f["eta_t_bar"], udrho=f["udrho"],
                      forcing_bt=f["forcing_bt"], thick=f["thicku"], patm=f["patm"],
                      rho_g=f["rho_g"], pme=f["pme"], river=f["river"], source=f["source"],
                      smooth=f["eta_smooth"])
            got = o.update_ocean_barotropic_variant(12, c.dtbt, cgrid=False, press=False,
                                                    fields=fl, stepping_a=stepping_a)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)


def test_cgrid_volume_budget(oracle_lib):
    c = _regional(False)
    o = oracle_lib.Oracle(c)
    rng = np.random.default_rng(3)
    fl = _fields(c, rng, True, scale=0.5)
    nts = 10
    xbar, ud, conv = o.update_ocean_barotropic_variant(nts, c.dtbt, cgrid=True, press=False,
                                                       fields=fl, stepping_a=True)
    dat = c.grid["dat"]
    tm = c.grid["tmask"][0]
    sl = (slice(1, -1), slice(1, -1))
    S0 = np.sum((dat * fl["xbar_tau"])[sl])
    F = np.sum((dat * tm * fl["pme"])[sl])
    want = S0 + c.dtbt / 1035.0 * F * nts / 2.0
    got = np.sum((dat * xbar)[sl])
    scale = np.sum(np.abs(dat * fl["xbar_tau"])[sl])
    assert abs(got - want) <= 1e-12 * scale, (got, want)
    assert abs(got - S0) > 1e-9 * scale      # the budget is not trivially S0


def test_cgrid_inertial_oscillation_ab3(oracle_lib):
    ni, nj, nts = 12, 40, 4
    c = uniform_case(ni, nj, 2, cyclic_x=True, dx=1.0e4, dy=1.0e4, f=1.0e-4)
    o = oracle_lib.Oracle(c)
    sh = (nj + 2, ni + 2)
    m = _tmasken(c)
    U, V = 0.7, -0.3
    z = np.zeros(sh)
    fl = dict(x_tau=z, xbar_tau=z, udrho=np.stack([U * m[0], V * m[1]]),
              forcing_bt=np.zeros((2,) + sh), thick=100.0 * m, tmasken=m, patm=z,
              rho_g=np.full(sh, 1.0e4), pme=z, river=z, source=z, smooth=z)
    dt, f0 = 60.0, 1.0e-4
    for stepping_a in (True, False):
        _, ud, _ = o.update_ocean_barotropic_variant(nts, dt, cgrid=True, press=False,
                                                     fields=fl, stepping_a=stepping_a)
        ab = (23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0)
        u, v = [U], [V]
        for n in range(nts):
            fu = [f0 * v[max(n - q, 0)] for q in range(3)]
            fv = [-f0 * u[max(n - q, 0)] for q in range(3)]
            u.append(u[n] + dt * sum(a * x for a, x in zip(ab, fu)))
            v.append(v[n] + dt * sum(a * x for a, x in zip(ab, fv)))
        w = [1.0] * nts if stepping_a else [float(nts - s) for s in range(nts)]
        r = 1.0 / nts if stepping_a else 2.0 / (nts * (nts + 1))
        mu = r * sum(wi * ui for wi, ui in zip(w, u[1:]))
        mv = r * sum(wi * vi for wi, vi in zip(w, v[1:]))
        j = nj // 2
        np.testing.assert_allclose(ud[0, j, 1:-1], mu, rtol=1e-13)
        np.testing.assert_allclose(ud[1, j, 1:-1], mv, rtol=1e-13)
        assert abs(mu - U) > 1e-4 * abs(U)       # the rotation is visible


def test_cgrid_pressure_gradient_known_answer(oracle_lib):
    ni, nj = 10, 6
    c = uniform_case(ni, nj, 2, cyclic_x=False, dx=2.0, dy=1.0)
    o = oracle_lib.Oracle(c)
    sh = (nj + 2, ni + 2)
    m = _tmasken(c)
    s, H, R, dt = 0.125, 8.0, 4.0, 0.5
    ii = np.arange(ni + 2, dtype=np.float64)
    eta = s * (ii * ii)[None, :] * np.ones(sh)
    z = np.zeros(sh)
    fl = dict(x_tau=eta, xbar_tau=eta, udrho=np.zeros((2,) + sh),
              forcing_bt=np.zeros((2,) + sh), thick=H * np.ones((2,) + sh), tmasken=m, patm=z,
              rho_g=np.full(sh, R), pme=z, river=z, source=z, smooth=z)
    _, ud, _ = o.update_ocean_barotropic_variant(1, dt, cgrid=True, press=False, fields=fl,
                                                 stepping_a=True, gamma=0.0)
    # d(eta)/dx on the east face of cell i: s*((i+1)^2 - i^2)/dx = s*(2i+1)/2
    want = -dt * H * R * s * (2.0 * ii + 1.0) / 2.0 * np.ones(sh)
    assert np.array_equal(ud[0][m[0] == 1.0], want[m[0] == 1.0])
    assert np.all(ud[0][m[0] == 0.0] == 0.0)
    assert np.all(ud[1] == 0.0)


def test_pressure_based_is_scaled_depth_based(oracle_lib):
    rg = 1035.0 * 9.80
    for cgrid in (False, True):
        c = _regional(True)
        o = oracle_lib.Oracle(c)
        rng = np.random.default_rng(4)
        fl = _fields(c, rng, cgrid)
        d = o.update_ocean_barotropic_variant(10, c.dtbt, cgrid=cgrid, press=False, fields=fl)
        fp = dict(fl)
        fp["x_tau"] = rg * fl["x_tau"]
        fp["xbar_tau"] = rg * fl["xbar_tau"]
        fp["thick"] = 1035.0 * fl["thick"]
        fp["pme"] = fl["pme"]            # steady: grav*(pme ...) vs (pme ...)/rho0 * rho0*grav
        p = o.update_ocean_barotropic_variant(10, c.dtbt, cgrid=cgrid, press=True, fields=fp)
        sl = (slice(1, -1), slice(1, -1))
        np.testing.assert_allclose(p[0][sl], rg * d[0][sl], rtol=1e-9,
                                   atol=1e-9 * np.abs(rg * d[0]).max())
        for n in range(2):
            np.testing.assert_allclose(p[1][n][sl], d[1][n][sl], rtol=1e-9,
                                       atol=1e-9 * np.abs(d[1]).max())
        np.testing.assert_allclose(p[2][sl], d[2][sl], rtol=1e-9, atol=1e-9 * np.abs(d[2]).max())
