"""The CPU restatement of the surface-height step and of the barotropic
closure pinned by closed forms and invariants.

eta_and_pbot_tendency + eta_and_pbot_update (ocean_barotropic.F90:2667-2700,
2290-2363):

* unit response: with rho0 = 1024 (rho0r exact) and one of conv, pme,
  river, source set to a dyadic value q, deta_dt = q/1024 exactly in wet
  columns and 0 on land (a dropped term, a wrong tmask level or rho0 in
  place of rho0r fails);
* the step is x_taum1 + dtime*deta_dt on the compute domain, the halo of a
  cyclic-x domain wraps on the compute rows, closed halo rows and corners
  keep the array's previous contents;
* pressure-based: dpbot_dt - dpatm_dt = grav*rho0*deta_dt (same inputs, to
  rounding) and pbot_t - anompb = rho0*grav*ht.

Closure (ocean_barotropic.F90:2906-2921, 2956-2985; operators
ocean_operators.F90:462-477, 631-676):

* REMAP_BT_TO_BU of a constant is that constant wherever daur is the
  reciprocal of the four quarter-cell areas (random widths: a mispaired
  width or a dropped quarter fails), masked by umask(k=1), zero in the last
  row and column before the halo update;
* a unit impulse at one T point reaches exactly the four U cells around it
  with its quarter-cell weight;
* GRAD_BAROTROPIC_P of p = a*X(i) + b*Y(j) (X, Y cumulative U-cell
  widths) is (a, b) at wet U points, zero on land and in the last row and
  column; on the C-grid (X, Y cumulative dxte, dytn) it is (a, b) at wet
  east and north faces;
* ps = rho_g*eta_t_bar + patm; pressure-based pbot_u of a constant depth
  with zero anomaly is rho0*grav*H.
"""
import numpy as np

from _cases import uniform_case

NI, NJ = 12, 9      # data domain of a 10 x 7 case


def _case(cyclic_x=True):
    c = uniform_case(10, 7, 4, cyclic_x=cyclic_x)
    kmt = c.grid["kmt"]
    kmt[3, 4] = 0
    kmt[5, 7] = 0
    if cyclic_x:
        kmt[:, 0] = kmt[:, -2]
        kmt[:, -1] = kmt[:, 1]
    kmu = np.zeros_like(kmt)
    kmu[:-1, :-1] = np.minimum.reduce([kmt[:-1, :-1], kmt[:-1, 1:], kmt[1:, :-1], kmt[1:, 1:]])
    if cyclic_x:
        kmu[:, -1] = kmu[:, 1]
    c.grid["kmu"] = kmu
    kk = np.arange(1, 5)[:, None, None]
    c.grid["tmask"] = (kk <= kmt[None]).astype(np.float64)
    c.grid["umask"] = (kk <= kmu[None]).astype(np.float64)
    return c


def test_unit_response_of_each_term(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    z = np.zeros((NJ, NI))
    tm = c.grid["tmask"][0]
    for which in range(4):
        ins = [z.copy() for _ in range(4)]
        ins[which][:] = 3.0 * 2.0 ** -3
        x, tend, _ = o.eta_and_pbot_update(z, *ins, dtime=2.0 ** 5, rho0=1024.0)
        want = tm * (3.0 * 2.0 ** -3 / 1024.0)
        assert np.array_equal(tend, want), which
        assert np.array_equal(x[1:-1, 1:-1], (2.0 ** 5 * want)[1:-1, 1:-1]), which


def test_step_and_halo(oracle_lib):
    for cyc in (True, False):
        c = _case(cyc)
        o = oracle_lib.Oracle(c)
        rng = np.random.default_rng(5)
        x0 = rng.standard_normal((NJ, NI))
        conv, pme, river, src = (rng.standard_normal((NJ, NI)) for _ in range(4))
        prev = np.full((NJ, NI), -7.25)
        x, tend, _ = o.eta_and_pbot_update(x0, conv, pme, river, src, 1800.0, x_taup1=prev)
        assert np.array_equal(x[1:-1, 1:-1], (x0 + 1800.0 * tend)[1:-1, 1:-1])
        # rows j = 0 and j = nj+1 (closed in y) keep their previous contents
        assert np.all(x[0, :] == -7.25) and np.all(x[-1, :] == -7.25)
        if cyc:
            assert np.array_equal(x[1:-1, 0], x[1:-1, -2])
            assert np.array_equal(x[1:-1, -1], x[1:-1, 1])
        else:
            assert np.all(x[:, 0] == -7.25) and np.all(x[:, -1] == -7.25)
        # the tendency covers the data domain (halo included)
        tm = c.grid["tmask"][0]
        assert np.all(tend[tm == 0.0] == 0.0)
        assert np.all(tend[tm == 1.0] != 0.0)


def test_pressure_based_relations(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    rng = np.random.default_rng(6)
    x0 = rng.standard_normal((NJ, NI))
    ins = [rng.standard_normal((NJ, NI)) for _ in range(4)]
    dpatm = rng.standard_normal((NJ, NI))
    ht = rng.uniform(100.0, 5000.0, (NJ, NI))
    _, te, _ = o.eta_and_pbot_update(x0, *ins, 1800.0)
    xa, tp, pbot = o.eta_and_pbot_update(x0, *ins, 1800.0, pressure_based=True, dpatm_dt=dpatm,
                                         ht=ht)
    np.testing.assert_allclose(tp - dpatm, 9.80 * 1035.0 * te, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pbot - xa, 1035.0 * 9.80 * ht, rtol=1e-15)
    assert np.array_equal(xa[1:-1, 1:-1], (x0 + 1800.0 * tp)[1:-1, 1:-1])


def _ucell(rng):
    dte, dtw, dus, dun = (rng.uniform(0.5, 2.0, (NJ, NI)) for _ in range(4))
    area = np.zeros((NJ, NI))
    area[:-1, :-1] = (dte[:-1, :-1] * dus[:-1, :-1] + dtw[:-1, 1:] * dus[:-1, :-1]
                      + dte[1:, :-1] * dun[:-1, :-1] + dtw[1:, 1:] * dun[:-1, :-1])
    area[-1, :] = 1.0
    area[:, -1] = 1.0
    return dict(dte=dte, dtw=dtw, dus=dus, dun=dun, daur=1.0 / area)


def test_remap_constant_and_impulse(oracle_lib):
    for cyc in (True, False):
        c = _case(cyc)
        o = oracle_lib.Oracle(c)
        rng = np.random.default_rng(7)
        uc = _ucell(rng)
        z = np.zeros((NJ, NI))
        um = c.grid["umask"][0]
        eu, ps, gps, gab = o.barotropic_closure(np.full((NJ, NI), 0.625), z, z, uc)
        assert gab is None
        inner = eu[:-1, :-1]
        np.testing.assert_allclose(inner[um[:-1, :-1] == 1.0], 0.625, rtol=4e-16)
        assert np.all(eu[um == 0.0] == 0.0)
        assert np.all(eu[-1, :] == 0.0)
        if cyc:
            assert np.array_equal(eu[1:-1, -1], eu[1:-1, 1])
            assert np.array_equal(eu[1:-1, 0], eu[1:-1, -2])
        else:
            assert np.all(eu[:, -1] == 0.0)
        # impulse at T(i0, j0)
        i0, j0 = 6, 4
        a = z.copy()
        a[j0, i0] = 1.0
        eu, _, _, _ = o.barotropic_closure(a, z, z, uc)
        want = np.zeros((NJ, NI))
        want[j0, i0] = uc["dte"][j0, i0] * uc["dus"][j0, i0] * uc["daur"][j0, i0]
        want[j0, i0 - 1] = uc["dtw"][j0, i0] * uc["dus"][j0, i0 - 1] * uc["daur"][j0, i0 - 1]
        want[j0 - 1, i0] = uc["dte"][j0, i0] * uc["dun"][j0 - 1, i0] * uc["daur"][j0 - 1, i0]
        want[j0 - 1, i0 - 1] = (uc["dtw"][j0, i0] * uc["dun"][j0 - 1, i0 - 1]
                                * uc["daur"][j0 - 1, i0 - 1])
        want *= um
        np.testing.assert_allclose(eu, want, rtol=1e-15, atol=0.0)


def test_gradient_of_linear_pressure(oracle_lib):
    c = _case(False)
    rng = np.random.default_rng(8)
    dxu = rng.uniform(0.5, 2.0, NI)
    dyu = rng.uniform(0.5, 2.0, NJ)
    c.grid["dxu"] = np.broadcast_to(dxu, (NJ, NI)).copy()
    c.grid["dyu"] = np.broadcast_to(dyu[:, None], (NJ, NI)).copy()
    c.grid["dxur"] = 1.0 / c.grid["dxu"]
    c.grid["dyur"] = 1.0 / c.grid["dyu"]
    X = np.concatenate([[0.0], np.cumsum(dxu[:-1])])
    Y = np.concatenate([[0.0], np.cumsum(dyu[:-1])])
    o = oracle_lib.Oracle(c)
    uc = _ucell(rng)
    a, b = 0.75, -1.25
    patm = a * X[None, :] + b * Y[:, None]
    z = np.zeros((NJ, NI))
    eu, ps, gps, gab = o.barotropic_closure(z, patm, np.full((NJ, NI), 3.0), uc,
                                            anompb_bar=patm)
    assert np.array_equal(ps, patm)
    um = c.grid["umask"][0]
    wet = um[:-1, :-1] == 1.0
    for n, v in ((0, a), (1, b)):
        np.testing.assert_allclose(gps[n][:-1, :-1][wet], v, rtol=1e-12)
        assert np.all(gps[n][um == 0.0] == 0.0)
        assert np.all(gps[n][-1, :] == 0.0) and np.all(gps[n][:, -1] == 0.0)
    assert np.array_equal(gab, gps)
    # ps = rho_g*eta_t_bar + patm
    eta = rng.standard_normal((NJ, NI))
    rg = rng.uniform(9000.0, 11000.0, (NJ, NI))
    _, ps, _, _ = o.barotropic_closure(eta, patm, rg, uc)
    assert np.array_equal(ps, rg * eta + patm)


def test_pressure_based_pbot_u(oracle_lib):
    c = _case()
    o = oracle_lib.Oracle(c)
    rng = np.random.default_rng(9)
    uc = _ucell(rng)
    z = np.zeros((NJ, NI))
    H = 3000.0
    pu, _, _, gab = o.barotropic_closure(z, z, z, uc, anompb_bar=z, pressure_based=True,
                                         ht=np.full((NJ, NI), H))
    um = c.grid["umask"][0]
    np.testing.assert_allclose(pu[:-1, :-1][um[:-1, :-1] == 1.0], 1035.0 * 9.80 * H, rtol=4e-16)
    assert np.all(pu[um == 0.0] == 0.0)
    assert np.all(gab == 0.0)


def test_cgrid_gradient_of_linear_pressure(oracle_lib):
    from _cases import bt_tmasken
    c = _case(False)
    rng = np.random.default_rng(10)
    dxte = rng.uniform(0.5, 2.0, NI)
    dytn = rng.uniform(0.5, 2.0, NJ)
    c.grid["dxte"] = np.broadcast_to(dxte, (NJ, NI)).copy()
    c.grid["dytn"] = np.broadcast_to(dytn[:, None], (NJ, NI)).copy()
    X = np.concatenate([[0.0], np.cumsum(dxte[:-1])])
    Y = np.concatenate([[0.0], np.cumsum(dytn[:-1])])
    o = oracle_lib.Oracle(c)
    uc = _ucell(rng)
    m = bt_tmasken(c)
    a, b = -0.5, 2.25
    patm = a * X[None, :] + b * Y[:, None]
    z = np.zeros((NJ, NI))
    _, ps, gps, gab = o.barotropic_closure(z, patm, z, uc, anompb_bar=patm, tmasken=m)
    for n, v in ((0, a), (1, b)):
        np.testing.assert_allclose(gps[n][m[n] == 1.0], v, rtol=1e-12)
        assert np.all(gps[n][m[n] == 0.0] == 0.0)
        assert np.all(gps[n][-1, :] == 0.0) and np.all(gps[n][:, -1] == 0.0)
    assert np.array_equal(gab, gps)
    # the B-grid gradient of the same field differs (it is not the C-grid one)
    _, _, gb, _ = o.barotropic_closure(z, patm, z, uc)
    assert not np.array_equal(gb, gps)
