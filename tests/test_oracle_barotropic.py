"""Pins the barotropic restatement (ocean_barotropic.F90:3225-3579 with the
update_ocean_barotropic tail :2804-2894; DIV_UD and GRAD_BAROTROPIC_P of
ocean_operators.F90:507-676):

* a state at rest with no forcing stays at rest;
* the flux-form divergence conserves mass: sum(dat*tmask*eta_bar) is
  unchanged without surface fluxes;
* pure inertial motion on an f-plane (uniform eta, no forcing, away from the
  walls) rotates U by exactly 2*atan(f*dtbt/2) per subcycle with |U| kept:
  the closed form of the Crank-Nicolson Coriolis step, averaged with the
  stepping-B weights nts-s+1 (or uniform weights for stepping A);
* constant fields average to themselves (sum of weights = 1/rnts);
* one subcycle with gamma = 0 and f = 0 gives U = -dtbt*H*grad(p_s) exactly
  for a quadratic surface height (pins the B-grid gradient's sign, index
  offsets, metric and mask).
"""
import math

import numpy as np
import pytest

from _cases import uniform_case


def _bt_fields(c, eta=0.0, u1=0.0, u2=0.0, hu=4000.0):
    NJ, NI = c.nj + 2, c.ni + 2
    z2 = np.zeros((NJ, NI))
    tm = c.grid["tmask"][0]
    um = c.grid["umask"][0]
    f = dict(c.fields)
    f.update(eta_t=eta * tm, eta_t_bar=eta * tm, udrho=np.stack([u1 * um, u2 * um]),
             forcing_bt=np.zeros((2, NJ, NI)), thicku=hu * um, patm=z2.copy(), pme=z2.copy(),
             river=z2.copy(), source=z2.copy(), eta_smooth=z2.copy(),
             rho_g=np.full((NJ, NI), 9.8 * 1035.0))
    return f


def test_rest_state_stays_at_rest(oracle_lib):
    c = uniform_case(24, 16, 3, dx=1e4, dy=1e4, f=1e-4)
    o = oracle_lib.Oracle(c)
    eb, ud, conv = o.update_ocean_barotropic(20, 10.0, fields=_bt_fields(c))
    assert np.all(eb == 0) and np.all(ud == 0) and np.all(conv == 0)


def test_mass_conservation(oracle_lib):
    from mom5_amd.synth import make_case
    case = make_case(48, 32, 6, nts=24, seed=41)
    o = oracle_lib.Oracle(case)
    eb, ud, _ = o.update_ocean_barotropic(case.nts, case.dtbt)
    w = (case.grid["dat"] * case.grid["tmask"][0])[1:-1, 1:-1]
    m0 = (w * case.fields["eta_t_bar"][1:-1, 1:-1]).sum()
    m1 = (w * eb[1:-1, 1:-1]).sum()
    assert abs(m1 - m0) <= 1e-12 * (w * np.abs(case.fields["eta_t_bar"][1:-1, 1:-1])).sum()
    assert np.abs(eb - case.fields["eta_t_bar"]).max() > 1e-6   # something did happen


@pytest.mark.parametrize("stepping_a", [False, True])
def test_inertial_oscillation_closed_form(oracle_lib, stepping_a):
    fcor, dtbt, nts = 1.2e-4, 300.0, 6
    c = uniform_case(24, 30, 3, dx=1e4, dy=1e4, f=fcor)
    o = oracle_lib.Oracle(c)
    fl = _bt_fields(c, eta=0.25, u1=3.0, u2=-1.0)
    eb, ud, _ = o.update_ocean_barotropic(nts, dtbt, stepping_a=stepping_a, fields=fl)
    phi = 2.0 * math.atan(fcor * dtbt / 2.0)
    U = np.array([3.0, -1.0])
    acc = np.zeros(2)
    for s in range(1, nts + 1):
        R = np.array([[math.cos(s * phi), math.sin(s * phi)],
                      [-math.sin(s * phi), math.cos(s * phi)]])
        w = 1.0 if stepping_a else float(nts - s + 1)
        acc += w * (R @ U)
    rnts = 1.0 / nts if stepping_a else 2.0 / (nts * (nts + 1))
    ref = rnts * acc
    mid = (slice(12, 19), slice(4, 20))      # rows far from the closed walls
    assert np.allclose(ud[0][mid], ref[0], rtol=1e-13)
    assert np.allclose(ud[1][mid], ref[1], rtol=1e-13)
    assert np.allclose(eb[mid], 0.25, rtol=1e-14)


def test_constant_fields_average_to_themselves(oracle_lib):
    c = uniform_case(20, 24, 3, dx=1e4, dy=1e4, f=0.0)
    o = oracle_lib.Oracle(c)
    fl = _bt_fields(c, eta=0.5, u1=2.0, u2=0.0)
    eb, ud, _ = o.update_ocean_barotropic(16, 60.0, fields=fl)
    mid = (slice(10, 15), slice(3, 18))
    assert np.allclose(ud[0][mid], 2.0, rtol=1e-14)
    assert np.allclose(ud[1][mid], 0.0, atol=1e-14)
    assert np.allclose(eb[mid], 0.5, rtol=1e-15)


@pytest.mark.parametrize("stepping_a", [False, True])
def test_bgrid_pressure_gradient_known_answer(oracle_lib, stepping_a):
    """One subcycle with gamma = 0 and f = 0 on a closed box: the predictor
    leaves eta unchanged, so U = -dtbt*thicku*GRAD_BAROTROPIC_P(rho_g*eta)
    (ocean_barotropic.F90:3409,3464-3486; ocean_operators.F90:650-660).
    For eta quadratic in x and y the B-grid difference of the two-point
    averages is the exact derivative at the U point (i+1/2, j+1/2), so
    U = -dtbt*H*rho_g*(2a*x_u, 2b*y_u) at wet U points and 0 elsewhere.
    A flipped sign, a backward difference (2i-1 in place of 2i+1), dyur in
    place of dxur (dx != dy) or a missing umask fails."""
    dx, dy, dtbt, H, rho_g = 2.0e4, 1.5e4, 20.0, 3000.0, 1.0e4
    c = uniform_case(18, 14, 2, cyclic_x=False, dx=dx, dy=dy, f=0.0)
    o = oracle_lib.Oracle(c)
    NJ, NI = c.nj + 2, c.ni + 2
    jj, ii = np.meshgrid(np.arange(NJ, dtype=np.float64), np.arange(NI, dtype=np.float64),
                         indexing="ij")
    a, b = 3.0e-11, -2.0e-11
    eta = a * (ii * dx) ** 2 + b * (jj * dy) ** 2
    fl = _bt_fields(c, hu=H)
    fl.update(eta_t=eta, eta_t_bar=eta, rho_g=np.full((NJ, NI), rho_g))
    _, ud, _ = o.update_ocean_barotropic(1, dtbt, gamma=0.0, stepping_a=stepping_a, fields=fl)
    um = c.grid["umask"][0]
    want_x = -dtbt * H * rho_g * 2.0 * a * (ii + 0.5) * dx * um
    want_y = -dtbt * H * rho_g * 2.0 * b * (jj + 0.5) * dy * um
    scale = np.abs(want_x).max()
    assert um.sum() > 100
    assert np.abs(ud[0] - want_x).max() <= 1e-12 * scale
    assert np.abs(ud[1] - want_y).max() <= 1e-12 * np.abs(want_y).max()
    assert np.all(ud[:, um == 0] == 0)
