"""Pins the small routines: compute_tmask_limit (ocean_tracer.F90:
2788-2892) on hand-made cases with known flags, the explicit update
(:2341-2350) on closed forms, the masked content sum, and the single-PE
halo semantics of mpp_update_domains."""
import ctypes as C

import numpy as np

from _cases import uniform_case


def test_tmask_limit_single_outlier(oracle_lib):
    c = uniform_case(8, 6, 4)
    o = oracle_lib.Oracle(c)
    T = np.full((4, 8, 10), 10.0) * c.grid["tmask"]
    S = np.full((4, 8, 10), 35.0) * c.grid["tmask"]
    T[2, 3, 4] = 50.0                       # one hot cell at (i=4, j=3, k=3)
    L = o.tmask_limit([T, S], [0.0, 30.0], [40.0, 40.0], 0, 1, ts_same=False)
    lt = L[0]
    expect = {(2, 3, 4), (2, 3, 3), (2, 2, 4), (1, 3, 4)}   # self, west, south, above
    got = {tuple(x) for x in np.argwhere(lt == 1.0)}
    assert got == expect
    assert not L[1].any()
    L2 = o.tmask_limit([T, S], [0.0, 30.0], [40.0, 40.0], 0, 1, ts_same=True)
    assert np.array_equal(L2[0], L2[1]) and L2[1].sum() == 4


def test_tmask_limit_ignores_land_values(oracle_lib):
    kmt = np.zeros((8, 10), dtype=np.int32)
    kmt[1:-1, 1:-1] = 4
    kmt[3, 5] = 0
    kmt[:, 0] = kmt[:, -2]
    kmt[:, -1] = kmt[:, 1]
    c = uniform_case(8, 6, 4, kmt=kmt)
    o = oracle_lib.Oracle(c)
    T = np.full((4, 8, 10), 10.0)
    T[:, 3, 5] = -99.0                      # land column holding junk
    L = o.tmask_limit([T], [0.0], [40.0], -1, -1, ts_same=False)
    assert not L[0].any()


def test_explicit_update_closed_form(oracle_lib):
    c = uniform_case(6, 4, 3, rho=2.0)
    o = oracle_lib.Oracle(c)
    T = np.full((3, 6, 8), 3.0)
    th = np.full((3, 6, 8), 4.0)
    out = o.update_explicit(T, th, 0.5)
    # (rho*T + dt*th) / rho = (6 + 2) / 2 = 4
    assert np.all(out[:, 1:-1, 1:-1] == 4.0)
    assert np.all(out[:, 0, :] == 0.0)


def test_content_of_unit_tracer_is_wet_mass(oracle_lib):
    from mom5_amd.synth import make_case
    case = make_case(30, 20, 6, nts=10, seed=51)
    o = oracle_lib.Oracle(case)
    one = np.ones((6, 22, 32))
    tm = case.grid["tmask"][:, 1:-1, 1:-1]
    mass = case.fields["rho_dzt"][:, 1:-1, 1:-1] * case.grid["dat"][None, 1:-1, 1:-1] * tm
    import math
    assert o.tracer_content(one) == math.fsum(mass.ravel())


def test_halo_update_single_pe(oracle_lib):
    L = oracle_lib.lib()
    dom = oracle_lib.Dom(isd=0, ied=5, jsd=0, jed=4, isc=1, iec=4, jsc=1, jec=3, nk=1,
                         cyclic_x=1, cyclic_y=0)
    h = 2
    a = np.arange((4 + 2 * h) * (3 + 2 * h), dtype=np.float64).reshape(3 + 2 * h, 4 + 2 * h)
    b = a.copy()
    L.orc_halo_update(C.byref(dom), h, b.ctypes.data_as(C.POINTER(C.c_double)), 1, 3)
    # compute rows: halo columns wrap; halo rows and corners untouched
    inner = a[h:-h]
    assert np.array_equal(b[h:-h, :h], inner[:, -2 * h:-h])
    assert np.array_equal(b[h:-h, -h:], inner[:, h:2 * h])
    assert np.array_equal(b[:h], a[:h]) and np.array_equal(b[-h:], a[-h:])
